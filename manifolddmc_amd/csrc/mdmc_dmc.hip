// mdmc_dmc.hip -- grid preparation (padding / refinement) and triangle output.
//
// Restates (MarchingCubes/VisualDMC/src/DMC):
//   refineGrid                      UniformGrid.inl:279-319
// The cell classification, dual vertices and quads are in mdmc_gridpass.hip.
#include "mdmc_device.h"
#include "mdmc_internal.h"

namespace mdmc {

namespace {

__device__ __forceinline__ int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

// one thread per point of every processed grid; grid g owns blocks [d[g].base, d[g + 1].base),
// so the grid a block works on, and with it every size below, is uniform across the block
__global__ void __launch_bounds__(kBlock) kPrepareGrids(const PrepareDesc* __restrict__ d, int G, int refinement, int pad) {
    const int64_t blk = blockIdx.x;
    int lo = 0, hi = G;  // d[lo].base <= blk
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (d[mid].base <= blk) lo = mid;
        else hi = mid;
    }
    const PrepareDesc& p = d[lo];
    const int onx = p.onx, ony = p.ony, onz = p.onz, inx = p.inx, iny = p.iny, inz = p.inz;
    const int64_t i = (blk - p.base) * kBlock + threadIdx.x;
    if (i >= (int64_t)onx * ony * onz) return;
    int x = (int)(i % onx);
    int y = (int)((i / onx) % ony);
    int z = (int)(i / ((int64_t)onx * ony));
    if (pad) {
        x = clampi(x, 1, onx - 2) - 1;
        y = clampi(y, 1, ony - 2) - 1;
        z = clampi(z, 1, onz - 2) - 1;
    }
    const float rr = (float)(refinement + 1);
    const float fx = (float)x / rr, fy = (float)y / rr, fz = (float)z / rr;
    const int mx = (int)floorf(fx), my = (int)floorf(fy), mz = (int)floorf(fz);
    float s[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int cx = clampi(mx + (k & 1), 0, inx - 1), cy = clampi(my + ((k >> 1) & 1), 0, iny - 1), cz = clampi(mz + ((k >> 2) & 1), 0, inz - 1);
        s[k] = p.in[(int64_t)cz * inx * iny + (int64_t)cy * inx + cx];
    }
    const f3 dd = mk3(fx - (float)mx, fy - (float)my, fz - (float)mz);
    p.out[i] = trilinear(s, dd);
}

__global__ void kTriangles(const int4* __restrict__ he, const int32_t* __restrict__ face_he, int64_t F, uint32_t* __restrict__ idx) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    int h = face_he[f];
    int v[4];
    for (int i = 0; i < 4; ++i) {
        const int4 e = he[h];
        v[i] = e.x;
        h = e.z;
    }
    uint32_t* o = idx + 6 * f;
    o[0] = v[0]; o[1] = v[1]; o[2] = v[2];
    o[3] = v[2]; o[4] = v[3]; o[5] = v[0];
}

}  // namespace

int64_t prepareBlocks(int64_t points) { return (points + kBlock - 1) / kBlock; }

void launchPrepareGrids(const PrepareDesc* descDev, int G, int64_t blocks, int refinement, int pad, hipStream_t s) {
    if (blocks > 0) kPrepareGrids<<<(unsigned)blocks, kBlock, 0, s>>>(descDev, G, refinement, pad);
}

void launchDmcTrianglesFromFaces(const int4* he, const int32_t* face_he, int64_t F, uint32_t* idx, hipStream_t s) {
    if (F > 0) kTriangles<<<gridFor(F), kBlock, 0, s>>>(he, face_he, F, idx);
}

}  // namespace mdmc
