// mdmc_dmc.hip -- grid preparation (padding / refinement) and triangle output.
//
// Restates (MarchingCubes/VisualDMC/src/DMC):
//   refineGrid                      UniformGrid.inl:279-319
// The cell classification, dual vertices and quads are in mdmc_gridpass.hip.
#include "mdmc_device.h"
#include "mdmc_internal.h"

#include <algorithm>

namespace mdmc {

namespace {

__device__ __forceinline__ int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

// One block per row (y, z) of every processed grid, its threads striding over x; grid g owns
// rows [d[g].base, d[g + 1].base), so the grid a block works on, and with it every size
// below, is uniform across the block.
//
// MODE 2 (refinement > 0): the trilinear sample at x / (r + 1).
// MODE 1 (no refinement, exact): the sample position is the point itself: x / 1 = x exactly,
//   floor(x) = x and the weights are +0, the same values the general path computes, so the
//   8 corner loads and the trilinear blend stay (a non-finite neighbour still reaches the
//   result through 0 * s).  Runs only when *special is set, as a grid-stride loop.
// MODE 0 (no refinement, plain): when every input value v has |v| < 2^126 and is not -0, each
//   difference of two of them is finite, so s0 + 0 * (s1 - s0) = s0 and every level of the
//   zero-weight blend returns its first operand: the result is the first corner, bit for bit.
//   This mode copies that corner and sets *special when it meets any other value (every input
//   point is some output point's first corner, so none is missed); MODE 1 then redoes all.
__device__ __forceinline__ bool plainValue(float v) {
    const uint32_t b = __float_as_uint(v);
    return (b & 0x7FFFFFFFu) < 0x7E800000u && b != 0x80000000u;
}

// one output row; its points x = lane, lane + width, ... (a block, or one wave in MODE 0)
template <int MODE>
__device__ __forceinline__ void prepareRow(const PrepareDesc* __restrict__ d, int G, int refinement, int pad, int64_t row, int64_t* special,
                                           int lane, int width) {
    int lo = 0, hi = G;  // d[lo].base <= row
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (d[mid].base <= row) lo = mid;
        else hi = mid;
    }
    const PrepareDesc& p = d[lo];
    const int onx = p.onx, ony = p.ony, onz = p.onz, inx = p.inx, iny = p.iny, inz = p.inz;
    const int r = (int)(row - p.base);  // < ony * onz < 2^31
    int y = r % ony, z = r / ony;
    float* out = p.out + ((int64_t)z * ony + y) * onx;
    if (pad) {
        y = clampi(y, 1, ony - 2) - 1;
        z = clampi(z, 1, onz - 2) - 1;
    }
    const int64_t nxy = (int64_t)inx * iny;
    if (MODE == 0) {
        const float* in = p.in + (int64_t)z * nxy + (int64_t)y * inx;
        bool plain = true;
        for (int x = lane; x < onx; x += width) {
            const float v = in[pad ? clampi(x, 1, onx - 2) - 1 : x];
            plain = plain && plainValue(v);
            out[x] = v;
        }
        if (!plain) *special = 1;
        return;
    }
    const float rr = (float)(refinement + 1);
    int my = y, mz = z;
    float dy = 0.f, dz = 0.f;
    if (MODE == 2) {
        const float fy = (float)y / rr, fz = (float)z / rr;
        my = (int)floorf(fy);
        mz = (int)floorf(fz);
        dy = fy - (float)my;
        dz = fz - (float)mz;
    }
    const int y0 = clampi(my, 0, iny - 1), y1 = clampi(my + 1, 0, iny - 1);
    const int z0 = clampi(mz, 0, inz - 1), z1 = clampi(mz + 1, 0, inz - 1);
    const float* r00 = p.in + (int64_t)z0 * nxy + (int64_t)y0 * inx;
    const float* r10 = p.in + (int64_t)z0 * nxy + (int64_t)y1 * inx;
    const float* r01 = p.in + (int64_t)z1 * nxy + (int64_t)y0 * inx;
    const float* r11 = p.in + (int64_t)z1 * nxy + (int64_t)y1 * inx;
    for (int x = lane; x < onx; x += width) {
        const int px = pad ? clampi(x, 1, onx - 2) - 1 : x;
        int mx = px;
        float dx = 0.f;
        if (MODE == 2) {
            const float fx = (float)px / rr;
            mx = (int)floorf(fx);
            dx = fx - (float)mx;
        }
        const int x0 = clampi(mx, 0, inx - 1), x1 = clampi(mx + 1, 0, inx - 1);
        const float s[8] = {r00[x0], r00[x1], r10[x0], r10[x1], r01[x0], r01[x1], r11[x0], r11[x1]};
        out[x] = trilinear(s, mk3(dx, dy, dz));
    }
}

// MODE 0 copies one row per wave (kBlock / 64 rows per block); the blending modes take a
// block per row
constexpr int kPlainRowsPerBlock = kBlock / 64;
template <int MODE>
__global__ void __launch_bounds__(kBlock) kPrepareGrids(const PrepareDesc* __restrict__ d, int G, int refinement, int pad, int64_t rows,
                                                        int64_t* special) {
    if (MODE == 1) {
        if (*(volatile int64_t*)special == 0) return;
        for (int64_t row = blockIdx.x; row < rows; row += gridDim.x) prepareRow<1>(d, G, refinement, pad, row, special, threadIdx.x, kBlock);
        return;
    }
    if (MODE == 0) {
        const int64_t row = (int64_t)blockIdx.x * kPlainRowsPerBlock + (threadIdx.x >> 6);
        if (row < rows) prepareRow<0>(d, G, refinement, pad, row, special, threadIdx.x & 63, 64);
        return;
    }
    prepareRow<MODE>(d, G, refinement, pad, blockIdx.x, special, threadIdx.x, kBlock);
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

// LazyEngine::Color::HSVtoRGB (LazyEngine/Renderer/Materials/Colors.h:15-56)
__device__ __forceinline__ f3 hsvToRgb(float h, float sat, float val) {
    float f = h / 60.0f;
    const float hi = floorf(f);
    f = f - hi;
    const float p = val * (1 - sat);
    const float q = val * (1 - sat * f);
    const float t = val * (1 - sat * (1 - f));
    if (hi == 0.0f || hi == 6.0f) return mk3(val, t, p);
    if (hi == 1.0f) return mk3(q, val, p);
    if (hi == 2.0f) return mk3(p, val, t);
    if (hi == 3.0f) return mk3(p, q, val);
    if (hi == 4.0f) return mk3(t, p, val);
    return mk3(val, p, q);
}

// HalfedgeMesh insertVertices (HalfedgeMesh.cu:716-760): every halfedge writes the render
// record of its origin vertex -- position, normal, colour by the vertex flags -- into the GL
// vertex layout {vec3 position, vec3 normal, vec4 colour} (Mesh.h:15-19).  All writers of a
// vertex store the same record; a vertex without outgoing halfedges keeps the zeros the
// buffer was filled with.
__global__ void kRenderVertices(const int4* __restrict__ he, int64_t H, const float* __restrict__ pos, const float* __restrict__ nrm,
                                const uint32_t* __restrict__ vflags, float* __restrict__ out) {
    const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= H) return;
    const int v = he[h].x;
    const uint32_t fl = vflags[v];
    const bool nm = fl & VF_NM, weird = fl & VF_WEIRD, fans = fl & VF_FANS;
    float hue = 180.f;  // cyan
    if (nm) hue = 0.f;            // on a non-manifold edge: red
    if (weird) hue = 60.f;        // on an edge with 3 faces: yellow
    if (fans) hue = 270.f;        // non-manifold vertex: purple
    if (weird && nm) hue = 300.f;  // magenta
    if (nm && fans) hue = 240.f;   // dark blue
    const f3 c = hsvToRgb(hue, 0.8f, 0.8f);
    float* o = out + 10 * (int64_t)v;
    o[0] = pos[3 * (int64_t)v];
    o[1] = pos[3 * (int64_t)v + 1];
    o[2] = pos[3 * (int64_t)v + 2];
    o[3] = nrm[3 * (int64_t)v];
    o[4] = nrm[3 * (int64_t)v + 1];
    o[5] = nrm[3 * (int64_t)v + 2];
    o[6] = c.x;
    o[7] = c.y;
    o[8] = c.z;
    o[9] = 1.f;
}

}  // namespace

int64_t prepareBlocks(const int od[3]) { return (int64_t)od[1] * od[2]; }

void launchPrepareGrids(const PrepareDesc* descDev, int G, int64_t blocks, int refinement, int pad, int64_t* special, hipStream_t s) {
    if (blocks <= 0) return;
    if (refinement > 0) {
        kPrepareGrids<2><<<(unsigned)blocks, kBlock, 0, s>>>(descDev, G, refinement, pad, blocks, special);
        return;
    }
    // *special is zero on entry; the exact pass returns at once unless the copy set it
    kPrepareGrids<0><<<(unsigned)((blocks + kPlainRowsPerBlock - 1) / kPlainRowsPerBlock), kBlock, 0, s>>>(descDev, G, refinement, pad, blocks, special);
    kPrepareGrids<1><<<(unsigned)std::min<int64_t>(blocks, 2048), kBlock, 0, s>>>(descDev, G, refinement, pad, blocks, special);
}

void launchDmcTrianglesFromFaces(const int4* he, const int32_t* face_he, int64_t F, uint32_t* idx, hipStream_t s) {
    if (F > 0) kTriangles<<<gridFor(F), kBlock, 0, s>>>(he, face_he, F, idx);
}

void launchRenderVertices(const int4* he, int64_t H, const float* pos, const float* nrm, const uint32_t* vflags, float* out, hipStream_t s) {
    if (H > 0) kRenderVertices<<<gridFor(H), kBlock, 0, s>>>(he, H, pos, nrm, vflags, out);
}

}  // namespace mdmc
