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
// so the grid a block works on, and with it every size below, is uniform across the block.
// The block's first point is split into (x0, y0, z0) once (uniform, scalar); each thread
// carries x0 + tid into y and z.  Without refinement the sample position is the point
// itself: x / 1 = x exactly, floor(x) = x and the weights are +0, the same values the
// general path computes, so the 8 corner loads and the trilinear blend stay (a non-finite
// neighbour still reaches the result through 0 * s).
template <bool REFINE>
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
    const int64_t i0 = (blk - p.base) * kBlock;
    const int64_t i = i0 + threadIdx.x;
    if (i >= (int64_t)onx * ony * onz) return;
    const int64_t r0 = i0 / onx;
    int x = (int)(i0 - r0 * onx) + (int)threadIdx.x;
    int y = (int)(r0 % ony);
    int z = (int)(r0 / ony);
    while (x >= onx) {  // once at most when onx >= kBlock
        x -= onx;
        if (++y == ony) {
            y = 0;
            ++z;
        }
    }
    if (pad) {
        x = clampi(x, 1, onx - 2) - 1;
        y = clampi(y, 1, ony - 2) - 1;
        z = clampi(z, 1, onz - 2) - 1;
    }
    int mx = x, my = y, mz = z;
    f3 dd = mk3(0.f, 0.f, 0.f);
    if (REFINE) {
        const float rr = (float)(refinement + 1);
        const float fx = (float)x / rr, fy = (float)y / rr, fz = (float)z / rr;
        mx = (int)floorf(fx);
        my = (int)floorf(fy);
        mz = (int)floorf(fz);
        dd = mk3(fx - (float)mx, fy - (float)my, fz - (float)mz);
    }
    const int x0 = clampi(mx, 0, inx - 1), x1 = clampi(mx + 1, 0, inx - 1);
    const int y0 = clampi(my, 0, iny - 1), y1 = clampi(my + 1, 0, iny - 1);
    const int z0 = clampi(mz, 0, inz - 1), z1 = clampi(mz + 1, 0, inz - 1);
    const int64_t nxy = (int64_t)inx * iny;
    const float* r00 = p.in + (int64_t)z0 * nxy + (int64_t)y0 * inx;
    const float* r10 = p.in + (int64_t)z0 * nxy + (int64_t)y1 * inx;
    const float* r01 = p.in + (int64_t)z1 * nxy + (int64_t)y0 * inx;
    const float* r11 = p.in + (int64_t)z1 * nxy + (int64_t)y1 * inx;
    const float s[8] = {r00[x0], r00[x1], r10[x0], r10[x1], r01[x0], r01[x1], r11[x0], r11[x1]};
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

int64_t prepareBlocks(int64_t points) { return (points + kBlock - 1) / kBlock; }

void launchPrepareGrids(const PrepareDesc* descDev, int G, int64_t blocks, int refinement, int pad, hipStream_t s) {
    if (blocks <= 0) return;
    if (refinement > 0) kPrepareGrids<true><<<(unsigned)blocks, kBlock, 0, s>>>(descDev, G, refinement, pad);
    else kPrepareGrids<false><<<(unsigned)blocks, kBlock, 0, s>>>(descDev, G, refinement, pad);
}

void launchDmcTrianglesFromFaces(const int4* he, const int32_t* face_he, int64_t F, uint32_t* idx, hipStream_t s) {
    if (F > 0) kTriangles<<<gridFor(F), kBlock, 0, s>>>(he, face_he, F, idx);
}

void launchRenderVertices(const int4* he, int64_t H, const float* pos, const float* nrm, const uint32_t* vflags, float* out, hipStream_t s) {
    if (H > 0) kRenderVertices<<<gridFor(H), kBlock, 0, s>>>(he, H, pos, nrm, vflags, out);
}

}  // namespace mdmc
