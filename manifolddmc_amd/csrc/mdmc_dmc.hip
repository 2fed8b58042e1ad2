// mdmc_dmc.hip -- the grid pass: padding/refinement, cell classification and
// dual-vertex generation, quad assembly.
//
// Restates (MarchingCubes/VisualDMC/src/DMC):
//   refineGrid                      UniformGrid.inl:279-319
//   performDualMarchingCubes        DualMarchingCubes.cu:38-67
//   CellIntersection::sliceP        CellIntersection.inl:368-439
//   UniformGrid::computeGradients   UniformGrid.inl:325-379
//   mapQuadrilaterals               DualMarchingCubes.cu:72-85
// The reference appends vertices with an atomic counter and scatters them into
// a 64-bit-keyed quad hash.  Here a counting pass and a prefix sum give every
// cell its vertex range up front (vertices ordered by cell, then contour) and
// the cell at the lower end of a crossed edge writes that edge's quad directly
// (quads ordered by edge key 3*point+axis), so no hash table and no atomics
// are needed and the output is deterministic.
#include "mdmc_device.h"
#include "mdmc_internal.h"

namespace mdmc {

namespace {

__device__ __forceinline__ int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

__global__ void __launch_bounds__(kBlock) kPrepareGrid(const float* __restrict__ in, int inx, int iny, int inz, float* __restrict__ out,
                                                       int onx, int ony, int onz, int refinement, int pad) {
    const int64_t total = (int64_t)onx * ony * onz;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
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
        s[k] = in[(int64_t)cz * inx * iny + (int64_t)cy * inx + cx];
    }
    const f3 d = mk3(fx - (float)mx, fy - (float)my, fz - (float)mz);
    out[i] = trilinear(s, d);
}

__device__ __forceinline__ void loadCorners(const GridDev& g, int x, int y, int z, float s[8]) {
    const int64_t nx = g.nx, nxy = (int64_t)g.nx * g.ny;
    const int64_t b = (int64_t)z * nxy + (int64_t)y * nx + x;
    s[0] = g.v[b];
    s[1] = g.v[b + 1];
    s[2] = g.v[b + nx];
    s[3] = g.v[b + nx + 1];
    s[4] = g.v[b + nxy];
    s[5] = g.v[b + nxy + 1];
    s[6] = g.v[b + nxy + nx];
    s[7] = g.v[b + nxy + nx + 1];
}

__device__ __forceinline__ int mcCase(const float s[8], float iso) {
    int mc = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) mc |= (s[k] >= iso ? 1 : 0) << k;
    return mc;
}

__device__ __forceinline__ bool crossed(float a, float b, float iso) { return (a >= iso) != (b >= iso); }

// owned quads of the cell at (x,y,z): its edges 0 (x), 3 (y), 8 (z) when they are
// crossed and all four cells around them exist (the quad hash bucket is complete)
__device__ __forceinline__ int ownedMask(const float s[8], float iso, int x, int y, int z) {
    int m = 0;
    if (y >= 1 && z >= 1 && crossed(s[0], s[1], iso)) m |= 1;
    if (x >= 1 && z >= 1 && crossed(s[0], s[2], iso)) m |= 2;
    if (x >= 1 && y >= 1 && crossed(s[0], s[4], iso)) m |= 4;
    return m;
}

// block-wide exclusive scan of one int per thread (256 threads, 4 waves)
__device__ __forceinline__ int blockExclusive(int v, int* lds, int& total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(x, d, 64);
        if (lane >= d) x += t;
    }
    if (lane == 63) lds[wid] = x;
    __syncthreads();
    int add = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        const int s = lds[w];
        if (w < wid) add += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return x - v + add;
}

__global__ void __launch_bounds__(kBlock) kDmcCount(GridDev g, float iso, uint16_t* __restrict__ lp16, int32_t* __restrict__ blockV,
                                                    int32_t* __restrict__ blockQ) {
    __shared__ int lds[8];
    const int CX = g.nx - 1, CY = g.ny - 1, CZ = g.nz - 1;
    const int64_t C = (int64_t)CX * CY * CZ;
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    int nv = 0, nq = 0;
    if (c < C) {
        const int x = (int)(c % CX), y = (int)((c / CX) % CY), z = (int)(c / ((int64_t)CX * CY));
        float s[8];
        loadCorners(g, x, y, z, s);
        const int mc = mcCase(s, iso);
        if (mc != 0 && mc != 255) {
            nv = polyCount(cellPolygon(mc, s, iso));
            nq = __popc(ownedMask(s, iso, x, y, z));
        }
    }
    int tv = 0, tq = 0;
    const int ev = blockExclusive(nv, lds, tv);
    blockExclusive(nq, lds + 4, tq);
    if (c < C) lp16[c] = (uint16_t)ev;
    if (threadIdx.x == 0) {
        blockV[blockIdx.x] = tv;
        blockQ[blockIdx.x] = tq;
    }
}

// vertex id of the contour of cell (x,y,z) that holds its local edge e
__device__ __forceinline__ int32_t neighborVertex(const GridDev& g, float iso, int x, int y, int z, int e, const uint16_t* __restrict__ lp16,
                                                  const int64_t* __restrict__ blockVOff) {
    const int CX = g.nx - 1, CY = g.ny - 1;
    const int64_t c = (int64_t)z * CX * CY + (int64_t)y * CX + x;
    float s[8];
    loadCorners(g, x, y, z, s);
    const uint64_t w = cellPolygon(mcCase(s, iso), s, iso);
    return (int32_t)(blockVOff[c / kBlock] + lp16[c] + polyContourOf(w, e));
}

__global__ void __launch_bounds__(kBlock) kDmcGenerate(GridDev g, float iso, int surfIter, const uint16_t* __restrict__ lp16,
                                                       const int64_t* __restrict__ blockVOff, const int64_t* __restrict__ blockQOff,
                                                       float* __restrict__ pos, float* __restrict__ nrm, int4* __restrict__ quads,
                                                       uint32_t* __restrict__ qattr) {
    __shared__ int lds[4];
    const int CX = g.nx - 1, CY = g.ny - 1, CZ = g.nz - 1;
    const int64_t C = (int64_t)CX * CY * CZ;
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    int x = 0, y = 0, z = 0, mc = 0, owned = 0;
    float s[8];
    uint64_t w = 0;
    if (c < C) {
        x = (int)(c % CX);
        y = (int)((c / CX) % CY);
        z = (int)(c / ((int64_t)CX * CY));
        loadCorners(g, x, y, z, s);
        mc = mcCase(s, iso);
        if (mc != 0 && mc != 255) {
            w = cellPolygon(mc, s, iso);
            owned = ownedMask(s, iso, x, y, z);
        }
    }
    int tq = 0;
    const int qex = blockExclusive(__popc(owned), lds, tq);
    if (c >= C || w == 0) return;

    const int64_t vid0 = blockVOff[blockIdx.x] + lp16[c];
    // ---- dual vertices (sliceP)
    {
        const int nx = g.nx, ny = g.ny, nz = g.nz;
        const int64_t NX = nx, NXY = (int64_t)nx * ny;
        auto at = [&](int px, int py, int pz) { return g.v[(int64_t)pz * NXY + (int64_t)py * NX + px]; };
        float gxv[8], gyv[8], gzv[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int cx = k & 1, cy = (k >> 1) & 1, cz = (k >> 2) & 1;
            const int px = x + cx, py = y + cy, pz = z + cz;
            const float gx = cx ? at(clampi(px + 1, 0, nx - 1), py, pz) - s[k & ~1] : s[k | 1] - at(clampi(px - 1, 0, nx - 1), py, pz);
            const float gy = cy ? at(px, clampi(py + 1, 0, ny - 1), pz) - s[k & ~2] : s[k | 2] - at(px, clampi(py - 1, 0, ny - 1), pz);
            const float gz = cz ? at(px, py, clampi(pz + 1, 0, nz - 1)) - s[k & ~4] : s[k | 4] - at(px, py, clampi(pz - 1, 0, nz - 1));
            gxv[k] = (0.5f * gx) / g.dx;
            gyv[k] = (0.5f * gy) / g.dy;
            gzv[k] = (0.5f * gz) / g.dz;
        }
        const int n = polyCount(w);
        int j = 0;
        for (int t = 0; t < n; ++t) {
            const int sz = polySize(w, t);
            f3 p = mk3(0.f, 0.f, 0.f);
            for (int i = 0; i < sz; ++i, ++j) {
                const int e = polyEdge(w, j);
                const int a = edgeA(e), b = edgeB(e), ax = edgeAxis(e);
                const float tt = (iso - s[a]) / (s[b] - s[a]);
                f3 off;
                off.x = 0.f + (tt * (ax == 0 ? 1.f : 0.f) + (float)(a & 1));
                off.y = 0.f + (tt * (ax == 1 ? 1.f : 0.f) + (float)((a >> 1) & 1));
                off.z = 0.f + (tt * (ax == 2 ? 1.f : 0.f) + (float)((a >> 2) & 1));
                p = p + off;
            }
            p = divs(p, (float)sz);
            if (surfIter) movePointToSurface(s, iso, p);
            f3 nn = mk3(trilinear(gxv, p), trilinear(gyv, p), trilinear(gzv, p));
            nn = normalize3(nn);
            const f3 wp = mk3(g.ox, g.oy, g.oz) + (mk3((float)x, (float)y, (float)z) + p) * mk3(g.dx, g.dy, g.dz);
            const int64_t v = vid0 + t;
            pos[3 * v + 0] = wp.x;
            pos[3 * v + 1] = wp.y;
            pos[3 * v + 2] = wp.z;
            nrm[3 * v + 0] = nn.x;
            nrm[3 * v + 1] = nn.y;
            nrm[3 * v + 2] = nn.z;
        }
    }
    // ---- quads owned by this cell, in axis order x, y, z (edge key 3*point + axis)
    if (owned) {
        int64_t q = blockQOff[blockIdx.x] + qex;
        const int edgy = (s[0] < iso) ? 1 : 0;
        const int base = (x & 1) | ((y & 1) << 1) | ((z & 1) << 2);
        // per axis: own edge, then (dx,dy,dz,edge) of the three other cells around the edge
        const int ownE[3] = {0, 3, 8};
        const int nb[3][3][4] = {{{0, 1, 0, 2}, {0, 0, 1, 4}, {0, 1, 1, 6}},
                                 {{1, 0, 0, 1}, {0, 0, 1, 7}, {1, 0, 1, 5}},
                                 {{1, 0, 0, 9}, {0, 1, 0, 11}, {1, 1, 0, 10}}};
        for (int a = 0; a < 3; ++a) {
            if (!((owned >> a) & 1)) continue;
            int slot[4];
            slot[edgeSlot(ownE[a], edgy)] = (int32_t)(vid0 + polyContourOf(w, ownE[a]));
            for (int k = 0; k < 3; ++k) {
                const int e = nb[a][k][3];
                slot[edgeSlot(e, edgy)] = neighborVertex(g, iso, x - nb[a][k][0], y - nb[a][k][1], z - nb[a][k][2], e, lp16, blockVOff);
            }
            quads[q] = make_int4(slot[0], slot[1], slot[2], slot[3]);
            qattr[q] = (uint32_t)(3 * base + a);  // QuadrilateralAttribute 0x1F after setColor(colour)
            ++q;
        }
    }
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

void launchPrepareGrid(const float* in, const int inDims[3], float* out, const int outDims[3], int refinement, int pad, hipStream_t s) {
    const int64_t total = (int64_t)outDims[0] * outDims[1] * outDims[2];
    kPrepareGrid<<<gridFor(total), kBlock, 0, s>>>(in, inDims[0], inDims[1], inDims[2], out, outDims[0], outDims[1], outDims[2], refinement, pad);
}

void launchDmcCount(const GridDev& g, float iso, uint16_t* lp16, int32_t* blockV, int32_t* blockQ, hipStream_t s) {
    const int64_t C = (int64_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1);
    kDmcCount<<<gridFor(C), kBlock, 0, s>>>(g, iso, lp16, blockV, blockQ);
}

void launchDmcGenerate(const GridDev& g, float iso, int surfIter, const uint16_t* lp16, const int64_t* blockVOff, const int64_t* blockQOff,
                       float* pos, float* nrm, int4* quads, uint32_t* qattr, hipStream_t s) {
    const int64_t C = (int64_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1);
    kDmcGenerate<<<gridFor(C), kBlock, 0, s>>>(g, iso, surfIter, lp16, blockVOff, blockQOff, pos, nrm, quads, qattr);
}

void launchDmcTrianglesFromFaces(const int4* he, const int32_t* face_he, int64_t F, uint32_t* idx, hipStream_t s) {
    if (F > 0) kTriangles<<<gridFor(F), kBlock, 0, s>>>(he, face_he, F, idx);
}

}  // namespace mdmc
