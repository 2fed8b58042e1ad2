// mdmc_adjoint.hip -- dL/dfield from dL/d(dual vertex position), topology frozen.
//
// No reference counterpart: the reference has no backward pass.  The forward it
// differentiates is the dual vertex of CellIntersection::sliceP
// (CellIntersection.inl:368-439) without the move-to-surface iteration:
//   p_v = origin + (ijk + (1/n) sum_k (t_k E_k + C_k)) * delta,
//   t_k = (iso - s_a) / (s_b - s_a)   for edge k = (a, b) of the vertex's contour,
// with E_k the unit vector of the edge's axis.  Hence
//   dp_v/ds_a = (1/n) delta_ax E_k (t_k - 1) / (s_b - s_a),
//   dp_v/ds_b = (1/n) delta_ax E_k (-t_k) / (s_b - s_a).
// Intersection weights t_k are recomputed from the field, not stored; the forward
// keeps only the topology (cell -> first vertex and contour count, vertex -> cell).
//
// Three passes, no atomics, so the result is deterministic:
//   1. per vertex: its contribution to each of its cell's 8 corners;
//   2. per point of the processed grid: the sum over the (up to 8) cells around it;
//   3. per point of the input grid: the adjoint of padding / refinement
//      (refineGrid, UniformGrid.inl:279-319: clamp-duplicated border, trilinear).
#include "mdmc_adjoint.h"
#include "mdmc_device.h"

namespace mdmc {

namespace {

__device__ __forceinline__ int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

__global__ void kAdjointVertex(GridDev g, float iso, const int32_t* __restrict__ adjMap, const int32_t* __restrict__ adjCell, int64_t V,
                               const float* __restrict__ dpos, float* __restrict__ contrib) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int CX = g.nx - 1, CY = g.ny - 1;
    const int64_t c = adjCell[v];
    const int t = (int)(v - (adjMap[c] >> 2));
    const int x = (int)(c % CX), y = (int)((c / CX) % CY), z = (int)(c / ((int64_t)CX * CY));
    const int64_t NX = g.nx, NXY = (int64_t)g.nx * g.ny;
    const int64_t b = (int64_t)z * NXY + (int64_t)y * NX + x;
    float s[8];
    s[0] = g.v[b];
    s[1] = g.v[b + 1];
    s[2] = g.v[b + NX];
    s[3] = g.v[b + NX + 1];
    s[4] = g.v[b + NXY];
    s[5] = g.v[b + NXY + 1];
    s[6] = g.v[b + NXY + NX];
    s[7] = g.v[b + NXY + NX + 1];
    int mc = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) mc |= (s[k] >= iso ? 1 : 0) << k;
    const uint64_t w = cellPolygon(mc, s, iso);
    int j = 0;
    for (int cc = 0; cc < t; ++cc) j += polySize(w, cc);
    const int n = polySize(w, t);
    const float gw[3] = {dpos[3 * v] * g.dx, dpos[3 * v + 1] * g.dy, dpos[3 * v + 2] * g.dz};
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const float inv = 1.f / (float)n;
    for (int i = 0; i < n; ++i, ++j) {
        const int e = polyEdge(w, j);
        const int a = edgeA(e), bb = edgeB(e), ax = edgeAxis(e);
        const float den = s[bb] - s[a];
        const float tt = (iso - s[a]) / den;
        const float ga = gw[ax] * inv;
        // acc[a] += ga * (tt - 1) / den; acc[bb] += ga * (-tt) / den  (unrolled selects keep acc in registers)
        const float da = ga * (tt - 1.f) / den, db = -ga * tt / den;
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] += (k == a ? da : 0.f) + (k == bb ? db : 0.f);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) contrib[8 * v + k] = acc[k];
}

// processed point q: sum over the cells q - (dx, dy, dz) of the contributions to their corner (dx, dy, dz)
__global__ void kAdjointGather(int nx, int ny, int nz, const int32_t* __restrict__ adjMap, const float* __restrict__ contrib, float* __restrict__ out,
                               int accumulate) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t P = (int64_t)nx * ny * nz;
    if (q >= P) return;
    const int x = (int)(q % nx), y = (int)((q / nx) % ny), z = (int)(q / ((int64_t)nx * ny));
    const int CX = nx - 1, CY = ny - 1, CZ = nz - 1;
    float sum = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
        const int cx = x - dx, cy = y - dy, cz = z - dz;
        if (cx < 0 || cy < 0 || cz < 0 || cx >= CX || cy >= CY || cz >= CZ) continue;
        const int32_t m = adjMap[((int64_t)cz * CY + cy) * CX + cx];
        if (m < 0) continue;
        const int64_t fv = m >> 2;
        const int cnt = (m & 3) + 1;
        for (int i = 0; i < cnt; ++i) sum += contrib[8 * (fv + i) + k];
    }
    out[q] = accumulate ? out[q] + sum : sum;
}

// the input coordinates and weights one processed coordinate draws from (refineGrid on
// one axis, after undoing the padding): value = (1-d) * in[clamp(m)] + d * in[clamp(m+1)]
__device__ __forceinline__ float axisWeight(int u, int j, int n, float rr) {
    const float f = (float)u / rr;
    const int m = (int)floorf(f);
    const float d = f - (float)m;
    float wgt = 0.f;
    if (clampi(m, 0, n - 1) == j) wgt += 1.f - d;
    if (clampi(m + 1, 0, n - 1) == j) wgt += d;
    return wgt;
}

// processed coordinates on one axis that read input coordinate j, with their weights;
// calls f(processed coordinate, weight)
template <class Fn>
__device__ __forceinline__ void axisSources(int j, int n, int on, int r, int pad, Fn&& f) {
    const float rr = (float)(r + 1);
    const int un = on - 2 * pad;  // unpadded processed size
    const int lo = max(0, (j - 1) * (r + 1)), hi = min(un - 1, (j + 1) * (r + 1));
    for (int u = lo; u <= hi; ++u) {
        const float wgt = axisWeight(u, j, n, rr);
        if (wgt == 0.f) continue;
        if (pad) {
            f(u + 1, wgt);
            if (u == 0) f(0, wgt);
            if (u == un - 1) f(on - 1, wgt);
        } else {
            f(u, wgt);
        }
    }
}

__global__ void kPrepareAdjoint(int inx, int iny, int inz, int onx, int ony, int onz, int r, int pad, const float* __restrict__ dproc,
                                float* __restrict__ din, int accumulate) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t Pin = (int64_t)inx * iny * inz;
    if (i >= Pin) return;
    const int jx = (int)(i % inx), jy = (int)((i / inx) % iny), jz = (int)(i / ((int64_t)inx * iny));
    float sum = 0.f;
    axisSources(jz, inz, onz, r, pad, [&](int pz, float wz) {
        axisSources(jy, iny, ony, r, pad, [&](int py, float wy) {
            axisSources(jx, inx, onx, r, pad, [&](int px, float wx) {
                sum += (wz * wy * wx) * dproc[((int64_t)pz * ony + py) * onx + px];
            });
        });
    });
    din[i] = accumulate ? din[i] + sum : sum;
}

}  // namespace

void launchAdjointVertex(const GridDev& g, float iso, const int32_t* adjMap, const int32_t* adjCell, int64_t V, const float* dpos, float* contrib,
                         hipStream_t s) {
    if (V > 0) kAdjointVertex<<<gridFor(V), kBlock, 0, s>>>(g, iso, adjMap, adjCell, V, dpos, contrib);
}

void launchAdjointGather(const GridDev& g, const int32_t* adjMap, const float* contrib, float* out, int accumulate, hipStream_t s) {
    const int64_t P = (int64_t)g.nx * g.ny * g.nz;
    kAdjointGather<<<gridFor(P), kBlock, 0, s>>>(g.nx, g.ny, g.nz, adjMap, contrib, out, accumulate);
}

void launchPrepareAdjoint(const int inDims[3], const int outDims[3], int r, int pad, const float* dproc, float* din, int accumulate, hipStream_t s) {
    const int64_t Pin = (int64_t)inDims[0] * inDims[1] * inDims[2];
    kPrepareAdjoint<<<gridFor(Pin), kBlock, 0, s>>>(inDims[0], inDims[1], inDims[2], outDims[0], outDims[1], outDims[2], r, pad, dproc, din,
                                                    accumulate);
}

}  // namespace mdmc
