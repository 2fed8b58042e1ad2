// mdmc_gridpass.hip -- the DMC grid pass in one kernel.
//
// Restates (MarchingCubes/VisualDMC/src/DMC):
//   performDualMarchingCubes        DualMarchingCubes.cu:38-67
//   CellIntersection::sliceP        CellIntersection.inl:368-439
//   UniformGrid::computeGradients   UniformGrid.inl:325-379
//   QuadrilateralDeviceHashTable / mapQuadrilaterals   DeviceHashTables.h:288-348, DualMarchingCubes.cu:72-85
//
// The reference appends vertices with one global atomic counter and scatters
// their ids into a 64-bit keyed quad hash that a second kernel compacts.  Here
// one kernel walks the cells in linear order in tiles of 2048:
//   1. find the tile's active cells (MC case neither 0 nor 255) and compact
//      them, in cell order, into LDS; then compute polygon and owned edges of
//      the active cells only, with every lane busy;
//   2. scan the per-cell vertex and quad counts over the active cells;
//   3. get the tile's global vertex / quad offsets by decoupled look-back over
//      the tiles before it (tiles are claimed with an atomic ticket, so every
//      predecessor is already running and the wait always ends);
//   4. publish each active cell's first vertex id and edge->contour map in a
//      ring buffer that later tiles read for the quads they own;
//   5. compute the tile's dual vertices, one vertex per thread;
//   6. write the quads whose edge starts at one of the tile's cells, one per
//      thread, taking the other three vertex ids from LDS or from the ring.
// Vertices come out ordered by (cell, contour) and quads by edge key
// 3*point+axis, with no hash table and no global append counter.
//
// Cross-tile communication uses relaxed device-scope atomics only: every word
// another tile reads carries its own validity (the look-back status packs flag
// and both counts; a ring entry carries a tag naming the cell it holds), so no
// release/acquire fence -- which on this chip writes back or invalidates whole
// caches -- is needed.  Every wait is bounded; a timeout raises an error flag
// instead of hanging.
#include <algorithm>

#include "mdmc_device.h"
#include "mdmc_gridpass.h"

namespace mdmc {

namespace {

constexpr int GP_THREADS = 256;
constexpr int GP_CPT = 8;
constexpr int GP_TILE = GP_THREADS * GP_CPT;
constexpr int GP_WAVES = GP_THREADS / 64;
constexpr int GP_RUNS = GP_TILE / 64;  // runs of 64 consecutive cells, one ballot each
constexpr int kSpinLimit = 1 << 22;

constexpr uint64_t FLAG_AGG = 1ULL, FLAG_INC = 2ULL;
__device__ __forceinline__ uint64_t packStatus(uint64_t flag, uint64_t v, uint64_t q) { return (flag << 62) | (q << 31) | v; }
__device__ __forceinline__ uint64_t stFlag(uint64_t s) { return s >> 62; }
__device__ __forceinline__ uint64_t stV(uint64_t s) { return s & 0x7FFFFFFFULL; }
__device__ __forceinline__ uint64_t stQ(uint64_t s) { return (s >> 31) & 0x7FFFFFFFULL; }

__device__ __forceinline__ uint64_t loadRelaxed(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void storeRelaxed(unsigned long long* p, uint64_t v) {
    __hip_atomic_store(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int loadRelaxedI(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void storeRelaxedI(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// bounded wait for a flag set by an earlier tile
__device__ __forceinline__ void waitFlag(const int* flag, int* err) {
    for (int it = 0; it < kSpinLimit; ++it) {
        if (loadRelaxedI(flag)) return;
        __builtin_amdgcn_s_sleep(1);
    }
    atomicOr(err, GP_ERR_TIMEOUT);
}

// Ring entry of cell c: bits 0-30 first vertex id, 31-54 the contour (2 bits) of each of
// the 12 edges, 55-63 tag = (c >> ringShift) + 1.  The ring is cleared before the pass and
// holds fewer than 511 generations, so a tag never matches an older or missing entry.
__device__ __forceinline__ uint64_t ringTag(int64_t c, int shift) { return (uint64_t)((c >> shift) + 1) & 0x1FFULL; }

// the problem that holds global tile t, or global cell c (binary search over the bases)
__device__ __forceinline__ int problemOfTile(const GridPassArgs& a, int64_t t) {
    int lo = 0, hi = a.nProblems;  // probs[lo].tileBase <= t
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.probs[mid].tileBase <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ int64_t tileOfCell(const GridPassArgs& a, int64_t c) {
    int lo = 0, hi = a.nProblems;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.probs[mid].cellBase <= c) lo = mid;
        else hi = mid;
    }
    const ProblemDev& p = a.probs[lo];
    const int64_t local = c - p.cellBase;
    return p.tileBase + (local < p.cells ? local : p.cells - 1) / GP_TILE;
}

__device__ __forceinline__ int ringVertex(const GridPassArgs& a, int64_t n, int e) {
    const uint64_t tag = ringTag(n, a.ringShift);
    const unsigned long long* p = a.ring + (n & (a.R - 1));
    uint64_t v = loadRelaxed(p);
    for (int it = 0; (v >> 55) != tag; ++it) {
        if (it >= kSpinLimit) {
            atomicOr(a.err, GP_ERR_TIMEOUT);
            return -1;
        }
        __builtin_amdgcn_s_sleep(1);
        v = loadRelaxed(p);
    }
    return (int)(v & 0x7FFFFFFFULL) + (int)((v >> (31 + 2 * e)) & 3ULL);
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

__device__ __forceinline__ int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }
__device__ __forceinline__ bool crossed(float a, float b, float iso) { return (a >= iso) != (b >= iso); }

// bit 0..2: owned x / y / z edge (edges 0, 3, 8) crossed with all four cells present; bit 3: edgy (s0 < iso)
__device__ __forceinline__ int ownedBits(const float s[8], float iso, int x, int y, int z) {
    int m = 0;
    if (y >= 1 && z >= 1 && crossed(s[0], s[1], iso)) m |= 1;
    if (x >= 1 && z >= 1 && crossed(s[0], s[2], iso)) m |= 2;
    if (x >= 1 && y >= 1 && crossed(s[0], s[4], iso)) m |= 4;
    if (s[0] < iso) m |= 8;
    return m;
}

// contour index of every crossed edge, 2 bits per edge
__device__ __forceinline__ uint32_t contourMap(uint64_t w) {
    uint32_t map = 0;
    int j = 0;
    for (int c = 0; c < 4; ++c) {
        const int s = polySize(w, c);
        for (int i = 0; i < s; ++i, ++j) map |= (uint32_t)c << (2 * polyEdge(w, j));
    }
    return map;
}

// the last index l in [0, n) with pre[l] <= key (pre non-decreasing, pre[0] = 0 <= key)
__device__ __forceinline__ int upperIndex(const uint16_t* pre, int n, int key) {
    int lo = 0, hi = n;  // invariant: pre[lo] <= key, answer in [lo, hi)
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((int)pre[mid] <= key) lo = mid;
        else hi = mid;
    }
    return lo;
}

// sliceP for contour t of one cell: dual vertex position (world) and normal
__device__ __forceinline__ void dualVertex(const GridDev& g, float iso, int surfIter, int x, int y, int z, const float s[8], uint64_t w, int t, f3& wp,
                                           f3& nn) {
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
    int j = 0;
    for (int c = 0; c < t; ++c) j += polySize(w, c);
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
    nn = normalize3(mk3(trilinear(gxv, p), trilinear(gyv, p), trilinear(gzv, p)));
    wp = mk3(g.ox, g.oy, g.oz) + (mk3((float)x, (float)y, (float)z) + p) * mk3(g.dx, g.dy, g.dz);
}

// coordinates of cell c0 + l from those of c0 (x0, y0, z0) and l < GP_TILE, in 32-bit arithmetic
struct TileBase {
    int x0, y0, z0;
};
__device__ __forceinline__ void cellXYZ(const TileBase& b, int l, int CX, int CY, int& x, int& y, int& z) {
    const unsigned xs = (unsigned)(b.x0 + l);
    const unsigned r = xs / (unsigned)CX;
    x = (int)(xs - r * (unsigned)CX);
    const unsigned ys = (unsigned)b.y0 + r;
    const unsigned zz = ys / (unsigned)CY;
    y = (int)(ys - zz * (unsigned)CY);
    z = b.z0 + (int)zz;
}

__global__ void __launch_bounds__(GP_THREADS) kGridPass(GridPassArgs a) {
    // per active cell (index in cell order): polygon, owned-edge bits, tile cell, scans
    __shared__ uint64_t s_poly[GP_TILE];
    __shared__ uint16_t s_cell[GP_TILE];
    __shared__ uint16_t s_vpre[GP_TILE];
    __shared__ uint16_t s_qpre[GP_TILE];  // also holds the MC case of each tile cell in phase 1
    __shared__ uint8_t s_own[GP_TILE];
    // per tile cell: its active index (valid for active cells)
    __shared__ uint16_t s_aidx[GP_TILE];
    __shared__ unsigned long long s_mask[GP_RUNS];
    __shared__ int s_mpre[GP_RUNS + 1];
    __shared__ int s_wave[GP_WAVES];
    __shared__ long long s_tile;
    __shared__ int s_prob;
    __shared__ long long s_voff, s_qoff;
    uint8_t* const s_mc = reinterpret_cast<uint8_t*>(s_qpre);

    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    if (tid == 0) {
        const long long tk = (long long)atomicAdd(a.ticket, 1u);
        s_tile = tk;
        s_prob = tk < a.numTiles ? problemOfTile(a, tk) : 0;
    }
    __syncthreads();
    const int64_t t = s_tile;
    if (t >= a.numTiles) return;  // uniform across the block
    const int prob = s_prob;
    const ProblemDev& P = a.probs[prob];
    const GridDev g = P.g;
    const float iso = P.iso;
    const int CX = P.CX, CY = P.CY;
    const int64_t C = P.cells;
    const int64_t c0 = (t - P.tileBase) * GP_TILE;  // first cell of the tile inside its problem
    const int64_t gc0 = P.cellBase + c0;             // ... and in the concatenation

    TileBase tb;
    {
        const int64_t r0 = c0 / CX;
        tb.x0 = (int)(c0 - r0 * CX);
        tb.y0 = (int)(r0 % CY);
        tb.z0 = (int)(r0 / CY);
    }

    // ---- 1a. MC case of every cell; a ballot per run of 64 cells marks the active ones
#pragma unroll
    for (int k = 0; k < GP_CPT; ++k) {
        const int l = k * GP_THREADS + tid;
        int mc = 0;
        if (c0 + l < C) {
            int x, y, z;
            cellXYZ(tb, l, CX, CY, x, y, z);
            float s[8];
            loadCorners(g, x, y, z, s);
#pragma unroll
            for (int q = 0; q < 8; ++q) mc |= (s[q] >= iso ? 1 : 0) << q;
        }
        const uint64_t m = __ballot(mc != 0 && mc != 255);
        if (lane == 0) s_mask[k * GP_WAVES + wid] = m;
        s_mc[l] = (uint8_t)mc;
    }
    __syncthreads();
    if (tid < 64) {
        const int v = tid < GP_RUNS ? __popcll(s_mask[tid]) : 0;
        int x = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d, 64);
            if (tid >= d) x += y;
        }
        if (tid < GP_RUNS) s_mpre[tid] = x - v;
        if (tid == GP_RUNS - 1) s_mpre[GP_RUNS] = x;
    }
    __syncthreads();
    const int NA = s_mpre[GP_RUNS];
#pragma unroll
    for (int k = 0; k < GP_CPT; ++k) {
        const int run = k * GP_WAVES + wid;
        const uint64_t m = s_mask[run];
        if ((m >> lane) & 1ULL) {
            const int ai = s_mpre[run] + __popcll(m & ((1ULL << lane) - 1ULL));
            const int l = k * GP_THREADS + tid;
            s_cell[ai] = (uint16_t)l;
            s_aidx[l] = (uint16_t)ai;
        }
    }
    __syncthreads();

    // ---- 1b. polygon and owned edges of the active cells
    for (int ai = tid; ai < NA; ai += GP_THREADS) {
        const int l = s_cell[ai];
        int x, y, z;
        cellXYZ(tb, l, CX, CY, x, y, z);
        float s[8];
        loadCorners(g, x, y, z, s);
        s_poly[ai] = cellPolygon(s_mc[l], s, iso);
        s_own[ai] = (uint8_t)ownedBits(s, iso, x, y, z);
    }
    __syncthreads();

    // ---- 2. prefix of vertex / quad counts over the active cells (thread tid owns 8*tid .. 8*tid+7)
    int cnt = 0, VT = 0, QT = 0;
    int nvq[GP_CPT];
#pragma unroll
    for (int j = 0; j < GP_CPT; ++j) {
        const int ai = tid * GP_CPT + j;
        nvq[j] = 0;
        if (ai < NA) nvq[j] = polyCount(s_poly[ai]) | (__popc(s_own[ai] & 7) << 16);
        cnt += nvq[j];
    }
    {
        int x = cnt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (lane == 63) s_wave[wid] = x;
        __syncthreads();
        int add = 0, tot = 0;
#pragma unroll
        for (int q = 0; q < GP_WAVES; ++q) {
            const int v = s_wave[q];
            if (q < wid) add += v;
            tot += v;
        }
        int run = x - cnt + add;
#pragma unroll
        for (int j = 0; j < GP_CPT; ++j) {
            const int ai = tid * GP_CPT + j;
            if (ai < NA) {
                s_vpre[ai] = (uint16_t)(run & 0xFFFF);
                s_qpre[ai] = (uint16_t)(run >> 16);
            }
            run += nvq[j];
        }
        VT = tot & 0xFFFF;
        QT = tot >> 16;
    }

    // ---- 3. decoupled look-back (wave 0)
    if (tid < 64) {
        if (t == 0) {
            if (lane == 0) {
                storeRelaxed(&a.status[0], packStatus(FLAG_INC, (uint64_t)VT, (uint64_t)QT));
                s_voff = 0;
                s_qoff = 0;
            }
        } else {
            if (lane == 0) storeRelaxed(&a.status[t], packStatus(FLAG_AGG, (uint64_t)VT, (uint64_t)QT));
            int64_t accV = 0, accQ = 0;
            int64_t i = t - 1;
            int spins = 0;
            bool doneLB = false;
            while (!doneLB) {
                const int64_t jdx = i - lane;
                const uint64_t st = (jdx >= 0) ? loadRelaxed(&a.status[jdx]) : packStatus(FLAG_INC, 0, 0);
                const uint64_t f = stFlag(st);
                const uint64_t incMask = __ballot(f == FLAG_INC);
                const uint64_t zeroMask = __ballot(f == 0);
                const int firstInc = incMask ? __ffsll((unsigned long long)incMask) - 1 : 64;
                const uint64_t below = (firstInc >= 64) ? ~0ULL : ((1ULL << firstInc) - 1ULL);
                if (zeroMask & (below | (firstInc < 64 ? (1ULL << firstInc) : 0ULL))) {
                    if (++spins > kSpinLimit) {
                        if (lane == 0) atomicOr(a.err, GP_ERR_TIMEOUT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                int64_t v = (lane <= firstInc) ? (int64_t)stV(st) : 0;
                int64_t q = (lane <= firstInc) ? (int64_t)stQ(st) : 0;
#pragma unroll
                for (int d = 32; d > 0; d >>= 1) {
                    v += __shfl_down(v, d, 64);
                    q += __shfl_down(q, d, 64);
                }
                v = __shfl(v, 0, 64);
                q = __shfl(q, 0, 64);
                accV += v;
                accQ += q;
                if (firstInc < 64) doneLB = true;
                else i -= 64;
            }
            if (lane == 0) {
                if (accV + VT > 0x7FFFFFFFLL || accQ + QT > 0x7FFFFFFFLL) atomicOr(a.err, GP_ERR_RANGE);
                storeRelaxed(&a.status[t], packStatus(FLAG_INC, (uint64_t)(accV + VT) & 0x7FFFFFFFULL, (uint64_t)(accQ + QT) & 0x7FFFFFFFULL));
                s_voff = accV;
                s_qoff = accQ;
            }
        }
    }
    // ring reuse: the slots this tile overwrites held cells [gc0 - R, gc0 + T - R) of the
    // concatenation; every tile that reads them (cells up to `reach` further on) must have finished
    {
        const int64_t oldHi = gc0 + GP_TILE - 1 - a.R;
        if (oldHi >= 0) {
            const int64_t lo = gc0 - a.R + 1;
            const int64_t firstReader = tileOfCell(a, lo < 0 ? 0 : lo);
            const int64_t hiCell = oldHi + a.reach;
            int64_t lastReader = tileOfCell(a, hiCell < a.totalCells ? hiCell : a.totalCells - 1);
            if (lastReader > t - 1) lastReader = t - 1;
            for (int64_t tr = firstReader + tid; tr <= lastReader; tr += GP_THREADS) waitFlag(&a.done[tr], a.err);
        }
    }
    __syncthreads();
    const int64_t voff = s_voff, qoff = s_qoff;

    // ---- 4. publish first vertex id + contour map of the tile's active cells
    for (int ai = tid; ai < NA; ai += GP_THREADS) {
        const int64_t c = gc0 + s_cell[ai];
        const uint64_t ent = (ringTag(c, a.ringShift) << 55) | ((uint64_t)contourMap(s_poly[ai]) << 31) | (uint64_t)((voff + s_vpre[ai]) & 0x7FFFFFFF);
        storeRelaxed(a.ring + (c & (a.R - 1)), ent);
    }
    // for the adjoint: cell -> (first vertex << 2 | contours - 1), -1 for empty cells
    if (a.adjMap) {
#pragma unroll
        for (int k = 0; k < GP_CPT; ++k) {
            const int l = k * GP_THREADS + tid;
            const int64_t c = c0 + l;
            if (c < C) {
                int val = -1;
                if ((s_mask[k * GP_WAVES + wid] >> lane) & 1ULL) {
                    const int ai = s_aidx[l];
                    val = (int)(((voff + s_vpre[ai]) << 2) | (int64_t)(polyCount(s_poly[ai]) - 1));
                }
                a.adjMap[c] = val;
            }
        }
    }

    // ---- 5. dual vertices, one per thread
    for (int i = tid; i < VT; i += GP_THREADS) {
        const int ai = upperIndex(s_vpre, NA, i);
        const int tt = i - s_vpre[ai];
        int x, y, z;
        cellXYZ(tb, s_cell[ai], CX, CY, x, y, z);
        float s[8];
        loadCorners(g, x, y, z, s);
        f3 wp, nn;
        dualVertex(g, iso, a.surfIter, x, y, z, s, s_poly[ai], tt, wp, nn);
        const int64_t v = voff + i;
        if (v < a.capV) {
            a.pos[3 * v + 0] = wp.x;
            a.pos[3 * v + 1] = wp.y;
            a.pos[3 * v + 2] = wp.z;
            a.nrm[3 * v + 0] = nn.x;
            a.nrm[3 * v + 1] = nn.y;
            a.nrm[3 * v + 2] = nn.z;
            if (a.adjCell) a.adjCell[v] = (int)(c0 + s_cell[ai]);
            if (a.vpid) a.vpid[v] = prob;
        } else {
            atomicOr(a.err, GP_ERR_OVERFLOW);
        }
    }

    // ---- 6. quads whose edge starts at a cell of this tile.  Their other cells lie
    // 1, CX, CX+1, CX*CY, CX*CY+1 or CX*CY+CX cells back: in LDS or in the ring.
    for (int jq = tid; jq < QT; jq += GP_THREADS) {
        const int ai = upperIndex(s_qpre, NA, jq);
        int r = jq - s_qpre[ai];
        const int own = s_own[ai];
        int axis = 0;
        for (int b = 0; b < 3; ++b) {
            if ((own >> b) & 1) {
                if (r == 0) {
                    axis = b;
                    break;
                }
                --r;
            }
        }
        const int edgy = (own >> 3) & 1;
        int x, y, z;
        cellXYZ(tb, s_cell[ai], CX, CY, x, y, z);
        // own edge, then (dx, dy, dz, edge) of the three other cells around it, packed
        // 4 bits each (dx | dy<<1 | dz<<2, edge) per cell: x: (y-1: e2) (z-1: e4) (y-1,z-1: e6);
        // y: (x-1: e1) (z-1: e7) (x-1,z-1: e5); z: (x-1: e9) (y-1: e11) (x-1,y-1: e10)
        const int ownE = axis == 0 ? 0 : (axis == 1 ? 3 : 8);
        const uint32_t nbTable = axis == 0 ? 0x664422u : (axis == 1 ? 0x557411u : 0xA3B291u);
        int s0 = -1, s1 = -1, s2 = -1, s3 = -1;
        {
            const int sl = edgeSlot(ownE, edgy);
            const int vid = (int)(voff + s_vpre[ai] + polyContourOf(s_poly[ai], ownE));
            s0 = sl == 0 ? vid : s0;
            s1 = sl == 1 ? vid : s1;
            s2 = sl == 2 ? vid : s2;
            s3 = sl == 3 ? vid : s3;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t ent = (nbTable >> (8 * k)) & 0xFFu;
            const int d = (int)(ent & 0xF), e = (int)(ent >> 4);
            const int64_t n = (int64_t)(z - ((d >> 2) & 1)) * CX * CY + (int64_t)(y - ((d >> 1) & 1)) * CX + (x - (d & 1));
            int vid;
            if (n >= c0) {
                const int an = s_aidx[n - c0];
                vid = (int)(voff + s_vpre[an] + polyContourOf(s_poly[an], e));
            } else {
                vid = ringVertex(a, P.cellBase + n, e);
            }
            const int sl = edgeSlot(e, edgy);
            s0 = sl == 0 ? vid : s0;
            s1 = sl == 1 ? vid : s1;
            s2 = sl == 2 ? vid : s2;
            s3 = sl == 3 ? vid : s3;
        }
        const int64_t q = qoff + jq;
        if (q < a.capQ) {
            a.quads[q] = make_int4(s0, s1, s2, s3);
            a.qattr[q] = (uint32_t)(3 * ((x & 1) | ((y & 1) << 1) | ((z & 1) << 2)) + axis);
        } else {
            atomicOr(a.err, GP_ERR_OVERFLOW);
        }
    }
    // every ring value read above has been consumed by a store, so the reads are complete
    __syncthreads();
    if (tid == 0) storeRelaxedI(&a.done[t], 1);
}

__global__ void kGridPassTotals(const unsigned long long* status, int64_t numTiles, int64_t* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const uint64_t s = status[numTiles - 1];
        out[0] = (int64_t)stV(s);
        out[1] = (int64_t)stQ(s);
        out[2] = (int64_t)stFlag(s);
    }
}

}  // namespace

int gridPassTile() { return GP_TILE; }

void gridPassRing(int64_t totalCells, int64_t reach, int64_t* R, int* shift) {
    // at least 8 z-slices of cells (the quads of a tile reach one slice back), and few
    // enough generations that the 9-bit tag of a ring entry never wraps
    const int64_t want = std::max<int64_t>({8 * (reach + GP_TILE), (totalCells + 509) / 510, (int64_t)1 << 16});
    int s = 16;
    while (((int64_t)1 << s) < want) ++s;
    *R = (int64_t)1 << s;
    *shift = s;
}

void launchGridPass(const GridPassArgs& a, hipStream_t s) {
    if (a.numTiles <= 0) return;
    kGridPass<<<(unsigned)a.numTiles, GP_THREADS, 0, s>>>(a);
}

void launchGridPassTotals(const unsigned long long* status, int64_t numTiles, int64_t* out, hipStream_t s) {
    kGridPassTotals<<<1, 1, 0, s>>>(status, numTiles, out);
}

}  // namespace mdmc
