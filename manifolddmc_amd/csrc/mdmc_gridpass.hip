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
//   1. classify the tile's cells (MC case, polygon, owned edges) into LDS;
//   2. scan the per-cell vertex and quad counts inside the tile;
//   3. get the tile's global vertex / quad offsets by decoupled look-back over
//      the tiles before it (tiles are claimed with an atomic ticket, so every
//      predecessor is already running and the wait always ends);
//   4. publish each active cell's first vertex id and edge->contour map in a
//      ring buffer that later tiles read for the quads they own;
//   5. compute the tile's dual vertices, one vertex per thread;
//   6. write the quads whose edge starts at one of the tile's cells, one per
//      thread, taking the other three vertex ids from LDS or from the ring.
// Vertices come out ordered by (cell, contour) and quads by edge key
// 3*point+axis, with no hash table and no global append counter.  Every wait is
// bounded; a timeout raises an error flag instead of hanging.
#include "mdmc_device.h"
#include "mdmc_gridpass.h"

namespace mdmc {

namespace {

constexpr int GP_THREADS = 256;
constexpr int GP_CPT = 8;
constexpr int GP_TILE = GP_THREADS * GP_CPT;
constexpr int kSpinLimit = 1 << 22;

constexpr uint64_t FLAG_AGG = 1ULL, FLAG_INC = 2ULL;
__device__ __forceinline__ uint64_t packStatus(uint64_t flag, uint64_t v, uint64_t q) { return (flag << 62) | (q << 31) | v; }
__device__ __forceinline__ uint64_t stFlag(uint64_t s) { return s >> 62; }
__device__ __forceinline__ uint64_t stV(uint64_t s) { return s & 0x7FFFFFFFULL; }
__device__ __forceinline__ uint64_t stQ(uint64_t s) { return (s >> 31) & 0x7FFFFFFFULL; }

// Polling uses relaxed agent-scope loads (they bypass the CU's L1 without invalidating
// it); one acquire fence per tile then orders the reads that depend on the flags.
__device__ __forceinline__ uint64_t loadRelaxed(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int loadRelaxedI(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void fenceAcquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void storeRelease(unsigned long long* p, uint64_t v) {
    __hip_atomic_store(p, (unsigned long long)v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void storeReleaseI(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT); }

// bounded wait for a flag set by an earlier tile (relaxed polling; the caller fences)
__device__ __forceinline__ void waitFlag(const int* flag, int* err) {
    for (int it = 0; it < kSpinLimit; ++it) {
        if (loadRelaxedI(flag)) return;
        __builtin_amdgcn_s_sleep(1);
    }
    atomicOr(err, GP_ERR_TIMEOUT);
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

__global__ void __launch_bounds__(GP_THREADS) kGridPass(GridPassArgs a) {
    __shared__ uint64_t s_poly[GP_TILE];
    __shared__ uint16_t s_vpre[GP_TILE];
    __shared__ uint16_t s_qpre[GP_TILE];
    __shared__ uint8_t s_own[GP_TILE];
    __shared__ int s_wave[GP_THREADS / 64];
    __shared__ long long s_tile;
    __shared__ long long s_voff, s_qoff;
    __shared__ int s_VT, s_QT;

    const int tid = threadIdx.x;
    if (tid == 0) s_tile = (long long)atomicAdd(a.ticket, 1u);
    __syncthreads();
    const int64_t t = s_tile;
    if (t >= a.numTiles) return;  // uniform across the block
    const GridDev& g = a.g;
    const float iso = a.iso;
    const int CX = a.CX, CY = a.CY;
    const int64_t C = a.C;
    const int64_t c0 = t * GP_TILE;

    // ---- 1. classify
    for (int k = 0; k < GP_CPT; ++k) {
        const int l = k * GP_THREADS + tid;
        const int64_t c = c0 + l;
        uint64_t w = 0;
        int own = 0;
        if (c < C) {
            const int x = (int)(c % CX), y = (int)((c / CX) % CY), z = (int)(c / ((int64_t)CX * CY));
            float s[8];
            loadCorners(g, x, y, z, s);
            int mc = 0;
#pragma unroll
            for (int q = 0; q < 8; ++q) mc |= (s[q] >= iso ? 1 : 0) << q;
            if (mc != 0 && mc != 255) {
                w = cellPolygon(mc, s, iso);
                own = ownedBits(s, iso, x, y, z);
            }
        }
        s_poly[l] = w;
        s_own[l] = (uint8_t)own;
    }
    __syncthreads();

    // ---- 2. prefix of vertex / quad counts in cell order (thread tid owns cells 8*tid .. 8*tid+7)
    int cnt = 0;
    int nvq[GP_CPT];
#pragma unroll
    for (int j = 0; j < GP_CPT; ++j) {
        const int l = tid * GP_CPT + j;
        const int nv = polyCount(s_poly[l]);
        const int nq = __popc(s_own[l] & 7);
        nvq[j] = nv | (nq << 16);
        cnt += nvq[j];
    }
    {
        const int lane = tid & 63, wid = tid >> 6;
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
        for (int q = 0; q < GP_THREADS / 64; ++q) {
            const int v = s_wave[q];
            if (q < wid) add += v;
            tot += v;
        }
        int run = x - cnt + add;
#pragma unroll
        for (int j = 0; j < GP_CPT; ++j) {
            const int l = tid * GP_CPT + j;
            s_vpre[l] = (uint16_t)(run & 0xFFFF);
            s_qpre[l] = (uint16_t)(run >> 16);
            run += nvq[j];
        }
        if (tid == 0) {
            s_VT = tot & 0xFFFF;
            s_QT = tot >> 16;
        }
    }
    __syncthreads();
    const int VT = s_VT, QT = s_QT;

    // ---- 3. decoupled look-back (wave 0)
    if (tid < 64) {
        const int lane = tid;
        if (t == 0) {
            if (lane == 0) {
                storeRelease(&a.status[0], packStatus(FLAG_INC, (uint64_t)VT, (uint64_t)QT));
                s_voff = 0;
                s_qoff = 0;
            }
        } else {
            if (lane == 0) storeRelease(&a.status[t], packStatus(FLAG_AGG, (uint64_t)VT, (uint64_t)QT));
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
                storeRelease(&a.status[t], packStatus(FLAG_INC, (uint64_t)(accV + VT) & 0x7FFFFFFFULL, (uint64_t)(accQ + QT) & 0x7FFFFFFFULL));
                s_voff = accV;
                s_qoff = accQ;
            }
        }
    }
    // ring reuse: the slots this tile overwrites held cells [c0 - R, c0 + T - R); every tile
    // that reads them (cells up to CX*CY + CX further on) must have finished
    {
        const int64_t lo = c0 - a.R + 1;
        if (lo + GP_TILE - 1 >= 0) {
            const int64_t firstReader = (lo < 0 ? 0 : lo) / GP_TILE;
            int64_t lastReader = (c0 + GP_TILE - a.R + (int64_t)CX * CY + CX) / GP_TILE;
            if (lastReader > t - 1) lastReader = t - 1;
            for (int64_t tr = firstReader + tid; tr <= lastReader; tr += GP_THREADS) waitFlag(&a.done[tr], a.err);
        }
    }
    __syncthreads();
    fenceAcquire();
    const int64_t voff = s_voff, qoff = s_qoff;

    // ---- 4. publish first vertex id + contour map of the tile's active cells
    for (int k = 0; k < GP_CPT; ++k) {
        const int l = k * GP_THREADS + tid;
        const uint64_t w = s_poly[l];
        if (w) {
            const int64_t c = c0 + l;
            a.ring[c % a.R] = make_int2((int)(voff + s_vpre[l]), (int)contourMap(w));
        }
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) storeReleaseI(&a.ready[t], 1);

    // ---- 5. dual vertices, one per thread
    for (int i = tid; i < VT; i += GP_THREADS) {
        const int l = upperIndex(s_vpre, GP_TILE, i);
        const int tt = i - s_vpre[l];
        const int64_t c = c0 + l;
        const int x = (int)(c % CX), y = (int)((c / CX) % CY), z = (int)(c / ((int64_t)CX * CY));
        float s[8];
        loadCorners(g, x, y, z, s);
        f3 wp, nn;
        dualVertex(g, iso, a.surfIter, x, y, z, s, s_poly[l], tt, wp, nn);
        const int64_t v = voff + i;
        if (v < a.capV) {
            a.pos[3 * v + 0] = wp.x;
            a.pos[3 * v + 1] = wp.y;
            a.pos[3 * v + 2] = wp.z;
            a.nrm[3 * v + 0] = nn.x;
            a.nrm[3 * v + 1] = nn.y;
            a.nrm[3 * v + 2] = nn.z;
        } else {
            atomicOr(a.err, GP_ERR_OVERFLOW);
        }
    }

    // ---- 6. quads whose edge starts at a cell of this tile.  Their other cells lie
    // 1, CX, CX+1, CX*CY, CX*CY+1 or CX*CY+CX cells back; wait once for those tiles.
    if (QT > 0) {
        if (tid < 12) {
            const int64_t CXY = (int64_t)CX * CY;
            const int64_t offs[6] = {1, CX, (int64_t)CX + 1, CXY, CXY + 1, CXY + CX};
            const int64_t off = offs[tid >> 1];
            const int64_t cell = (tid & 1) ? c0 + GP_TILE - 1 - off : c0 - off;
            if (cell >= 0) {
                const int64_t tn = cell / GP_TILE;
                if (tn < t) waitFlag(&a.ready[tn], a.err);
            }
        }
        __syncthreads();
        fenceAcquire();
    }
    for (int jq = tid; jq < QT; jq += GP_THREADS) {
        const int l = upperIndex(s_qpre, GP_TILE, jq);
        int r = jq - s_qpre[l];
        const int own = s_own[l];
        int axis = 0;
        for (int b = 0; b < 3; ++b) {
            if ((own >> b) & 1) {
                if (r == 0) { axis = b; break; }
                --r;
            }
        }
        const int edgy = (own >> 3) & 1;
        const int64_t c = c0 + l;
        const int x = (int)(c % CX), y = (int)((c / CX) % CY), z = (int)(c / ((int64_t)CX * CY));
        // own edge, then (dx, dy, dz, edge) of the three other cells around it, packed
        // 4 bits each (dx | dy<<1 | dz<<2, edge) per cell: x: (y-1: e2) (z-1: e4) (y-1,z-1: e6);
        // y: (x-1: e1) (z-1: e7) (x-1,z-1: e5); z: (x-1: e9) (y-1: e11) (x-1,y-1: e10)
        const int ownE = axis == 0 ? 0 : (axis == 1 ? 3 : 8);
        const uint32_t nbTable = axis == 0 ? 0x664422u : (axis == 1 ? 0x557411u : 0xA3B291u);
        int s0 = -1, s1 = -1, s2 = -1, s3 = -1;
        auto put = [&](int sl, int v) {
            if (sl == 0) s0 = v;
            else if (sl == 1) s1 = v;
            else if (sl == 2) s2 = v;
            else s3 = v;
        };
        put(edgeSlot(ownE, edgy), (int)(voff + s_vpre[l] + polyContourOf(s_poly[l], ownE)));
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t ent = (nbTable >> (8 * k)) & 0xFFu;
            const int d = (int)(ent & 0xF), e = (int)(ent >> 4);
            const int64_t n = (int64_t)(z - ((d >> 2) & 1)) * CX * CY + (int64_t)(y - ((d >> 1) & 1)) * CX + (x - (d & 1));
            int vid;
            if (n >= c0) {
                const int ln = (int)(n - c0);
                vid = (int)(voff + s_vpre[ln] + polyContourOf(s_poly[ln], e));
            } else {
                const int2 rent = a.ring[n % a.R];
                vid = rent.x + (int)((((uint32_t)rent.y) >> (2 * e)) & 3u);
            }
            put(edgeSlot(e, edgy), vid);
        }
        const int64_t q = qoff + jq;
        if (q < a.capQ) {
            a.quads[q] = make_int4(s0, s1, s2, s3);
            a.qattr[q] = (uint32_t)(3 * ((x & 1) | ((y & 1) << 1) | ((z & 1) << 2)) + axis);
        } else {
            atomicOr(a.err, GP_ERR_OVERFLOW);
        }
    }
    __syncthreads();
    if (tid == 0) {
        __threadfence();
        storeReleaseI(&a.done[t], 1);
    }
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

void launchGridPass(const GridPassArgs& a, hipStream_t s) {
    if (a.numTiles <= 0) return;
    kGridPass<<<(unsigned)a.numTiles, GP_THREADS, 0, s>>>(a);
}

void launchGridPassTotals(const unsigned long long* status, int64_t numTiles, int64_t* out, hipStream_t s) {
    kGridPassTotals<<<1, 1, 0, s>>>(status, numTiles, out);
}

}  // namespace mdmc
