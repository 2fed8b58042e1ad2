// mdmc_gridpass.hip -- the DMC grid pass: classify, scan, emit.
//
// Restates (MarchingCubes/VisualDMC/src/DMC):
//   performDualMarchingCubes        DualMarchingCubes.cu:38-67
//   CellIntersection::sliceP        CellIntersection.inl:368-439
//   UniformGrid::computeGradients   UniformGrid.inl:325-379
//   QuadrilateralDeviceHashTable / mapQuadrilaterals   DeviceHashTables.h:288-348, DualMarchingCubes.cu:72-85
//
// The reference appends vertices with one global atomic counter and scatters
// their ids into a 64-bit keyed quad hash that a second kernel compacts.  Here the
// cells are cut into tiles of consecutive cells (1024, 2048 or 4096) and three
// launches with no waits between blocks build the same arrays:
//   classify  per tile: MC case of every cell, a wave ballot per 64 cells and a
//             block scan compact the active cells (case neither 0 nor 255) in
//             cell order; their polygon and owned edges are computed with every
//             lane busy and stored per tile, with the tile's vertex / quad counts;
//             each active cell also gets a 32-bit map entry (its vertex index in
//             the tile, the contour of the edges later cells read);
//   scan      exclusive prefix of the tile counts: every tile's first vertex and
//             quad id (the only host read-back: the exact output sizes);
//   emit      per tile: each vertex's cell and contour, and the quads whose edge
//             starts at one of the tile's cells, taking the other three vertex ids
//             from LDS or from the map entries of earlier tiles; then one thread
//             per vertex over the whole pass computes the dual vertex (edge mean,
//             move-to-surface, gradient normal) -- a tile holds too few vertices
//             to keep a block's lanes busy on that long computation.
// Vertices come out ordered by (cell, contour) and quads by edge key 3*point+axis,
// with no hash table and no global append counter.  Blocks take tiles either in
// launch order or XCD-chunked (block b runs on XCD b % 8; each XCD gets one
// contiguous run of tiles, so the rows and slices a tile shares with its
// neighbours stay in that XCD's L2).
#include <algorithm>

#include "mdmc_arena.h"
#include "mdmc_device.h"
#include "mdmc_gridpass.h"

namespace mdmc {

namespace {

constexpr int GP_THREADS = 256;
constexpr int GP_WAVES = GP_THREADS / 64;

// the problem that holds global tile t (binary search over the tile bases)
__device__ __forceinline__ int problemOfTile(const ProblemDev* probs, int n, int64_t t) {
    int lo = 0, hi = n;  // probs[lo].tileBase <= t
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (probs[mid].tileBase <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// block b -> tile: launch order, or XCD-chunked (blocks b, b + 8, b + 16, ... run on one XCD
// and take one contiguous run of numTiles / 8 tiles, the first numTiles % 8 runs one longer)
__device__ __forceinline__ int64_t tileOfBlock(int64_t b, int64_t numTiles, bool chunked) {
    if (!chunked) return b;
    const int64_t x = b & 7, i = b >> 3, q = numTiles >> 3, r = numTiles & 7;
    return x * q + (x < r ? x : r) + i;
}

// Polygons of all 256 cases with the decider outcomes folded in: case mc's polygon is
// c_polyTab[(c_polyBase[mc] & 0xFFFF) + k], k = the decider bits of the faces in the
// mask c_polyBase[mc] >> 16, packed in face order.  Cases the table marks unambiguous
// have an empty mask and their table polygon; the others hold the asymptotic
// polygon of every outcome (656 entries in all).  Filled by initGridPassTables.
constexpr int kPolyTabSize = 656;
__constant__ uint64_t c_polyTab[kPolyTabSize];
__constant__ uint32_t c_polyBase[256];
// per table entry: the contour of each crossed edge, 2 bits per edge (contourMap of the word),
// and the number of contours - 1 in bits 24-25
__constant__ uint32_t c_polyCmap[kPolyTabSize];
// per table entry and contour t: the contour's edges as a 12-bit mask, at 4 * entry + t
__constant__ uint16_t c_contourEdges[4 * kPolyTabSize];

// the table entry of a cell of case mc: the decider bits of the case's ambiguous faces (only
// those need the corner values; s is read for them alone)
__device__ __forceinline__ int deciderIndex(uint32_t b, const float s[8], float iso) {
    const uint32_t mask = b >> 16;
    int k = 0, bit = 0;
#define MDMC_FACE(F)                                          \
    if ((mask >> F) & 1u) {                                   \
        if (faceDecider<F>(s, iso)) k |= 1 << bit;            \
        ++bit;                                                \
    }
    MDMC_FACE(0) MDMC_FACE(1) MDMC_FACE(2) MDMC_FACE(3) MDMC_FACE(4) MDMC_FACE(5)
#undef MDMC_FACE
    return (int)(b & 0xFFFFu) + k;
}
__device__ __forceinline__ uint64_t cellPolygonTab(int mc, const float s[8], float iso) {
    return c_polyTab[deciderIndex(c_polyBase[mc], s, iso)];
}

// An active cell's record (classify -> emit), 28 bits: table entry (bits 0-9), cell in the
// tile (10-21), owned x / y / z edge crossed with all four cells present (22-24), edgy: the
// cell's first corner below the iso-value (25), contours - 1 (26-27)
__device__ __forceinline__ int recIndex(uint32_t r) { return (int)(r & 0x3FFu); }
__device__ __forceinline__ int recCell(uint32_t r) { return (int)((r >> 10) & 0xFFFu); }
__device__ __forceinline__ int recOwn(uint32_t r) { return (int)((r >> 22) & 0xFu); }
__device__ __forceinline__ int recVerts(uint32_t r) { return (int)((r >> 26) & 3u) + 1; }
__device__ __forceinline__ int recQuads(uint32_t r) { return __popc((r >> 22) & 7u); }

// n / d and n % d for n < 2^24, d < 2^24 (d and its float reciprocal uniform): the float
// quotient is off by at most one, fixed with the exact remainder
struct SmallDiv {
    unsigned d;
    float rcp;
};
__device__ __forceinline__ SmallDiv smallDiv(unsigned d) { return SmallDiv{d, 1.0f / (float)d}; }
__device__ __forceinline__ unsigned divSmall(const SmallDiv& D, unsigned n, unsigned& rem) {
    int q = (int)((float)n * D.rcp);
    int r = (int)n - q * (int)D.d;
    if (r < 0) {
        --q;
        r += (int)D.d;
    } else if (r >= (int)D.d) {
        ++q;
        r -= (int)D.d;
    }
    rem = (unsigned)r;
    return (unsigned)q;
}

// map entry: bits 0-13 vertex index in the tile, then 2 bits per edge 1,2,4,5,6,7,9,10,11
__device__ __forceinline__ uint32_t mapEntry(int vlocal, uint32_t cmap) {
    const uint32_t nb = ((cmap >> 2) & 0xFu) | (((cmap >> 8) & 0xFFu) << 4) | (((cmap >> 18) & 0x3Fu) << 12);
    return (uint32_t)vlocal | (nb << 14);
}
// slot of edge e in the map entry (edges 0, 3, 8 are never read from a neighbour)
__device__ __forceinline__ int mapContour(uint32_t m, int e) {
    const int k = (int)((0x876054320100ULL >> (4 * e)) & 0xF);
    return (int)((m >> (14 + 2 * k)) & 3u);
}

// The field pointer comes from a descriptor in memory, where the compiler cannot see that it
// addresses global memory; read through the global address space (global, not flat, loads).
typedef const __attribute__((address_space(1))) float GlobalFloat;
__device__ __forceinline__ GlobalFloat* fieldOf(const GridDev& g) { return (GlobalFloat*)g.v; }

__device__ __forceinline__ void loadCorners(const GridDev& g, int x, int y, int z, float s[8]) {
    const int64_t nx = g.nx, nxy = (int64_t)g.nx * g.ny;
    const int64_t b = (int64_t)z * nxy + (int64_t)y * nx + x;
    GlobalFloat* v = fieldOf(g);
    s[0] = v[b];
    s[1] = v[b + 1];
    s[2] = v[b + nx];
    s[3] = v[b + nx + 1];
    s[4] = v[b + nxy];
    s[5] = v[b + nxy + 1];
    s[6] = v[b + nxy + nx];
    s[7] = v[b + nxy + nx + 1];
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
// the same from the case bits (bit q: s[q] >= iso) and the edgy bit, with no corner values
__device__ __forceinline__ int ownedBitsCase(int mc, int edgy, int x, int y, int z) {
    int m = edgy ? 8 : 0;
    if (y >= 1 && z >= 1 && ((mc ^ (mc >> 1)) & 1)) m |= 1;
    if (x >= 1 && z >= 1 && ((mc ^ (mc >> 2)) & 1)) m |= 2;
    if (x >= 1 && y >= 1 && ((mc ^ (mc >> 4)) & 1)) m |= 4;
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

// the last index l in [0, n) with pre[l] <= key (pre non-decreasing, pre[0] = 0 <= key, n <=
// TILE): a fixed number of halving steps with selects, so that the lanes of a wave stay in step
template <int TILE>
__device__ __forceinline__ int upperIndex(const uint16_t* pre, int n, int key) {
    int lo = 0;
#pragma unroll
    for (int step = TILE / 2; step > 0; step >>= 1) {
        const int m = lo + step;
        lo = (m < n && (int)pre[m] <= key) ? m : lo;
    }
    return lo;
}

// a power of two whose reciprocal is a normal float (exponent field 1..253, zero mantissa)
__device__ __forceinline__ bool exactReciprocal(float d) {
    const uint32_t b = __float_as_uint(d);
    const uint32_t e = (b >> 23) & 0xFF;
    return (b & 0x7FFFFF) == 0 && e >= 1 && e <= 253;
}

// sliceP for contour t of one cell, in three parts: the 24 neighbour samples of the corner
// gradients (computeGradients), the edge mean of the contour's crossings (cell-local), and --
// after the move to the surface -- the normal and the world position.
__device__ __forceinline__ void neighbourSamples(const GridDev& g, int x, int y, int z, float nbx[8], float nby[8], float nbz[8]) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int64_t NX = nx, NXY = (int64_t)nx * ny;
    GlobalFloat* fv = fieldOf(g);
    auto at = [&](int px, int py, int pz) { return fv[(int64_t)pz * NXY + (int64_t)py * NX + px]; };
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int cx = k & 1, cy = (k >> 1) & 1, cz = (k >> 2) & 1;
        const int px = x + cx, py = y + cy, pz = z + cz;
        nbx[k] = at(clampi(cx ? px + 1 : px - 1, 0, nx - 1), py, pz);
        nby[k] = at(px, clampi(cy ? py + 1 : py - 1, 0, ny - 1), pz);
        nbz[k] = at(px, py, clampi(cz ? pz + 1 : pz - 1, 0, nz - 1));
    }
}

// contour t of polygon word w as a contour word: its edges from nibble 0, its size in the top nibble
__device__ __forceinline__ uint64_t contourWord(uint64_t w, int t) {
    int j = 0;
    for (int c = 0; c < t; ++c) j += polySize(w, c);
    const int sz = polySize(w, t);
    const uint64_t edges = (w >> (16 + 4 * j)) & ((1ULL << (4 * sz)) - 1ULL);
    return edges | ((uint64_t)sz << 60);
}

// Interior edges of cell (x, y, z) of a grid of CX x CY x CZ cells as a 12-bit mask: an edge is
// interior when all four cells around it are inside the grid, i.e. each coordinate across the
// edge lies in [1, C - 1] at the edge's corner, which is x >= 1 for an edge on the cell's
// low-x face and x <= CX - 2 for one on its high-x face (likewise y, z).  Edge e's faces come
// from its first corner and its axis (EdgeVertices, CellIntersection.inl:383): low x: edges 3,
// 7, 8, 11; high x: 1, 5, 9, 10; low y: 0, 4, 8, 9; high y: 2, 6, 10, 11; low z: 0, 1, 2, 3; high
// z: 4, 5, 6, 7.  A DMC vertex is in one quad per interior edge of its contour.
__device__ __forceinline__ uint32_t interiorEdgeMask(int x, int y, int z, int CX, int CY, int CZ) {
    uint32_t out = 0;
    if (x < 1) out |= 0x988u;
    if (x > CX - 2) out |= 0x622u;
    if (y < 1) out |= 0x311u;
    if (y > CY - 2) out |= 0xC44u;
    if (z < 1) out |= 0x00Fu;
    if (z > CZ - 2) out |= 0x0F0u;
    return ~out & 0xFFFu;
}

// the mean of the contour's edge crossings in cell-local coordinates (CellIntersection.inl:391-406)
__device__ __forceinline__ f3 edgeMean(const float s[8], float iso, uint64_t cw) {
    const int sz = (int)(cw >> 60);
    f3 p = mk3(0.f, 0.f, 0.f);
    for (int i = 0; i < sz; ++i) {
        const int e = (int)((cw >> (4 * i)) & 0xF);
        const int a = edgeA(e), b = edgeB(e), ax = edgeAxis(e);
        const float tt = (iso - s[a]) / (s[b] - s[a]);
        f3 off;
        off.x = 0.f + (tt * (ax == 0 ? 1.f : 0.f) + (float)(a & 1));
        off.y = 0.f + (tt * (ax == 1 ? 1.f : 0.f) + (float)((a >> 1) & 1));
        off.z = 0.f + (tt * (ax == 2 ? 1.f : 0.f) + (float)((a >> 2) & 1));
        p = p + off;
    }
    return divs(p, (float)sz);
}

__device__ __forceinline__ void finishVertex(const GridDev& g, int x, int y, int z, const float s[8], const float nbx[8], const float nby[8],
                                             const float nbz[8], f3 p, f3& wp, f3& nn) {
    // central differences ((upper - lower) * 0.5 / delta) at the 8 corners
    float gxv[8], gyv[8], gzv[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int cx = k & 1, cy = (k >> 1) & 1, cz = (k >> 2) & 1;
        const float gx = cx ? nbx[k] - s[k & ~1] : s[k | 1] - nbx[k];
        const float gy = cy ? nby[k] - s[k & ~2] : s[k | 2] - nby[k];
        const float gz = cz ? nbz[k] - s[k & ~4] : s[k | 4] - nbz[k];
        gxv[k] = 0.5f * gx;
        gyv[k] = 0.5f * gy;
        gzv[k] = 0.5f * gz;
    }
    // (0.5 g) / delta.  When every delta is a power of two its reciprocal is exact and the
    // product equals the quotient bit for bit, so that case skips the division sequence.
    if (exactReciprocal(g.dx) && exactReciprocal(g.dy) && exactReciprocal(g.dz)) {
        const float rx = 1.0f / g.dx, ry = 1.0f / g.dy, rz = 1.0f / g.dz;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            gxv[k] *= rx;
            gyv[k] *= ry;
            gzv[k] *= rz;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            gxv[k] /= g.dx;
            gyv[k] /= g.dy;
            gzv[k] /= g.dz;
        }
    }
    nn = normalize3(mk3(trilinear(gxv, p), trilinear(gyv, p), trilinear(gzv, p)));
    wp = mk3(g.ox, g.oy, g.oz) + (mk3((float)x, (float)y, (float)z) + p) * mk3(g.dx, g.dy, g.dz);
}

// coordinates of cell c0 + l from those of c0 (x0, y0, z0) and l < tile size, in 32-bit arithmetic
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

// Block scan of the per-active-cell counts (vertices | quads << 16, from the records); thread
// tid owns active indices CPT*tid .. CPT*tid + CPT - 1.  Writes the exclusive prefixes and
// returns the tile's totals.  A tile has at most 4*4096 vertices and 3*4096 quads, so both fit
// 16 bits and the packed sum never carries.
template <int CPT>
__device__ __forceinline__ void scanCounts(const uint32_t* s_rec, int NA, uint16_t* s_vpre, uint16_t* s_qpre, int* s_wave, int& VT, int& QT) {
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    int cnt = 0;
    int nvq[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int ai = tid * CPT + j;
        nvq[j] = 0;
        if (ai < NA) {
            const uint32_t r = s_rec[ai];
            nvq[j] = recVerts(r) | (recQuads(r) << 16);
        }
        cnt += nvq[j];
    }
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
    for (int j = 0; j < CPT; ++j) {
        const int ai = tid * CPT + j;
        if (ai < NA) {
            s_vpre[ai] = (uint16_t)(run & 0xFFFF);
            if (s_qpre) s_qpre[ai] = (uint16_t)(run >> 16);
        }
        run += nvq[j];
    }
    VT = tot & 0xFFFF;
    QT = tot >> 16;
}

struct TileCtx {
    int64_t t;
    int prob;
    int64_t c0;   // first cell of the tile inside its problem
    int64_t gc0;  // ... and in the concatenation
    TileBase tb;
};

template <int TILE>
__device__ __forceinline__ TileCtx tileCtxOf(const GridPassArgs& a, int64_t t) {
    TileCtx c;
    c.t = t;
    c.prob = problemOfTile(a.probs, a.nProblems, c.t);
    const ProblemDev& P = a.probs[c.prob];
    c.c0 = (c.t - P.tileBase) * TILE;
    c.gc0 = P.cellBase + c.c0;
    const int64_t r0 = c.c0 / P.CX;
    c.tb.x0 = (int)(c.c0 - r0 * P.CX);
    c.tb.y0 = (int)(r0 % P.CY);
    c.tb.z0 = (int)(r0 / P.CY);
    return c;
}

// The tile a block works on: with QUEUE false one per block (its index mapped as tileOfBlock
// maps it); with QUEUE true a persistent block takes tile after tile from a work queue (one
// counter bumped by the block's first thread, broadcast through LDS) until none is left.
// Returns -1 when the block is done.  The caller's previous tile must be finished (a barrier)
// before it asks for the next one.
template <bool QUEUE>
__device__ __forceinline__ int64_t nextTile(const GridPassArgs& a, int slot, int chunked, int iter, int64_t* s_t) {
    if (!QUEUE) return iter == 0 ? tileOfBlock(blockIdx.x, a.numTiles, chunked != 0) : -1;
    if (threadIdx.x == 0) *s_t = atomicAdd((unsigned long long*)&a.queue[slot], 1ULL);
    __syncthreads();
    const int64_t t = *s_t;
    return t < a.numTiles ? t : -1;
}

// coordinates of tile cell l (the tile's first cell at (x0, y0, z0)), with the small divider
__device__ __forceinline__ void tileCellXYZ(const TileBase& b, int l, const SmallDiv& dx, const SmallDiv& dy, int& x, int& y, int& z) {
    unsigned rx, ry;
    const unsigned r = divSmall(dx, (unsigned)(b.x0 + l), rx);
    const unsigned zz = divSmall(dy, (unsigned)b.y0 + r, ry);
    x = (int)rx;
    y = (int)ry;
    z = b.z0 + (int)zz;
}

// ---- classify: active cells of one tile, their records, counts and map entries
template <int CPT, bool QUEUE>
__global__ void __launch_bounds__(GP_THREADS) kClassify(GridPassArgs a, int chunked) {
    constexpr int TILE = GP_THREADS * CPT, RUNS = TILE / 64;
    __shared__ uint32_t s_rec[TILE];
    __shared__ uint32_t s_base[256];          // c_polyBase
    __shared__ uint32_t s_tab[kPolyTabSize];  // c_polyCmap
    __shared__ uint16_t s_cell[TILE];
    __shared__ uint16_t s_vpre[TILE];
    __shared__ uint16_t s_mc[TILE];  // MC case of each tile cell | edgy << 8
    __shared__ unsigned long long s_mask[RUNS];
    __shared__ int s_mpre[RUNS + 1];
    __shared__ int s_wave[GP_WAVES];

    __shared__ int64_t s_t;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    // the two small tables in LDS (read after the first barrier below)
    s_base[tid] = c_polyBase[tid];
    for (int i = tid; i < kPolyTabSize; i += GP_THREADS) s_tab[i] = c_polyCmap[i];
    for (int iter = 0;; ++iter) {
    const int64_t tile = nextTile<QUEUE>(a, 0, chunked, iter, &s_t);
    if (tile < 0) break;
    const TileCtx tc = tileCtxOf<TILE>(a, tile);
    const ProblemDev& P = a.probs[tc.prob];
    const GridDev g = P.g;
    const float iso = P.iso;
    const int CX = P.CX, CY = P.CY;
    const int64_t C = P.cells;

    // MC case of every cell; a ballot per run of 64 cells marks the active ones.  This thread's
    // cells are GP_THREADS apart: their coordinates advance from the first one's, with no
    // division per cell.  Cells past the grid's end (last tile only) read cell 0 and are
    // dropped, so that every corner load is issued before the first compare.
    const SmallDiv dX = smallDiv((unsigned)CX), dY = smallDiv((unsigned)CY);
    int cx[CPT], cy[CPT], cz[CPT];
    tileCellXYZ(tc.tb, tid, dX, dY, cx[0], cy[0], cz[0]);
#pragma unroll
    for (int k = 1; k < CPT; ++k) {
        cx[k] = cx[k - 1];
        cy[k] = cy[k - 1];
        cz[k] = cz[k - 1];
        if (CX >= GP_THREADS) {  // at most one row boundary per step
            cx[k] += GP_THREADS;
            if (cx[k] >= CX) {
                cx[k] -= CX;
                if (++cy[k] == CY) {
                    cy[k] = 0;
                    ++cz[k];
                }
            }
        } else {
            tileCellXYZ(tc.tb, k * GP_THREADS + tid, dX, dY, cx[k], cy[k], cz[k]);
        }
    }
    float cs[CPT][8];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const bool in = tc.c0 + k * GP_THREADS + tid < C;
        loadCorners(g, in ? cx[k] : 0, in ? cy[k] : 0, in ? cz[k] : 0, cs[k]);
    }
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int l = k * GP_THREADS + tid;
        int mc = 0, edgy = 0;
        if (tc.c0 + l < C) {
#pragma unroll
            for (int q = 0; q < 8; ++q) mc |= (cs[k][q] >= iso ? 1 : 0) << q;
            edgy = cs[k][0] < iso ? 1 : 0;
        }
        const uint64_t m = __ballot(mc != 0 && mc != 255);
        if (lane == 0) s_mask[k * GP_WAVES + wid] = m;
        s_mc[l] = (uint16_t)(mc | (edgy << 8));
    }
    __syncthreads();
    if (tid < 64) {
        const int v = tid < RUNS ? __popcll(s_mask[tid]) : 0;
        int x = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d, 64);
            if (tid >= d) x += y;
        }
        if (tid < RUNS) s_mpre[tid] = x - v;
        if (tid == RUNS - 1) s_mpre[RUNS] = x;
    }
    __syncthreads();
    const int NA = s_mpre[RUNS];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int run = k * GP_WAVES + wid;
        const uint64_t m = s_mask[run];
        if ((m >> lane) & 1ULL) s_cell[s_mpre[run] + __popcll(m & ((1ULL << lane) - 1ULL))] = (uint16_t)(k * GP_THREADS + tid);
    }
    __syncthreads();

    // records of the active cells, every lane busy: the corners are loaded again only for the
    // cases with an ambiguous face (the decider reads them); owned edges come from the case
    for (int ai = tid; ai < NA; ai += GP_THREADS) {
        const int l = s_cell[ai];
        const int mce = s_mc[l], mc = mce & 0xFF;
        int x, y, z;
        tileCellXYZ(tc.tb, l, dX, dY, x, y, z);
        const uint32_t b = s_base[mc];
        int idx = (int)(b & 0xFFFFu);
        if (b >> 16) {
            float s[8];
            loadCorners(g, x, y, z, s);
            idx = deciderIndex(b, s, iso);
        }
        const int own = ownedBitsCase(mc, mce >> 8, x, y, z);
        s_rec[ai] = (uint32_t)idx | ((uint32_t)l << 10) | ((uint32_t)own << 22) | ((s_tab[idx] >> 24) << 26);
    }
    __syncthreads();
    int VT, QT;
    scanCounts<CPT>(s_rec, NA, s_vpre, nullptr, s_wave, VT, QT);
    __syncthreads();

    const int64_t rb = tc.t * TILE;
    for (int ai = tid; ai < NA; ai += GP_THREADS) {
        const uint32_t r = s_rec[ai];
        a.rec[rb + ai] = r;
        a.cellMap[tc.gc0 + recCell(r)] = mapEntry(s_vpre[ai], s_tab[recIndex(r)]);
    }
    if (tid == 0) {
        a.tileV[tc.t] = VT;
        a.tileQ[tc.t] = QT;
        a.tileNA[tc.t] = NA;
        if (tc.t == a.numTiles - 1) {
            a.tileV[a.numTiles] = 0;
            a.tileQ[a.numTiles] = 0;
        }
    }
    __syncthreads();  // the next tile reuses the block's LDS
    }
}

// ---- emit: the tile's dual vertices and the quads its cells own, one active cell per thread
template <int CPT, bool QUEUE>
__global__ void __launch_bounds__(GP_THREADS) kEmit(GridPassArgs a, int chunked) {
    constexpr int TILE = GP_THREADS * CPT;
    __shared__ uint32_t s_rec[TILE];
    __shared__ uint32_t s_tab[kPolyTabSize];  // c_polyCmap: per table entry the contour of each crossed edge
    __shared__ uint16_t s_vpre[TILE];
    __shared__ uint16_t s_qpre[TILE];
    __shared__ uint16_t s_aidx[TILE];  // per tile cell: its active index (0xFFFF: inactive, adjoint map only)
    __shared__ int s_wave[GP_WAVES];

    __shared__ int64_t s_t;
    const int tid = threadIdx.x;
    for (int i = tid; i < kPolyTabSize; i += GP_THREADS) s_tab[i] = c_polyCmap[i];
    for (int iter = 0;; ++iter) {
    const int64_t tile = nextTile<QUEUE>(a, 1, chunked, iter, &s_t);
    if (tile < 0) break;
    const TileCtx tc = tileCtxOf<TILE>(a, tile);
    const ProblemDev& P = a.probs[tc.prob];
    const int CX = P.CX, CY = P.CY;
    const int64_t C = P.cells;
    const int64_t c0 = tc.c0;
    const int NA = a.tileNA[tc.t];
    const int64_t rb = tc.t * TILE;
    if (a.adjMap) {
        for (int l = tid; l < TILE; l += GP_THREADS) s_aidx[l] = 0xFFFF;
        __syncthreads();
    }
    for (int ai = tid; ai < NA; ai += GP_THREADS) {
        const uint32_t r = a.rec[rb + ai];
        s_rec[ai] = r;
        s_aidx[recCell(r)] = (uint16_t)ai;
    }
    __syncthreads();
    int VT, QT;
    scanCounts<CPT>(s_rec, NA, s_vpre, s_qpre, s_wave, VT, QT);
    __syncthreads();
    const int64_t voff = a.tileVOff[tc.t], qoff = a.tileQOff[tc.t];

    // for the adjoint: cell -> (first vertex << 2 | contours - 1), -1 for empty cells
    if (a.adjMap) {
        for (int l = tid; l < TILE; l += GP_THREADS) {
            const int64_t c = c0 + l;
            if (c < C) {
                const int ai = s_aidx[l];
                a.adjMap[c] = ai == 0xFFFF ? -1 : (int)(((voff + s_vpre[ai]) << 2) | (int64_t)(recVerts(s_rec[ai]) - 1));
            }
        }
    }

    const SmallDiv dX = smallDiv((unsigned)CX), dY = smallDiv((unsigned)CY);
    for (int ai = tid; ai < NA; ai += GP_THREADS) {
        const uint32_t r = s_rec[ai];
        const int l = recCell(r);
        const uint32_t cmap = s_tab[recIndex(r)];
        // per vertex: its cell, table entry and contour; kVertices computes position and normal
        const int nv = recVerts(r);
        const int64_t v0 = voff + s_vpre[ai];
        for (int t = 0; t < nv; ++t) {
            a.vcell[v0 + t] = (int)(c0 + l);
            a.vct[v0 + t] = (uint16_t)(recIndex(r) | (t << 10));
            if (a.vpid) a.vpid[v0 + t] = tc.prob;
        }
        // quads whose edge starts at this cell.  Their other cells lie 1, CX, CX+1, CX*CY,
        // CX*CY+1 or CX*CY+CX cells back: in this tile (LDS) or an earlier one (map).
        const int own = recOwn(r);
        if ((own & 7) == 0) continue;
        int x, y, z;
        tileCellXYZ(tc.tb, l, dX, dY, x, y, z);
        const int edgy = (own >> 3) & 1;
        int64_t q = qoff + s_qpre[ai];
        for (int axis = 0; axis < 3; ++axis) {
            if (!((own >> axis) & 1)) continue;
            // own edge, then (dx, dy, dz, edge) of the three other cells around it, packed
            // 4 bits each (dx | dy<<1 | dz<<2, edge) per cell: x: (y-1: e2) (z-1: e4) (y-1,z-1: e6);
            // y: (x-1: e1) (z-1: e7) (x-1,z-1: e5); z: (x-1: e9) (y-1: e11) (x-1,y-1: e10)
            const int ownE = axis == 0 ? 0 : (axis == 1 ? 3 : 8);
            const uint32_t nbTable = axis == 0 ? 0x664422u : (axis == 1 ? 0x557411u : 0xA3B291u);
            int s0 = -1, s1 = -1, s2 = -1, s3 = -1;
            {
                const int sl = edgeSlot(ownE, edgy);
                const int vid = (int)(v0 + (int)((cmap >> (2 * ownE)) & 3u));
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
                    vid = (int)(voff + s_vpre[an] + (int)((s_tab[recIndex(s_rec[an])] >> (2 * e)) & 3u));
                } else {
                    const uint32_t m = a.cellMap[P.cellBase + n];
                    vid = (int)(a.tileVOff[P.tileBase + n / TILE] + (int64_t)(m & 0x3FFFu) + mapContour(m, e));
                }
                const int sl = edgeSlot(e, edgy);
                s0 = sl == 0 ? vid : s0;
                s1 = sl == 1 ? vid : s1;
                s2 = sl == 2 ? vid : s2;
                s3 = sl == 3 ? vid : s3;
            }
            a.quads[q] = make_int4(s0, s1, s2, s3);
            a.qattr[q] = (uint32_t)(3 * ((x & 1) | ((y & 1) << 1) | ((z & 1) << 2)) + axis);
            ++q;
        }
    }
    __syncthreads();  // the next tile reuses the block's LDS
    }
}

// ---- row slabs
// A block owns bandW (<= SLAB_XMAX) cells of R consecutive rows and a chunk of slices, one LDS
// column per thread: column c is the cell x0 - 1 + c (c = 0 is the last cell of the band to
// the left), row r the cell row y0 - 1 + r (r = 0 is the last row of the block above).  The
// block marches z through its chunk with the two point planes of the current slice in LDS.
// Each row segment (slice, row, band) is a contiguous run of the linear cell order, so the
// count pass gives every segment its vertex / quad counts, one scan over the segments gives
// each its first vertex and quad, and the emit pass numbers its cells from those offsets:
// the same ids, in the same (cell, contour) / (edge key) order, as the tile shapes.  The emit
// pass also classifies the halo cells, whose ids it needs for its quads, and keeps the
// previous slice's ids and contour maps on chip for the quads' z - 1 corners.
struct SlabCtx {
    int prob;
    int band, rb, zc;
    int x0, xn, y0, yn, z0, z1;
};

template <int R>
__device__ __forceinline__ SlabCtx slabCtx(const GridPassArgs& a, int chunked) {
    SlabCtx c;
    const int64_t L = tileOfBlock(blockIdx.x, a.numBlocks, chunked != 0);
    int lo = 0, hi = a.nProblems;  // probs[lo].blockBase <= L
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.probs[mid].blockBase <= L) lo = mid;
        else hi = mid;
    }
    const ProblemDev& P = a.probs[lo];
    c.prob = lo;
    const int64_t rel = L - P.blockBase;
    c.band = (int)(rel % P.bands);
    const int64_t rest = rel / P.bands;
    c.rb = (int)(rest % P.rowBlocks);
    c.zc = (int)(rest / P.rowBlocks);
    const int CZ = P.g.nz - 1;
    c.x0 = c.band * P.bandW;
    c.xn = min(P.bandW, P.CX - c.x0);
    c.y0 = c.rb * R;
    c.yn = min(R, P.CY - c.y0);
    c.z0 = c.zc * P.zChunk;
    c.z1 = min(CZ, c.z0 + P.zChunk);
    return c;
}

__device__ __forceinline__ int64_t segIndex(const ProblemDev& P, int z, int y, int band) {
    return P.segBase + ((int64_t)z * P.CY + y) * P.bands + band;
}

// EMIT = false: vertex and quad counts per row segment.  EMIT = true: ids, quads and the
// vertex work items (cell, contour word) of the block's cells.
template <int R, bool EMIT>
__global__ void __launch_bounds__(SLAB_COLS) kSlab(GridPassArgs a, int chunked) {
    constexpr int NR = R + 1;             // cell rows: the halo row and the R own rows
    constexpr int PR = R + 2;             // point rows y0 - 1 .. y0 + R
    constexpr int NC = NR * SLAB_COLS;    // cell slots per slice: row * SLAB_COLS + column
    constexpr int NW = SLAB_COLS / 64;
    constexpr int RING = EMIT ? 2 : 1;
    __shared__ float s_pl[2][PR][SLAB_COLS];  // point planes, plane z in s_pl[z & 1]
    __shared__ uint8_t s_mc[NC];
    __shared__ uint16_t s_act[NC];            // active cell slots, row-major
    __shared__ uint16_t s_poly[NC];           // polygon table entry by active index
    __shared__ uint32_t s_cnt[NC];            // vertices | quads << 8 | owned bits (x, y, z edge, edgy) << 16
    __shared__ int32_t s_id[RING][NC];        // first vertex id of the cell, -1 without one; ring by z parity
    __shared__ uint32_t s_cm[RING][NC];       // contour of each crossed edge, 2 bits per edge
    __shared__ int32_t s_qb[EMIT ? NC : 1];   // first quad id of an own cell
    __shared__ int s_wn[NR][NW];              // per row and wave: active cells, then their base
    __shared__ uint32_t s_wsum[NR][NW];       // per row and wave: vertices | quads << 16
    __shared__ int64_t s_rowV[NR], s_rowQ[NR];
    __shared__ int s_nact;

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const SlabCtx bc = slabCtx<R>(a, chunked);
    const ProblemDev& P = a.probs[bc.prob];
    const GridDev g = P.g;
    const float iso = P.iso;
    const int CX = P.CX, CY = P.CY;
    const int64_t NX = g.nx, NXY = (int64_t)g.nx * g.ny;
    // this thread's column: cell x = x0 - 1 + tid, point x the same
    const int x = bc.x0 - 1 + tid;
    const bool colPoint = tid <= bc.xn + 1 && x >= 0 && x < g.nx;
    const bool colCell = EMIT ? (tid <= bc.xn && x >= 0) : (tid >= 1 && tid <= bc.xn);
    const int zs = EMIT ? max(bc.z0 - 1, 0) : bc.z0;

    auto loadPlane = [&](int zp) {
        float* dst = &s_pl[zp & 1][0][0];
#pragma unroll
        for (int k = 0; k < PR; ++k) {
            const int py = bc.y0 - 1 + k;
            float v = 0.f;
            if (colPoint && py >= 0 && py < g.ny && zp < g.nz) v = g.v[(int64_t)zp * NXY + (int64_t)py * NX + x];
            dst[k * SLAB_COLS + tid] = v;
        }
    };
    if (bc.z0 < bc.z1) loadPlane(zs);

    for (int z = zs; z < bc.z1; ++z) {
        const bool out = EMIT ? (z >= bc.z0) : true;  // the slice before the chunk only feeds the quads
        const int slot = z & 1;
        const int ring = EMIT ? slot : 0;
        loadPlane(z + 1);
        __syncthreads();
        const float* lo = &s_pl[z & 1][0][0];
        const float* hi = &s_pl[(z + 1) & 1][0][0];

        // classify this thread's column (one cell per row) and ballot the active cells
        uint64_t bal[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int y = bc.y0 - 1 + r;
            const bool cell = colCell && y >= 0 && r <= bc.yn && (EMIT || r >= 1);
            int mc = 0;
            if (cell) {
                const int k0 = r * SLAB_COLS + tid, k1 = (r + 1) * SLAB_COLS + tid;
                const float c8[8] = {lo[k0], lo[k0 + 1], lo[k1], lo[k1 + 1], hi[k0], hi[k0 + 1], hi[k1], hi[k1 + 1]};
#pragma unroll
                for (int q = 0; q < 8; ++q) mc |= (c8[q] >= iso ? 1 : 0) << q;
            }
            const bool act = cell && mc != 0 && mc != 255;
            bal[r] = __ballot(act);
            s_mc[r * SLAB_COLS + tid] = (uint8_t)mc;
            s_cnt[r * SLAB_COLS + tid] = 0;
            if (lane == 0) s_wn[r][wid] = __popcll(bal[r]);
        }
        __syncthreads();
        if (tid == 0) {
            int run = 0;
            for (int r = 0; r < NR; ++r)
                for (int w = 0; w < NW; ++w) {
                    const int n = s_wn[r][w];
                    s_wn[r][w] = run;
                    run += n;
                }
            s_nact = run;
        }
        __syncthreads();
        const uint64_t below = (1ULL << lane) - 1ULL;
#pragma unroll
        for (int r = 0; r < NR; ++r)
            if ((bal[r] >> lane) & 1ULL) s_act[s_wn[r][wid] + __popcll(bal[r] & below)] = (uint16_t)(r * SLAB_COLS + tid);
        __syncthreads();
        const int nact = s_nact;

        // polygons of the active cells, every lane busy
        for (int ai = tid; ai < nact; ai += SLAB_COLS) {
            const int c = s_act[ai];
            const int r = c / SLAB_COLS, col = c % SLAB_COLS;
            const int k0 = r * SLAB_COLS + col, k1 = (r + 1) * SLAB_COLS + col;
            const float sv[8] = {lo[k0], lo[k0 + 1], lo[k1], lo[k1 + 1], hi[k0], hi[k0 + 1], hi[k1], hi[k1 + 1]};
            const int idx = deciderIndex(c_polyBase[s_mc[c]], sv, iso);
            const uint64_t w = c_polyTab[idx];
            int own = 0;
            if (r >= 1 && col >= 1) own = ownedBits(sv, iso, bc.x0 - 1 + col, bc.y0 - 1 + r, z);
            s_cnt[c] = (uint32_t)polyCount(w) | ((uint32_t)__popc(own & 7) << 8) | ((uint32_t)own << 16);
            s_poly[ai] = (uint16_t)idx;
            if (EMIT) s_cm[ring][c] = contourMap(w);
        }
        __syncthreads();

        // per row: exclusive prefix over the columns of vertices and quads (wave scan, then waves)
        uint32_t pre[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const uint32_t cnt = s_cnt[r * SLAB_COLS + tid];
            const uint32_t val = (cnt & 0xFFu) | (((cnt >> 8) & 0xFFu) << 16);
            uint32_t xs = val;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t y = __shfl_up(xs, d, 64);
                if (lane >= d) xs += y;
            }
            pre[r] = xs - val;
            if (lane == 63) s_wsum[r][wid] = xs;
        }
        if (EMIT && tid < NR) {
            // the row's first vertex: the row segment's offset, less the halo cell's vertices (the
            // last ones of the band to the left); its first quad: the segment's offset
            const int y = bc.y0 - 1 + tid;
            if (y >= 0 && tid <= bc.yn) {
                const int64_t sg = segIndex(P, z, y, bc.band);
                s_rowV[tid] = a.segVOff[sg] - (int64_t)(s_cnt[tid * SLAB_COLS] & 0xFFu);
                s_rowQ[tid] = a.segQOff[sg];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            uint32_t add = 0, tot = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const uint32_t v = s_wsum[r][w];
                if (w < wid) add += v;
                tot += v;
            }
            pre[r] += add;
            if (!EMIT && tid == 0 && r >= 1 && r <= bc.yn && out) {
                const int64_t sg = segIndex(P, z, bc.y0 - 1 + r, bc.band);
                a.segV[sg] = (int32_t)(tot & 0xFFFFu);
                a.segQ[sg] = (int32_t)(tot >> 16);
            }
        }
        if (EMIT) {
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const int c = r * SLAB_COLS + tid;
                const uint32_t cnt = s_cnt[c];
                const int y = bc.y0 - 1 + r;
                const bool rowOk = y >= 0 && r <= bc.yn;
                s_id[ring][c] = (rowOk && (cnt & 0xFFu)) ? (int32_t)(s_rowV[r] + (int64_t)(pre[r] & 0xFFFFu)) : -1;
                if (rowOk && r >= 1) s_qb[c] = (int32_t)(s_rowQ[r] + (int64_t)(pre[r] >> 16));
                // for the adjoint: cell -> (first vertex << 2 | contours - 1), -1 for empty cells
                if (a.adjMap && out && colCell && tid >= 1 && r >= 1 && r <= bc.yn) {
                    const int64_t cidx = ((int64_t)z * CY + y) * CX + x;
                    a.adjMap[cidx] = (cnt & 0xFFu) ? (int32_t)(((int64_t)s_id[ring][c] << 2) | (int64_t)((cnt & 0xFFu) - 1)) : -1;
                }
            }
        }
        __syncthreads();

        if (EMIT && out) {
            // the vertex work items and the quads of the active own cells
            for (int ai = tid; ai < nact; ai += SLAB_COLS) {
                const int c = s_act[ai];
                const int r = c / SLAB_COLS, col = c % SLAB_COLS;
                if (r < 1 || col < 1) continue;
                const uint32_t cnt = s_cnt[c];
                const int cx = bc.x0 - 1 + col, cy = bc.y0 - 1 + r;
                const int id = s_id[ring][c];
                const int pidx = s_poly[ai];
                const int nv = (int)(cnt & 0xFFu);
                const int32_t cidx = (int32_t)(((int64_t)z * CY + cy) * CX + cx);
                for (int t = 0; t < nv; ++t) {
                    const int64_t v = (int64_t)id + t;
                    a.vcell[v] = cidx;
                    a.vct[v] = (uint16_t)(pidx | (t << 10));
                    if (a.vpid) a.vpid[v] = bc.prob;
                }
                const int own = (int)(cnt >> 16);
                if (!(own & 7)) continue;
                const int edgy = (own >> 3) & 1;
                const uint32_t cm = s_cm[ring][c];
                int q = s_qb[c];
                for (int axis = 0; axis < 3; ++axis) {
                    if (!((own >> axis) & 1)) continue;
                    // own edge, then (dx, dy, dz, edge) of the three other cells around it, 4 bits
                    // each: x: (y-1: e2) (z-1: e4) (y-1,z-1: e6); y: (x-1: e1) (z-1: e7) (x-1,z-1: e5);
                    // z: (x-1: e9) (y-1: e11) (x-1,y-1: e10)
                    const int ownE = axis == 0 ? 0 : (axis == 1 ? 3 : 8);
                    const uint32_t nbTable = axis == 0 ? 0x664422u : (axis == 1 ? 0x557411u : 0xA3B291u);
                    int s0 = -1, s1 = -1, s2 = -1, s3 = -1;
                    {
                        const int sl = edgeSlot(ownE, edgy);
                        const int vid = id + (int)((cm >> (2 * ownE)) & 3u);
                        s0 = sl == 0 ? vid : s0;
                        s1 = sl == 1 ? vid : s1;
                        s2 = sl == 2 ? vid : s2;
                        s3 = sl == 3 ? vid : s3;
                    }
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const uint32_t ent = (nbTable >> (8 * k)) & 0xFFu;
                        const int d = (int)(ent & 0xF), e = (int)(ent >> 4);
                        const int nc = c - (d & 1) - ((d >> 1) & 1) * SLAB_COLS;
                        const int nring = EMIT ? (((d >> 2) & 1) ? ring ^ 1 : ring) : 0;
                        const int vid = s_id[nring][nc] + (int)((s_cm[nring][nc] >> (2 * e)) & 3u);
                        const int sl = edgeSlot(e, edgy);
                        s0 = sl == 0 ? vid : s0;
                        s1 = sl == 1 ? vid : s1;
                        s2 = sl == 2 ? vid : s2;
                        s3 = sl == 3 ? vid : s3;
                    }
                    a.quads[q] = make_int4(s0, s1, s2, s3);
                    a.qattr[q] = (uint32_t)(3 * ((cx & 1) | ((cy & 1) << 1) | ((z & 1) << 2)) + axis);
                    ++q;
                }
            }
        }
        __syncthreads();
    }
    if (!EMIT && tid == 0 && blockIdx.x == 0) {
        a.segV[a.numSegs] = 0;  // the entry past the last segment: the scans' totals
        a.segQ[a.numSegs] = 0;
    }
}

// ---- dual vertices (sliceP), one per thread over all vertices of the pass
// (blocks XCD-chunked like the tiles when `chunked`: consecutive vertices share rows and
// slices of the field, which then stay in one XCD's L2).  The move to the surface takes
// 1.5 steps on average but a wave waits for its slowest lane (5.8 steps at the 512^3
// gyroid), so every lane runs at most kMoveSteps steps; the block's unfinished lanes leave
// their state in LDS, consecutive lanes finish them on full waves, and each vertex then
// computes its normal from its final point.  Each vertex's arithmetic is unchanged.
constexpr int kMoveSteps = 2;
constexpr int kMoveState = 17;  // s[8], iso, i1, v1, point, step

__global__ void __launch_bounds__(GP_THREADS) kVertices(GridPassArgs a, int chunked) {
    __shared__ float s_st[kMoveState][GP_THREADS];
    __shared__ int s_wave[GP_WAVES];
    const int64_t nb = (a.V + GP_THREADS - 1) / GP_THREADS;
    const int64_t v = tileOfBlock(blockIdx.x, nb, chunked != 0) * GP_THREADS + threadIdx.x;
    const bool live = v < a.V;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    GridDev g{};
    float iso = 0.f;
    int x = 0, y = 0, z = 0;
    float s[8], nbx[8], nby[8], nbz[8];
    Mover m{};
    m.done = true;
    f3 p = mk3(0.f, 0.f, 0.f);
    if (live) {
        const ProblemDev& P = a.probs[a.vpid ? a.vpid[v] : 0];
        g = P.g;
        iso = P.iso;
        const unsigned c = (unsigned)a.vcell[v];
        const unsigned r = c / (unsigned)P.CX;
        x = (int)(c - r * (unsigned)P.CX);
        y = (int)(r % (unsigned)P.CY);
        z = (int)(r / (unsigned)P.CY);
        loadCorners(g, x, y, z, s);
        const int ct = a.vct[v];
        const uint64_t cw = contourWord(c_polyTab[ct & 0x3FF], ct >> 10);
        p = edgeMean(s, iso, cw);
        // out-degree of the vertex in the halfedge mesh: one quad per interior edge of its contour
        if (a.outdeg) a.outdeg[v] = __popc(interiorEdgeMask(x, y, z, P.CX, P.CY, g.nz - 1) & c_contourEdges[4 * (ct & 0x3FF) + (ct >> 10)]);
        if (a.surfIter) {
            moverInit(m, s, p);
            for (int k = 0; k < kMoveSteps && !m.done; ++k) moverStep(m, s, iso);
        }
    }
    // unfinished lanes -> LDS slots, in lane order
    const bool later = live && !m.done;
    const uint64_t bal = __ballot(later);
    if (lane == 0) s_wave[wid] = __popcll(bal);
    __syncthreads();
    int base = 0, total = 0;
#pragma unroll
    for (int w = 0; w < GP_WAVES; ++w) {
        if (w < wid) base += s_wave[w];
        total += s_wave[w];
    }
    const int slot = base + __popcll(bal & ((1ULL << lane) - 1ULL));
    if (later) {
#pragma unroll
        for (int k = 0; k < 8; ++k) s_st[k][slot] = s[k];
        s_st[8][slot] = iso;
        s_st[9][slot] = m.i1.x;
        s_st[10][slot] = m.i1.y;
        s_st[11][slot] = m.i1.z;
        s_st[12][slot] = m.v1;
        s_st[13][slot] = m.point.x;
        s_st[14][slot] = m.point.y;
        s_st[15][slot] = m.point.z;
        s_st[16][slot] = __int_as_float(m.it);
    }
    __syncthreads();
    if ((int)threadIdx.x < total) {
        const int k0 = threadIdx.x;
        float ss[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) ss[k] = s_st[k][k0];
        const float is = s_st[8][k0];
        Mover mm;
        mm.i1 = mk3(s_st[9][k0], s_st[10][k0], s_st[11][k0]);
        mm.v1 = s_st[12][k0];
        mm.point = mk3(s_st[13][k0], s_st[14][k0], s_st[15][k0]);
        mm.it = __float_as_int(s_st[16][k0]);
        mm.done = false;
        while (!mm.done) moverStep(mm, ss, is);
        s_st[13][k0] = mm.point.x;
        s_st[14][k0] = mm.point.y;
        s_st[15][k0] = mm.point.z;
    }
    __syncthreads();
    if (!live) return;
    neighbourSamples(g, x, y, z, nbx, nby, nbz);
    if (a.surfIter) p = later ? moverResult(Mover{mk3(0.f, 0.f, 0.f), mk3(s_st[13][slot], s_st[14][slot], s_st[15][slot]), 0.f, 0, true}) : moverResult(m);
    f3 wp, nn;
    finishVertex(g, x, y, z, s, nbx, nby, nbz, p, wp, nn);
    a.pos[3 * v + 0] = wp.x;
    a.pos[3 * v + 1] = wp.y;
    a.pos[3 * v + 2] = wp.z;
    a.nrm[3 * v + 0] = nn.x;
    a.nrm[3 * v + 1] = nn.y;
    a.nrm[3 * v + 2] = nn.z;
}

// The initialize's outgoing lists from the grid pass's quads: halfedge 4q + i (from corner i
// to corner i + 1) goes to entry off[v] + rank of its origin v = corner i: out holds 4q + i and
// ends its destination (corner i + 1) and the origin of the halfedge before it (corner i - 1).  The rank is the number of interior edges of v's contour with
// a lower edge number than the quad's edge (any order that numbers a vertex's halfedges
// 0 .. out-degree - 1 will do: a list's order never matters, buckets compare ids).  The
// quad's owner cell (the one whose corner 0 starts the edge) is the largest of its corners'
// cells, and the other cells lie 1, CX or CX*CY cells back in two of the axes (kEmit's
// neighbour table), which gives each corner's local edge.  A quad with a vertex id outside
// [0, V), or whose edge is not an interior edge of a corner's contour or ranks past the
// corner's list, is counted as invalid and writes no entry for it (the caller then stops).  Also the face -> halfedge array and the
// face attributes (collectHalfedgeElements, HalfedgeMesh.cu:14-58).
__global__ void kQuadLists(const int4* __restrict__ quads, const uint32_t* __restrict__ qattr, int64_t Q, int64_t V,
                           const int32_t* __restrict__ vcell, const uint16_t* __restrict__ vct, const int32_t* __restrict__ vpid,
                           const ProblemDev* __restrict__ probs, const int64_t* __restrict__ off, int32_t* __restrict__ out,
                           int2* __restrict__ ends, int32_t* __restrict__ face_he, uint32_t* __restrict__ fattr, int64_t* invalid) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Q) return;
    const int4 v4 = quads[q];
    const int v[4] = {v4.x, v4.y, v4.z, v4.w};
    const uint32_t at = qattr[q];
    face_he[q] = (int32_t)(4 * q);
    fattr[q] = at;
    bool bad = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) bad = bad || v[i] < 0 || v[i] >= V;
    if (bad) {
        atomicAdd((unsigned long long*)invalid, 1ULL);
        return;
    }
    const ProblemDev& P = probs[vpid ? vpid[v[0]] : 0];
    const int CX = P.CX, CY = P.CY, CZ = P.g.nz - 1;
    const int64_t CXY = (int64_t)CX * CY;
    int c[4], t[4];
    int64_t o[4], o1[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        c[i] = vcell[v[i]];
        t[i] = vct[v[i]];
        o[i] = off[v[i]];
        o1[i] = off[v[i] + 1];
    }
    const int owner = max(max(c[0], c[1]), max(c[2], c[3]));
    const unsigned r = (unsigned)owner / (unsigned)CX;
    const int ox = (int)((unsigned)owner - r * (unsigned)CX), oy = (int)(r % (unsigned)CY), oz = (int)(r / (unsigned)CY);
    const int axis = (int)((at & QA_COLOR) % 3u);
    int nbad = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int64_t d = (int64_t)owner - c[i];
        const int dz = d >= CXY ? 1 : 0;
        d -= dz * CXY;
        const int dy = d >= CX ? 1 : 0;
        d -= dy * CX;
        const int dx = (int)d;
        // local edge of the corner cell: x axis (dy, dz) 0 2 4 6; y axis (dx, dz) 3 1 7 5; z axis (dx, dy) 8 9 11 10
        const int k = axis == 0 ? (dy | (dz << 1)) : (axis == 1 ? (dx | (dz << 1)) : (dx | (dy << 1)));
        const int e = (int)((axis == 0 ? 0x6420u : (axis == 1 ? 0x5713u : 0xAB98u)) >> (4 * k)) & 0xF;
        const uint32_t m = interiorEdgeMask(ox - dx, oy - dy, oz - dz, CX, CY, CZ) & c_contourEdges[4 * (t[i] & 0x3FF) + (t[i] >> 10)];
        const int64_t pos = o[i] + __popc(m & ((1u << e) - 1u));
        if (!((m >> e) & 1u) || pos >= o1[i] || pos < 0 || pos >= 4 * Q) {
            ++nbad;
            continue;
        }
        out[pos] = (int)(4 * q + i);
        ends[pos] = make_int2(v[(i + 1) & 3], v[(i + 3) & 3]);
    }
    if (nbad) atomicAdd((unsigned long long*)invalid, (unsigned long long)nbad);
}

}  // namespace

int gridPassTile(int shape) { return shape == GP_SHAPE_QUEUE ? 2048 : 1024 << ((shape - 1) % 3); }

void initGridPassTables() {
    static const uint64_t poly[256] = {MDMC_POLYGON_TABLE_DATA};
    static const uint64_t amb[4] = {MDMC_AMBIGUOUS_MASK_DATA};
    uint64_t tab[kPolyTabSize];
    uint32_t base[256];
    int n = 0;
    for (int mc = 0; mc < 256; ++mc) {
        uint32_t mask = 0;
        if ((amb[mc >> 6] >> (mc & 63)) & 1ULL)
            for (int f = 0; f < 6; ++f) {
                const int fc = faceCase(mc, f);
                if (fc == 6 || fc == 9) mask |= 1u << f;
            }
        base[mc] = (uint32_t)n | (mask << 16);
        const int bits = __builtin_popcount(mask);
        if (!((amb[mc >> 6] >> (mc & 63)) & 1ULL)) {
            tab[n++] = poly[mc];
            continue;
        }
        for (int k = 0; k < (1 << bits); ++k) {
            // spread the packed outcome bits k over the faces of the mask
            uint32_t dec = 0;
            int bit = 0;
            for (int f = 0; f < 6; ++f)
                if ((mask >> f) & 1u) {
                    if ((k >> bit) & 1) dec |= 1u << f;
                    ++bit;
                }
            tab[n++] = polygonFromFaceDecisions(mc, dec);
        }
    }
    if (n != kPolyTabSize) throw HipError("grid pass: polygon table size mismatch", -4);
    // per entry: the contour (0..3) of each of its edges, 2 bits per edge
    uint32_t cmap[kPolyTabSize];
    for (int i = 0; i < n; ++i) {
        uint32_t m = 0;
        int j = 0;
        for (int c = 0; c < 4; ++c) {
            const int sz = (int)((tab[i] >> (4 * c)) & 0xF);
            for (int k = 0; k < sz; ++k, ++j) m |= (uint32_t)c << (2 * (int)((tab[i] >> (16 + 4 * j)) & 0xF));
            if (sz > 0) m = (m & 0xFFFFFFu) | ((uint32_t)c << 24);
        }
        cmap[i] = m;
    }
    MDMC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_polyTab), tab, sizeof tab));
    MDMC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_polyBase), base, sizeof base));
    MDMC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_polyCmap), cmap, sizeof cmap));
    // per entry and contour: the contour's edges as a mask
    uint16_t cedges[4 * kPolyTabSize] = {};
    for (int i = 0; i < n; ++i) {
        int j = 0;
        for (int c = 0; c < 4; ++c) {
            const int sz = (int)((tab[i] >> (4 * c)) & 0xF);
            for (int k = 0; k < sz; ++k, ++j) cedges[4 * i + c] |= (uint16_t)(1u << (int)((tab[i] >> (16 + 4 * j)) & 0xF));
        }
    }
    MDMC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_contourEdges), cedges, sizeof cedges));
}

const char* gridPassShapeName(int shape) {
    static const char* names[] = {"auto", "1024 cells, XCD-chunked", "2048 cells, XCD-chunked", "4096 cells, XCD-chunked",
                                  "1024 cells, launch order", "2048 cells, launch order", "4096 cells, launch order",
                                  "row slabs of 4 rows, XCD-chunked", "row slabs of 2 rows, XCD-chunked", "2048 cells, work queue"};
    return (shape >= 0 && shape <= GP_SHAPE_LAST) ? names[shape] : "?";
}

// zero the two work-queue counters (classify, emit) of a queue-shape pass
__global__ void kResetQueue(unsigned long long* q) {
    if (threadIdx.x < 2) q[threadIdx.x] = 0;
}

// persistent blocks for a queue-shape kernel: as many as the chip holds at once, at most one
// per tile
template <class K>
static unsigned queueBlocks(K kernel, int64_t numTiles) {
    int dev = 0, cus = 0, perCU = 0;
    MDMC_HIP(hipGetDevice(&dev));
    MDMC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    MDMC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kernel, GP_THREADS, 0));
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(numTiles, (int64_t)std::max(perCU, 1) * std::max(cus, 1)));
}

void launchGridClassify(int shape, const GridPassArgs& a, hipStream_t s) {
    if (a.numTiles <= 0) return;
    const int chunked = shape <= 3 ? 1 : 0;
    const unsigned nb = (unsigned)a.numTiles;
    if (gridPassIsQueue(shape)) {
        kResetQueue<<<1, 64, 0, s>>>(reinterpret_cast<unsigned long long*>(a.queue));
        kClassify<8, true><<<queueBlocks(kClassify<8, true>, a.numTiles), GP_THREADS, 0, s>>>(a, 0);
        return;
    }
    switch (gridPassTile(shape)) {
        case 1024: kClassify<4, false><<<nb, GP_THREADS, 0, s>>>(a, chunked); break;
        case 2048: kClassify<8, false><<<nb, GP_THREADS, 0, s>>>(a, chunked); break;
        default: kClassify<16, false><<<nb, GP_THREADS, 0, s>>>(a, chunked); break;
    }
}

void launchSlabCount(int shape, const GridPassArgs& a, hipStream_t s) {
    if (a.numBlocks <= 0) return;
    if (gridPassSlabRows(shape) == 4) kSlab<4, false><<<(unsigned)a.numBlocks, SLAB_COLS, 0, s>>>(a, 1);
    else kSlab<2, false><<<(unsigned)a.numBlocks, SLAB_COLS, 0, s>>>(a, 1);
}

void launchSlabEmit(int shape, const GridPassArgs& a, hipStream_t s) {
    if (a.numBlocks <= 0) return;
    if (gridPassSlabRows(shape) == 4) kSlab<4, true><<<(unsigned)a.numBlocks, SLAB_COLS, 0, s>>>(a, 1);
    else kSlab<2, true><<<(unsigned)a.numBlocks, SLAB_COLS, 0, s>>>(a, 1);
    if (a.V > 0) kVertices<<<(unsigned)((a.V + GP_THREADS - 1) / GP_THREADS), GP_THREADS, 0, s>>>(a, 1);
}

void launchQuadLists(const int4* quads, const uint32_t* qattr, int64_t Q, int64_t V, const int32_t* vcell, const uint16_t* vct, const int32_t* vpid,
                     const ProblemDev* probs, const int64_t* off, int32_t* out, int2* ends, int32_t* face_he, uint32_t* fattr, int64_t* invalid,
                     hipStream_t s) {
    if (Q > 0) kQuadLists<<<(unsigned)((Q + 255) / 256), 256, 0, s>>>(quads, qattr, Q, V, vcell, vct, vpid, probs, off, out, ends, face_he, fattr, invalid);
}

void launchGridEmit(int shape, const GridPassArgs& a, hipStream_t s) {
    if (a.numTiles <= 0) return;
    const int chunked = shape <= 3 ? 1 : 0;
    const unsigned nb = (unsigned)a.numTiles;
    if (gridPassIsQueue(shape)) {
        kEmit<8, true><<<queueBlocks(kEmit<8, true>, a.numTiles), GP_THREADS, 0, s>>>(a, 0);
    } else {
        switch (gridPassTile(shape)) {
            case 1024: kEmit<4, false><<<nb, GP_THREADS, 0, s>>>(a, chunked); break;
            case 2048: kEmit<8, false><<<nb, GP_THREADS, 0, s>>>(a, chunked); break;
            default: kEmit<16, false><<<nb, GP_THREADS, 0, s>>>(a, chunked); break;
        }
    }
    if (a.V > 0) kVertices<<<(unsigned)((a.V + GP_THREADS - 1) / GP_THREADS), GP_THREADS, 0, s>>>(a, chunked);
}

}  // namespace mdmc
