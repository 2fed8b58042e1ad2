// mdmc_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the Manifold Dual Marching Cubes hot path of
// PhiliGuertler/ManifoldDMC (MarchingCubes/VisualDMC/src/DMC).  It is the
// checker for the HIP library: only tests/, the smoke test and bench.py's CPU
// baseline may load it.  The product never links or calls it.
//
// The reference runs every stage as CUDA kernels whose outputs are appended
// with atomics, so element ids and a few race-resolved choices depend on the
// thread schedule.  This restatement fixes one legal schedule and documents it
// at each stage:
//   * DMC vertices are appended in (cell linear index, contour index) order,
//     i.e. performDualMarchingCubes' threads run in ascending order.
//   * DMC quads are appended in ascending edge-key order (a legal order of the
//     atomic appends in mapQuadrilaterals).
//   * every other kernel runs its threads one after another in ascending
//     thread id ("sequential schedule"), except markElementsP3333, whose
//     per-face flag clear is taken to happen before any thread marks a
//     neighbour (see simplify3333).
//   * all compactions are stable.
// Arithmetic follows the reference's operation order with no contraction
// (build with -ffp-contract=off); glm semantics are restated where they matter
// (normalize = v * (1/sqrt(dot)), min/max/clamp as (y<x)?y:x, mix = x(1-a)+ya).
// The same code instantiated with double gives the FP64 oracle.
//
// Parity status: every public stage is pinned by the reference's Data-and-Results
// fixtures, closed forms or invariants in tests/.  The P3X3Y/P3333
// simplifications, which no reference fixture covers, are pinned by hand-built
// quad meshes whose simplified results were derived by hand
// (tests/test_oracle_simplify_pins.py, through orc_simplify_quads); each of those
// tests also runs against a planted mistake (orc_set_plant, below) and must fail
// there.  The plants exist only in this CPU checker and are off by default.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../manifolddmc_amd/csrc/mc_tables.inc"

namespace orc {

static const uint64_t kPolygon[256] = {MDMC_POLYGON_TABLE_DATA};
static const uint64_t kAmbiguous[4] = {MDMC_AMBIGUOUS_MASK_DATA};
static inline bool caseIsAmbiguous(int c) { return (kAmbiguous[c >> 6] >> (c & 63)) & 1ULL; }

// --------------------------------------------------------------------------
// option flags (same bit meaning as the product header)
enum : int {
    F_PAD1 = 1,
    F_SURF_ITER = 2,
    F_P3X3Y_COLOR = 4,
    F_P3X3Y_OLD = 8,
    F_P3333 = 16,
    F_SPLIT_NM = 32,
    F_DETECT_NM_VERTS = 64,
};
// SplittingStep, ManifoldSplitter.h:155-166
enum : int { S_NONE = 0, S_SUBDIVIDE = 1, S_COLLAPSE = 2, S_MERGE = 3, S_TWINS = 4, S_MOVE = 5, S_ITERATIVE = 6, S_ALL = 100 };

// Flags bits, Halfedge.h:14-111
enum : uint8_t { VF_NM = 1, VF_BOUNDARY = 2, VF_WEIRD = 4, VF_FANS = 8, VF_REMOVE = 16 };
// QuadrilateralAttribute bits, DeviceVectors.h:334-450
enum : uint16_t { QA_COLOR = 0x1F, QA_NM = 0x20, QA_PATTERN = 0x40, QA_REMOVE = 0x80, QA_BOUNDARY = 0x100 };
static inline uint16_t qaSetColor(uint16_t d, int color) { return (uint16_t)((d & 0xE0) | (uint16_t)color); }

// Planted mistakes for the tests that pin the simplifications: each value turns one
// plausible misreading of the reference on, so a test can show that it catches it.
enum : int {
    PLANT_NONE = 0,
    PLANT_SWAP_SURVIVOR = 1,   // P3X3Y keeps quad.z at the midpoint and maps quad.x onto it
    PLANT_COLOR_LT = 2,        // colour rule with '<' instead of '<=' (HalfedgeMesh.cu:425-428)
    PLANT_NO_RENORM = 3,       // merged normal not renormalised
    PLANT_FAR_CORNER = 4,      // P3333 far vertex taken one corner later (twin.next^3 instead of twin.next^2)
    PLANT_NO_ALL_FOUR = 5,     // P3333 without the all-valence-4 rejection
    PLANT_NO_NEIGHBOR_NM = 6,  // P3X3Y colour / P3333 without the neighbour NM test
    PLANT_OLD_NO_COUNT = 7,    // P3X3Y old without the elementCount == 1 test
    PLANT_MID_XZ = 8,          // P3X3Y midpoint of the two valence-3 corners instead of the two >=5 corners
};
static int g_plant = PLANT_NONE;

// --------------------------------------------------------------------------
template <class R>
struct V3 {
    R x, y, z;
};
template <class R> static inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> static inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> static inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> static inline V3<R> operator/(V3<R> a, V3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class R> static inline V3<R> scl(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> static inline V3<R> divs(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> static inline V3<R> neg(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> static inline R gmax(R a, R b) { return (a < b) ? b : a; }
template <class R> static inline R gmin(R a, R b) { return (b < a) ? b : a; }
template <class R> static inline R gclamp(R x, R lo, R hi) { return gmin(gmax(x, lo), hi); }
template <class R> static inline V3<R> vclamp(V3<R> v, R lo, R hi) { return {gclamp(v.x, lo, hi), gclamp(v.y, lo, hi), gclamp(v.z, lo, hi)}; }
// glm normalize: func_geometric.inl:82-90, compute_dot (x*x+y*y)+z*z, inversesqrt = 1/sqrt
template <class R> static inline V3<R> normalize(V3<R> v) {
    R d = v.x * v.x + v.y * v.y + v.z * v.z;
    R inv = R(1) / std::sqrt(d);
    return {v.x * inv, v.y * inv, v.z * inv};
}
// glm mix: x*(1-a) + y*a
template <class R> static inline R mixs(R a, R b, R t) { return a * (R(1) - t) + b * t; }
template <class R> static inline V3<R> mixv(V3<R> a, V3<R> b, R t) { return {mixs(a.x, b.x, t), mixs(a.y, b.y, t), mixs(a.z, b.z, t)}; }

// Interpolation<T>::interpolateTrilinearly, Interpolation.inl:48-56
template <class R> static inline R trilinear(const R s[8], V3<R> u) {
    return (R(1) - u.z) * ((R(1) - u.y) * (s[0] + u.x * (s[1] - s[0])) + u.y * (s[2] + u.x * (s[3] - s[2]))) +
           u.z * ((R(1) - u.y) * (s[4] + u.x * (s[5] - s[4])) + u.y * (s[6] + u.x * (s[7] - s[6])));
}
template <class R> static inline V3<R> trilinearV(const V3<R> s[8], V3<R> u) {
    R a[8], b[8], c[8];
    for (int i = 0; i < 8; ++i) { a[i] = s[i].x; b[i] = s[i].y; c[i] = s[i].z; }
    return {trilinear(a, u), trilinear(b, u), trilinear(c, u)};
}

// --------------------------------------------------------------------------
// grid
struct Grid {
    int nx = 0, ny = 0, nz = 0;
    float ox = 0, oy = 0, oz = 0;
    float dx = 1, dy = 1, dz = 1;
    std::vector<float> v;
    inline int64_t gid(int x, int y, int z) const { return (int64_t)z * nx * ny + (int64_t)y * nx + x; }
    inline float at(int x, int y, int z) const { return v[gid(x, y, z)]; }
    inline float atc(int x, int y, int z) const {  // clamped corner fetch
        x = x < 0 ? 0 : (x >= nx ? nx - 1 : x);
        y = y < 0 ? 0 : (y >= ny ? ny - 1 : y);
        z = z < 0 ? 0 : (z >= nz ? nz - 1 : z);
        return at(x, y, z);
    }
};

// UniformGridHost::readFromFile(file, refinement, addZeroBoundary) + refineGrid,
// UniformGrid.inl:573-608 and :279-319.  Padding clamps the index, so the
// border is duplicated, not zeroed.
static Grid prepareGrid(const float* data, const int dims[3], const float origin[3], const float delta[3], int refinement, bool pad) {
    Grid in;
    in.nx = dims[0]; in.ny = dims[1]; in.nz = dims[2];
    in.v.assign(data, data + (size_t)in.nx * in.ny * in.nz);
    Grid out;
    out.nx = dims[0] + (dims[0] - 1) * refinement + (pad ? 2 : 0);
    out.ny = dims[1] + (dims[1] - 1) * refinement + (pad ? 2 : 0);
    out.nz = dims[2] + (dims[2] - 1) * refinement + (pad ? 2 : 0);
    const float f = 1.f / (static_cast<float>(refinement) + 1.f);
    out.dx = delta[0] * f; out.dy = delta[1] * f; out.dz = delta[2] * f;
    out.ox = origin[0]; out.oy = origin[1]; out.oz = origin[2];
    if (pad) { out.ox -= out.dx; out.oy -= out.dy; out.oz -= out.dz; }
    out.v.resize((size_t)out.nx * out.ny * out.nz);
    const float rr = static_cast<float>(refinement + 1);
    for (int z = 0; z < out.nz; ++z)
        for (int y = 0; y < out.ny; ++y)
            for (int x = 0; x < out.nx; ++x) {
                int ix = x, iy = y, iz = z;
                if (pad) {
                    ix = gclamp(ix, 1, out.nx - 2) - 1;
                    iy = gclamp(iy, 1, out.ny - 2) - 1;
                    iz = gclamp(iz, 1, out.nz - 2) - 1;
                }
                const float fx = static_cast<float>(ix) / rr, fy = static_cast<float>(iy) / rr, fz = static_cast<float>(iz) / rr;
                const int mx = (int)std::floor(fx), my = (int)std::floor(fy), mz = (int)std::floor(fz);
                float s[8];
                for (int k = 0; k < 8; ++k) s[k] = in.atc(mx + (k & 1), my + ((k >> 1) & 1), mz + ((k >> 2) & 1));
                V3<float> d = {fx - (float)mx, fy - (float)my, fz - (float)mz};
                out.v[out.gid(x, y, z)] = trilinear(s, d);
            }
    return out;
}

// --------------------------------------------------------------------------
// per-edge constants of the unit cell (CellIntersection.inl:165-200, :362-366, :383)
// corner k sits at (k&1, (k>>1)&1, (k>>2)&1)
static const int kEdgeCorner[12][2] = {{0, 1}, {1, 3}, {2, 3}, {0, 2}, {4, 5}, {5, 7}, {6, 7}, {4, 6}, {0, 4}, {1, 5}, {3, 7}, {2, 6}};
static const int kEdgeAxis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
// quad slot of the cell's vertex for the edge, by the "edgy" bit (value at the edge's first corner < iso)
static const int kEdgeSlot[12][2] = {{0, 0}, {0, 0}, {3, 1}, {3, 1}, {1, 3}, {1, 3}, {2, 2}, {2, 2}, {0, 0}, {3, 1}, {2, 2}, {1, 3}};
// cube faces as used by marchingCubesAsymptoticPolygon (CellIntersection.inl:290-292): corners and edges in walk order
static const int kFaceCorner[6][4] = {{0, 2, 1, 3}, {5, 7, 4, 6}, {4, 0, 5, 1}, {2, 6, 3, 7}, {4, 6, 0, 2}, {1, 3, 5, 7}};
static const int kFaceEdge[6][4] = {{3, 2, 1, 0}, {5, 6, 7, 4}, {8, 0, 9, 4}, {11, 6, 10, 2}, {7, 11, 3, 8}, {1, 10, 5, 9}};

struct Polygon {
    int n = 0;        // number of contours
    int size[4] = {0, 0, 0, 0};
    int edge[12];     // edges of all contours back to back
    int start(int c) const { int s = 0; for (int i = 0; i < c; ++i) s += size[i]; return s; }
};

static Polygon polygonFromTable(int mcCase) {
    Polygon p;
    uint64_t w = kPolygon[mcCase];
    int total = 0;
    for (int c = 0; c < 4; ++c) {
        int s = (int)((w >> (4 * c)) & 0xF);
        if (s == 0) break;
        p.size[c] = s; p.n = c + 1; total += s;
    }
    for (int j = 0; j < total; ++j) p.edge[j] = (int)((w >> (16 + 4 * j)) & 0xF);
    return p;
}

// CellIntersection::marchingCubesAsymptoticPolygon + handleAsymptoticFaceCase, CellIntersection.inl:203-338
template <class R>
static Polygon polygonAsymptotic(const R s[8], R iso) {
    uint8_t seg[12];
    for (int i = 0; i < 12; ++i) seg[i] = 0xFF;
    auto connect = [&](int a, int b) {  // a -> b
        seg[a] = (uint8_t)((seg[a] & 0xF0) | (b & 0xF));
        seg[b] = (uint8_t)((seg[b] & 0x0F) | (a << 4));
    };
    for (int f = 0; f < 6; ++f) {
        int fc = 0;
        R val[4];
        const int* e = kFaceEdge[f];
        for (int i = 0; i < 4; ++i) {
            val[i] = s[kFaceCorner[f][i]];
            fc |= (val[i] >= iso ? 1 : 0) << i;
        }
        switch (fc) {
            case 1: connect(e[0], e[3]); break;
            case 2: connect(e[1], e[0]); break;
            case 3: connect(e[1], e[3]); break;
            case 4: connect(e[3], e[2]); break;
            case 5: connect(e[0], e[2]); break;
            case 7: connect(e[1], e[2]); break;
            case 8: connect(e[2], e[1]); break;
            case 10: connect(e[2], e[0]); break;
            case 11: connect(e[2], e[3]); break;
            case 12: connect(e[3], e[1]); break;
            case 13: connect(e[0], e[1]); break;
            case 14: connect(e[3], e[0]); break;
            case 6: {
                R d = (val[0] * val[3] - val[1] * val[2]) / (val[0] + val[3] - val[1] - val[2]);
                if (d >= iso) { connect(e[3], e[0]); connect(e[1], e[2]); }
                else { connect(e[1], e[0]); connect(e[3], e[2]); }
            } break;
            case 9: {
                R d = (val[0] * val[3] - val[1] * val[2]) / (val[0] + val[3] - val[1] - val[2]);
                if (d >= iso) { connect(e[0], e[1]); connect(e[2], e[3]); }
                else { connect(e[0], e[3]); connect(e[2], e[1]); }
            } break;
            default: break;
        }
    }
    Polygon p;
    int pos = 0;
    for (int e = 0; e < 12; ++e) {
        if (seg[e] == 0xFF) continue;
        int out = seg[e] & 0xF;
        int cnt = 0;
        p.edge[pos + cnt++] = e;
        int guard = 0;
        while (out != e && guard++ < 12 && out < 12 && pos + cnt < 12) {
            p.edge[pos + cnt++] = out;
            int in = out;
            out = seg[in] & 0xF;
            seg[in] = 0xFF;
        }
        p.size[p.n] = cnt;
        pos += cnt;
        p.n++;
        if (p.n == 4) break;
    }
    return p;
}

template <class R>
static Polygon cellPolygon(int mcCase, const R s[8], R iso) {
    return caseIsAmbiguous(mcCase) ? polygonAsymptotic(s, iso) : polygonFromTable(mcCase);
}

// VertexToSurfaceIterator, CellIntersection.inl:12-122
template <class R>
struct SurfaceMover {
    const R* s;
    R iso;
    R step;
    V3<R> gradient(V3<R> u) const {
        V3<R> g;
        g.x = (1 - u.z) * ((1 - u.y) * (s[1] - s[0]) + u.y * (s[3] - s[2])) + u.z * ((1 - u.y) * (s[5] - s[4]) + u.y * (s[7] - s[6]));
        g.y = (1 - u.z) * ((1 - u.x) * (s[2] - s[0]) + u.x * (s[3] - s[1])) + u.z * ((1 - u.x) * (s[6] - s[4]) + u.x * (s[7] - s[5]));
        g.z = (1 - u.y) * ((1 - u.x) * (s[4] - s[0]) + u.x * (s[5] - s[1])) + u.y * ((1 - u.x) * (s[6] - s[2]) + u.x * (s[7] - s[3]));
        return g;
    }
    // returns bit0 breakNow, bit1 x, bit2 y, bit3 z clamped
    int update(V3<R>& point, V3<R>& g, const V3<R>& i1, V3<R>& i2, R v1, R& v2) const {
        int r = 0;
        g = normalize(g);
        i2 = i1 + scl(step, g);
        if (i2.x <= R(0) || i2.x >= R(1)) r |= 2;
        if (i2.y <= R(0) || i2.y >= R(1)) r |= 4;
        if (i2.z <= R(0) || i2.z >= R(1)) r |= 8;
        i2 = vclamp(i2, R(0), R(1));
        v2 = trilinear(s, i2);
        if ((v1 <= iso && iso <= v2) || (v2 <= iso && iso <= v1)) {
            R e = v1;
            if (v1 != v2) e = (iso - v1) / (v2 - v1);
            point = i1 + scl(e, i2 - i1);
            r |= 1;
        }
        return r;
    }
    void move(V3<R>& point) const {
        V3<R> i1 = point;
        R v1 = trilinear(s, i1);
        R v2 = R(0);
        for (int it = 0; it < 30; ++it) {
            V3<R> g = gradient(i1);
            if (v1 > iso) g = neg(g);
            V3<R> i2 = i1;
            int r = update(point, g, i1, i2, v1, v2);
            if (r & 1) break;
            if (r & 14) {
                if (r & 2) g.x = R(0);
                else if (r & 4) g.y = R(0);
                else g.z = R(0);
                int r2 = update(point, g, i1, i2, v1, v2);
                if (r2 & 1) break;
            }
            i1 = i2;
            v1 = v2;
        }
        const R eps = R(0.0001f);
        point = vclamp(point, eps, R(1.f - 0.0001f));
    }
};

// --------------------------------------------------------------------------
// mesh state (the buffers owned by DualMarchingCubes / HalfedgeMesh)
struct HE {
    int o, f, n, t;
};
struct Bucket {
    int n = 0;
    int h[4] = {-1, -1, -1, -1};
};

template <class R>
struct Mesh {
    // VertexHostVector
    std::vector<V3<R>> pos, nrm;
    // QuadrilateralHostVector
    std::vector<int> quad;  // 4 per quad
    std::vector<uint16_t> qattr;
    // HalfedgeMesh
    std::vector<HE> he;
    std::vector<int> face_he;
    std::vector<uint16_t> fattr;
    std::vector<int> vert_he;
    std::vector<uint8_t> vflags;
    // edge hash: key -> bucket
    std::unordered_map<uint64_t, int> hmap;
    std::vector<Bucket> buckets;
    // bookkeeping
    int64_t anomalies = 0;
    int64_t splitIterations = 0;
    int64_t dmcVertices = 0, dmcQuads = 0;
};

// Utilities.h:15-22
static inline uint64_t edgeKey(int a, int b) {
    if (a < b) return ((uint64_t)(int64_t)a << 32) | ((uint64_t)(int64_t)b & 0xFFFFFFFFFFFFULL);
    return ((uint64_t)(int64_t)b << 32) | ((uint64_t)(int64_t)a & 0xFFFFFFFFFFFFULL);
}

template <class R>
static void hashClear(Mesh<R>& m) {
    m.hmap.clear();
    m.buckets.clear();
}
// HalfedgeDeviceHashTable::addHalfedge, HalfedgeHashTables.h:46-65; a fifth entry lands in slot 0 (glm operator[] default)
template <class R>
static void hashAdd(Mesh<R>& m, int a, int b, int h) {
    uint64_t k = edgeKey(a, b);
    auto it = m.hmap.find(k);
    int bi;
    if (it == m.hmap.end()) {
        bi = (int)m.buckets.size();
        m.hmap.emplace(k, bi);
        m.buckets.push_back(Bucket());
    } else bi = it->second;
    Bucket& b2 = m.buckets[bi];
    int off = b2.n++;
    if (off >= 4) { m.anomalies++; off = 0; }
    b2.h[off] = h;
}
template <class R>
static int hashNumFaces(const Mesh<R>& m, int a, int b) {
    auto it = m.hmap.find(edgeKey(a, b));
    return it == m.hmap.end() ? -1 : m.buckets[it->second].n;
}

// --------------------------------------------------------------------------
// DMC: performDualMarchingCubes + sliceP + mapQuadrilaterals
template <class R>
struct QuadSlots {
    int v[4] = {-1, -1, -1, -1};
    int color = -1;
};

template <class R>
static void runDMC(Mesh<R>& m, const Grid& g, R iso, bool surfIter) {
    std::unordered_map<uint64_t, QuadSlots<R>> qmap;
    qmap.reserve(1024);
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const V3<R> origin = {R(g.ox), R(g.oy), R(g.oz)};
    const V3<R> deltas = {R(g.dx), R(g.dy), R(g.dz)};
    auto cl = [](int d, int i) { return i < 0 ? 0 : (i >= d ? d - 1 : i); };
    for (int z = 0; z + 1 < nz; ++z)
        for (int y = 0; y + 1 < ny; ++y)
            for (int x = 0; x + 1 < nx; ++x) {
                R s[8];
                for (int k = 0; k < 8; ++k) s[k] = R(g.atc(x + (k & 1), y + ((k >> 1) & 1), z + ((k >> 2) & 1)));
                int mc = 0;
                for (int k = 0; k < 8; ++k) mc |= (s[k] >= iso ? 1 : 0) << k;
                if (mc == 0 || mc == 255) continue;
                Polygon poly = cellPolygon(mc, s, iso);
                // UniformGrid::computeGradients, UniformGrid.inl:325-379
                V3<R> nrm8[8];
                for (int k = 0; k < 8; ++k) {
                    const int cx = k & 1, cy = (k >> 1) & 1, cz = (k >> 2) & 1;
                    const int px = x + cx, py = y + cy, pz = z + cz;
                    R gx = cx ? R(g.at(cl(nx, px + 1), py, pz)) - s[k & ~1] : s[k | 1] - R(g.at(cl(nx, px - 1), py, pz));
                    R gy = cy ? R(g.at(px, cl(ny, py + 1), pz)) - s[k & ~2] : s[k | 2] - R(g.at(px, cl(ny, py - 1), pz));
                    R gz = cz ? R(g.at(px, py, cl(nz, pz + 1))) - s[k & ~4] : s[k | 4] - R(g.at(px, py, cl(nz, pz - 1)));
                    V3<R> gv = {gx, gy, gz};
                    nrm8[k] = scl(R(0.5f), gv) / deltas;
                }
                unsigned edgy = 0;
                const int base = (x & 1) | ((y & 1) << 1) | ((z & 1) << 2);
                for (int t = 0; t < poly.n; ++t) {
                    V3<R> p = {R(0), R(0), R(0)};
                    const int st = poly.start(t), sz = poly.size[t];
                    for (int i = 0; i < sz; ++i) {
                        const int e = poly.edge[st + i];
                        const int a = kEdgeCorner[e][0], b = kEdgeCorner[e][1];
                        const R tt = (iso - s[a]) / (s[b] - s[a]);
                        // computeLocalCoordinateOffset: component = 0 + (t*E + C)
                        const int ax = kEdgeAxis[e];
                        const R cxx = R(a & 1), cyy = R((a >> 1) & 1), czz = R((a >> 2) & 1);
                        V3<R> off;
                        off.x = R(0) + (tt * R(ax == 0 ? 1 : 0) + cxx);
                        off.y = R(0) + (tt * R(ax == 1 ? 1 : 0) + cyy);
                        off.z = R(0) + (tt * R(ax == 2 ? 1 : 0) + czz);
                        p = p + off;
                        if (s[a] < iso) edgy |= 1u << e;
                    }
                    p = divs(p, R(sz));
                    if (surfIter) {
                        SurfaceMover<R> mv{s, iso, R(0.05f)};
                        mv.move(p);
                    }
                    V3<R> n = normalize(trilinearV(nrm8, p));
                    V3<R> ijk = {R(x), R(y), R(z)};
                    V3<R> w = origin + (ijk + p) * deltas;
                    const int vid = (int)m.pos.size();
                    m.pos.push_back(w);
                    m.nrm.push_back(n);
                    for (int i = 0; i < sz; ++i) {
                        const int e = poly.edge[st + i];
                        const int a = kEdgeCorner[e][0];
                        int color = -1;
                        if (e == 0) color = 3 * base;
                        if (e == 3) color = 3 * base + 1;
                        if (e == 8) color = 3 * base + 2;
                        const uint64_t key = 3ULL * (uint64_t)g.gid(x + (a & 1), y + ((a >> 1) & 1), z + ((a >> 2) & 1)) + (uint64_t)kEdgeAxis[e];
                        const int slot = kEdgeSlot[e][(edgy >> e) & 1];
                        QuadSlots<R>& q = qmap[key];
                        q.v[slot] = vid;
                        if (color != -1) q.color = color;
                    }
                }
            }
    m.dmcVertices = (int64_t)m.pos.size();
    std::vector<uint64_t> keys;
    keys.reserve(qmap.size());
    for (auto& kv : qmap)
        if (kv.second.v[0] >= 0 && kv.second.v[1] >= 0 && kv.second.v[2] >= 0 && kv.second.v[3] >= 0) keys.push_back(kv.first);
    std::sort(keys.begin(), keys.end());
    m.quad.reserve(keys.size() * 4);
    for (uint64_t k : keys) {
        const QuadSlots<R>& q = qmap[k];
        for (int i = 0; i < 4; ++i) m.quad.push_back(q.v[i]);
        // QuadrilateralAttribute(): 0x1F, then setColor(hash colour; 0x1F if never written)
        m.qattr.push_back(qaSetColor(QA_COLOR, q.color == -1 ? 0x1F : q.color));
    }
    m.dmcQuads = (int64_t)keys.size();
}

// --------------------------------------------------------------------------
// HalfedgeMesh::initialize: collectHalfedgeElements + connectHalfedgeTwins + refresh*Flags, HalfedgeMesh.cu:14-291, :848-857
template <class R>
static void halfedgeInitialize(Mesh<R>& m) {
    const int Q = (int)(m.quad.size() / 4);
    const int V = (int)m.pos.size();
    m.he.assign((size_t)4 * Q, HE{-1, -1, -1, -1});
    m.face_he.assign(Q, -1);
    m.fattr.assign(Q, QA_COLOR);
    m.vert_he.assign(V, -1);
    m.vflags.assign(V, 0);
    hashClear(m);
    for (int q = 0; q < Q; ++q) {
        const int* v = &m.quad[4 * q];
        for (int i = 0; i < 4; ++i) m.he[4 * q + i] = HE{v[i], q, 4 * q + ((i + 1) & 3), -1};
        m.face_he[q] = 4 * q;
        m.fattr[q] = m.qattr[q];
        for (int i = 0; i < 4; ++i) m.vert_he[v[i]] = 4 * q + i;  // last writer in ascending order
        for (int i = 0; i < 4; ++i) hashAdd(m, v[i], v[(i + 1) & 3], 4 * q + i);
    }
    for (const Bucket& b : m.buckets) {
        switch (b.n) {
            case 2:
                m.he[b.h[0]].t = b.h[1];
                m.he[b.h[1]].t = b.h[0];
                break;
            case 3:
                m.he[b.h[0]].t = b.h[1];
                m.he[b.h[1]].t = b.h[2];
                m.he[b.h[2]].t = b.h[0];
                break;
            case 4: {
                const int v0 = m.he[b.h[0]].o, v1 = m.he[b.h[1]].o, v2 = m.he[b.h[2]].o;
                int pa, pb, pc, pd;
                if (v0 != v1) { pa = 0; pb = 1; pc = 2; pd = 3; }
                else if (v0 != v2) { pa = 0; pb = 2; pc = 1; pd = 3; }
                else { pa = 0; pb = 3; pc = 1; pd = 2; }
                m.he[b.h[pa]].t = b.h[pb];
                m.he[b.h[pb]].t = b.h[pa];
                m.he[b.h[pc]].t = b.h[pd];
                m.he[b.h[pd]].t = b.h[pc];
                for (int i = 0; i < 4; ++i) {
                    const int f = m.he[b.h[i]].f;
                    m.fattr[f] |= QA_NM;
                    m.qattr[f] |= QA_NM;
                }
            } break;
            default: break;
        }
    }
    for (const Bucket& b : m.buckets) {
        if (b.n == 4)
            for (int i = 0; i < 4; ++i) m.vflags[m.he[b.h[i]].o] |= VF_NM;
        if (b.n == 3)
            for (int i = 0; i < 3; ++i) m.vflags[m.he[b.h[i]].o] |= VF_WEIRD;
        if (b.n == 1) m.vflags[m.he[b.h[0]].o] |= VF_BOUNDARY;
    }
}

// HalfedgeMesh::refreshHashTable, HalfedgeMesh.cu:1063-1100 (twins untouched)
template <class R>
static void refreshHashTable(Mesh<R>& m) {
    hashClear(m);
    for (auto& a : m.fattr) a = (uint16_t)(a & ~(QA_PATTERN | QA_REMOVE | QA_NM));
    for (auto& f : m.vflags) f = 0;
    const int F = (int)m.face_he.size();
    for (int f = 0; f < F; ++f) {
        int hs[4], vs[4];
        int h = m.face_he[f];
        for (int i = 0; i < 4; ++i) { vs[i] = m.he[h].o; hs[i] = h; h = m.he[h].n; }
        for (int i = 0; i < 4; ++i) hashAdd(m, vs[i], vs[(i + 1) & 3], hs[i]);
    }
    for (const Bucket& b : m.buckets) {
        if (b.n == 4)
            for (int i = 0; i < 4; ++i) m.vflags[m.he[b.h[i]].o] |= VF_NM;
        if (b.n == 3)
            for (int i = 0; i < 3; ++i) m.vflags[m.he[b.h[i]].o] |= VF_WEIRD;
        if (b.n == 1) m.vflags[m.he[b.h[0]].o] |= VF_BOUNDARY;
    }
    for (const Bucket& b : m.buckets)
        if (b.n == 4)
            for (int i = 0; i < 4; ++i) m.fattr[m.he[b.h[i]].f] |= QA_NM;
}

// --------------------------------------------------------------------------
// simplifications (HalfedgeMesh.cu:296-683, :859-1048)
struct VertexMap {
    std::vector<int> valence, count, type, target, address;
    explicit VertexMap(int n) : valence(n, 0), count(n, 0), type(n, 0), target(n, -1), address(n, -1) {}
};
enum { VT_NEUTRAL = 0, VT_PATTERN = 1, VT_REMOVABLE = 3 };

template <class R>
static void computeValence(const Mesh<R>& m, VertexMap& vm) {
    for (const HE& h : m.he) {
        vm.valence[h.o] += 1;
        if (h.t == -1) vm.valence[m.he[h.n].o] += 1;
    }
}

// getNeighboringFaces (HalfedgeVectors.h); a boundary halfedge (no twin) has no neighbour: the
// reference reads halfedges[-1] there, which is undefined; this restatement reports "no neighbour".
template <class R>
static bool neighborFaces(const Mesh<R>& m, int face, int out[4]) {
    int h = m.face_he[face];
    for (int i = 0; i < 4; ++i) {
        const HE& e = m.he[h];
        if (e.t < 0) return false;
        out[i] = m.he[e.t].f;
        h = e.n;
    }
    return true;
}

template <class R>
static void handleMerge(Mesh<R>& m, VertexMap& vm, int x, int y, int z, int w, uint16_t& attr) {
    if (g_plant == PLANT_SWAP_SURVIVOR) std::swap(x, z);
    // the two corners averaged: the >=5 corners y and w
    const int ma = (g_plant == PLANT_MID_XZ) ? x : y, mb = (g_plant == PLANT_MID_XZ) ? z : w;
    V3<R> p = m.pos[ma];
    p = p + scl(R(0.5f), m.pos[mb] - p);
    m.pos[x] = p;
    V3<R> n = m.nrm[ma];
    n = n + scl(R(0.5f), m.nrm[mb] - n);
    if (g_plant != PLANT_NO_RENORM) n = normalize(n);
    m.nrm[x] = n;
    vm.type[z] = VT_REMOVABLE;
    vm.target[z] = x;
    attr |= QA_REMOVE;
}

template <class R>
static void simplifyRemoveAndRebuild(Mesh<R>& m, VertexMap& vm, bool useFaceAttr, bool p3333) {
    const int V = (int)m.pos.size();
    const int Q = (int)(m.quad.size() / 4);
    std::vector<V3<R>> np, nn;
    for (int v = 0; v < V; ++v) {
        const bool skip = p3333 ? (vm.type[v] == VT_PATTERN) : (vm.type[v] == VT_REMOVABLE);
        if (skip) continue;
        vm.address[v] = (int)np.size();
        np.push_back(m.pos[v]);
        nn.push_back(m.nrm[v]);
    }
    std::vector<int> nq;
    std::vector<uint16_t> na;
    for (int q = 0; q < Q; ++q) {
        const uint16_t a = useFaceAttr ? m.fattr[q] : m.qattr[q];
        if (a & QA_REMOVE) continue;
        int out[4];
        for (int i = 0; i < 4; ++i) {
            const int v = m.quad[4 * q + i];
            if (p3333) out[i] = (a & QA_PATTERN) ? vm.address[vm.target[v]] : vm.address[v];
            else out[i] = (vm.type[v] == VT_REMOVABLE) ? vm.address[vm.target[v]] : vm.address[v];
        }
        for (int i = 0; i < 4; ++i) nq.push_back(out[i]);
        uint16_t at = qaSetColor(QA_COLOR, QA_COLOR);
        if (a & QA_NM) at |= QA_NM;
        if (a & QA_BOUNDARY) at |= QA_BOUNDARY;
        na.push_back(at);
    }
    m.pos.swap(np);
    m.nrm.swap(nn);
    m.quad.swap(nq);
    m.qattr.swap(na);
    halfedgeInitialize(m);
}

// HalfedgeMesh::simplify3X3Y (colour based), HalfedgeMesh.cu:859-923
template <class R>
static void simplify3X3YColor(Mesh<R>& m) {
    VertexMap vm((int)m.pos.size());
    computeValence(m, vm);
    const int Q = (int)(m.quad.size() / 4);
    for (int q = 0; q < Q; ++q) {  // markElementsP3X3Y
        uint16_t& a = m.qattr[q];
        if ((a & QA_NM) || (a & QA_BOUNDARY)) continue;
        a = (uint16_t)(a & 0x3F);
        const int* v = &m.quad[4 * q];
        int val[4];
        for (int i = 0; i < 4; ++i) val[i] = vm.valence[v[i]];
        const bool f1 = val[0] == 3 && val[1] >= 5 && val[2] == 3 && val[3] >= 5;
        const bool f2 = val[0] >= 5 && val[1] == 3 && val[2] >= 5 && val[3] == 3;
        if (f1 || f2) a |= QA_PATTERN;
    }
    for (int q = 0; q < Q; ++q) {  // mergeVerticesP3X3YColor
        uint16_t& a = m.qattr[q];
        if (!(a & QA_PATTERN)) continue;
        int nb[4];
        if (!neighborFaces(m, q, nb)) continue;
        bool stop = false;
        for (int i = 0; i < 4; ++i)
            if (((m.qattr[nb[i]] & QA_NM) && g_plant != PLANT_NO_NEIGHBOR_NM) || (a & QA_BOUNDARY)) stop = true;
        if (stop) continue;
        const int color = a & QA_COLOR;
        for (int i = 0; i < 4; ++i) {
            const int nc = m.qattr[nb[i]] & QA_COLOR;
            const bool lower = (g_plant == PLANT_COLOR_LT) ? (nc < color) : (nc <= color);
            if ((m.qattr[nb[i]] & QA_PATTERN) && lower) stop = true;
        }
        if (stop) continue;
        const int* v = &m.quad[4 * q];
        int val[4];
        for (int i = 0; i < 4; ++i) val[i] = vm.valence[v[i]];
        if (val[0] == 3 && val[1] >= 5 && val[2] == 3 && val[3] >= 5) handleMerge(m, vm, v[0], v[1], v[2], v[3], a);
        else if (val[0] >= 5 && val[1] == 3 && val[2] >= 5 && val[3] == 3) handleMerge(m, vm, v[1], v[0], v[3], v[2], a);
    }
    simplifyRemoveAndRebuild(m, vm, false, false);
}

// HalfedgeMesh::simplify3X3YOld, HalfedgeMesh.cu:925-990
template <class R>
static void simplify3X3YOld(Mesh<R>& m) {
    VertexMap vm((int)m.pos.size());
    computeValence(m, vm);
    const int Q = (int)(m.quad.size() / 4);
    for (int q = 0; q < Q; ++q) {  // countVerticesP3X3Y
        uint16_t& a = m.qattr[q];
        a = (uint16_t)(a & 0x3F);
        const int* v = &m.quad[4 * q];
        int val[4];
        for (int i = 0; i < 4; ++i) val[i] = vm.valence[v[i]];
        if (val[0] == 3 && val[1] >= 5 && val[2] == 3 && val[3] >= 5) {
            vm.count[v[0]]++; vm.count[v[2]]++; a |= QA_PATTERN;
        } else if (val[0] >= 5 && val[1] == 3 && val[2] >= 5 && val[3] == 3) {
            vm.count[v[1]]++; vm.count[v[3]]++; a |= QA_PATTERN;
        }
    }
    for (int q = 0; q < Q; ++q) {  // mergeVerticesP3X3Y
        uint16_t& a = m.qattr[q];
        if ((a & QA_NM) || (a & QA_BOUNDARY)) continue;
        if (!(a & QA_PATTERN)) continue;
        int nb[4];
        if (!neighborFaces(m, q, nb)) continue;
        bool stop = false;
        for (int i = 0; i < 4; ++i)
            if ((m.qattr[nb[i]] & QA_NM) || (a & QA_BOUNDARY)) stop = true;
        if (stop) continue;
        const int* v = &m.quad[4 * q];
        int val[4];
        for (int i = 0; i < 4; ++i) val[i] = vm.valence[v[i]];
        const bool anyCount = g_plant == PLANT_OLD_NO_COUNT;
        if (val[0] == 3 && val[1] >= 5 && val[2] == 3 && val[3] >= 5 && ((vm.count[v[0]] == 1 && vm.count[v[2]] == 1) || anyCount))
            handleMerge(m, vm, v[0], v[1], v[2], v[3], a);
        else if (val[0] >= 5 && val[1] == 3 && val[2] >= 5 && val[3] == 3 && ((vm.count[v[1]] == 1 && vm.count[v[3]] == 1) || anyCount))
            handleMerge(m, vm, v[1], v[0], v[3], v[2], a);
    }
    simplifyRemoveAndRebuild(m, vm, false, false);
}

// HalfedgeMesh::simplify3333, HalfedgeMesh.cu:547-683, :992-1048.  The per-face
// clearPatternFlags of markElementsP3333 is taken to precede every neighbour
// marking (otherwise a later face could erase a removal set by an earlier one).
template <class R>
static void simplify3333(Mesh<R>& m) {
    VertexMap vm((int)m.pos.size());
    computeValence(m, vm);
    const int F = (int)m.face_he.size();
    for (int f = 0; f < F; ++f) {
        uint16_t& a = m.fattr[f];
        if ((a & QA_NM) || (a & QA_BOUNDARY)) continue;
        a = (uint16_t)(a & 0x3F);
    }
    for (int f = 0; f < F; ++f) {
        const uint16_t a0 = m.fattr[f];
        if ((a0 & QA_NM) || (a0 & QA_BOUNDARY)) continue;
        HE qe[4];
        qe[0] = m.he[m.face_he[f]];
        bool ok = true;
        for (int i = 0; i < 4 && ok; ++i) {
            if (i < 3) qe[i + 1] = m.he[qe[i].n];
            if (vm.valence[qe[i].o] != 3) ok = false;
        }
        if (!ok) continue;
        int nbf[4], nbv[4], nbval[4];
        for (int i = 0; i < 4 && ok; ++i) {
            const int tw = qe[i].t;
            if (tw < 0) { ok = false; break; }  // boundary: no neighbour (reference reads halfedges[-1])
            nbf[i] = m.he[tw].f;
            const uint16_t na = m.fattr[nbf[i]];
            if (((na & QA_NM) && g_plant != PLANT_NO_NEIGHBOR_NM) || (na & QA_BOUNDARY)) { ok = false; break; }
            int hh = m.he[m.he[tw].n].n;
            if (g_plant == PLANT_FAR_CORNER) hh = m.he[hh].n;
            nbv[i] = m.he[hh].o;
            nbval[i] = vm.valence[nbv[i]];
            if (nbval[i] == 3) { ok = false; break; }
        }
        if (!ok) continue;
        bool allFour = true;
        for (int i = 0; i < 4; ++i) allFour = allFour && nbval[i] == 4;
        if (allFour && g_plant != PLANT_NO_ALL_FOUR) continue;
        m.fattr[f] |= QA_PATTERN;
        for (int i = 0; i < 4; ++i) {
            m.fattr[nbf[i]] |= QA_REMOVE;
            vm.type[qe[i].o] = VT_PATTERN;
            vm.target[qe[i].o] = nbv[i];
        }
    }
    simplifyRemoveAndRebuild(m, vm, true, true);
}

// --------------------------------------------------------------------------
// ManifoldCreator::fixManifoldTwins -> reconnectHalfedges, ManifoldCreator.cu:39-312
// floor of a cell coordinate; a non-finite or huge one becomes -2^30 (outside every grid),
// the same value the HIP path uses, instead of an undefined conversion
template <class R>
static inline int cellIndex(R q) { return (q > R(-1.0e9) && q < R(1.0e9)) ? (int)std::floor(q) : -(1 << 30); }
template <class R>
static inline void cellOf(const Grid& g, V3<R> p, int c[3]) {
    V3<R> o = {R(g.ox), R(g.oy), R(g.oz)}, d = {R(g.dx), R(g.dy), R(g.dz)};
    V3<R> q = (p - o) / d;
    c[0] = cellIndex(q.x); c[1] = cellIndex(q.y); c[2] = cellIndex(q.z);
}
template <class R>
static inline V3<R> nudge(int it) {
    R x = (it & 1) ? R(1.f) : R(0.f), y = (it & 2) ? R(1.f) : R(0.f), z = (it & 4) ? R(1.f) : R(0.f);
    if (it & 8) x = -x;
    if (it & 16) y = -y;
    if (it & 32) z = -z;
    const R eps = R(0.0001f);
    return {x * eps, y * eps, z * eps};
}

template <class R>
static void fixManifoldTwins(Mesh<R>& m, const Grid& g, R iso) {
    for (const Bucket& b : m.buckets) {
        if (!(b.h[0] != -1 && b.h[1] != -1 && b.h[2] != -1 && b.h[3] != -1)) continue;
        const HE& h0 = m.he[b.h[0]];
        V3<R> p0 = m.pos[h0.o];
        const int nxt = h0.n;
        V3<R> p1 = m.pos[m.he[nxt].o];
        int c0[3], c1[3];
        cellOf(g, p0, c0);
        cellOf(g, p1, c1);
        bool found = false;
        for (int i = 0; i < 64 && !found; ++i) {
            V3<R> o0 = nudge<R>(i);
            for (int k = 0; k < 64; ++k) {
                V3<R> o1 = nudge<R>(k);
                int a[3], bb[3];
                cellOf(g, p0 + o0, a);
                cellOf(g, p1 + o1, bb);
                const int sum = std::abs(a[0] - bb[0]) + std::abs(a[1] - bb[1]) + std::abs(a[2] - bb[2]);
                if (sum == 1) {
                    p0 = p0 + o0; cellOf(g, p0, c0);
                    p1 = p1 + o1; cellOf(g, p1, c1);
                    found = true;
                    break;
                }
            }
        }
        const int64_t g0 = g.gid(c0[0], c0[1], c0[2]), g1 = g.gid(c1[0], c1[1], c1[2]);
        int ca[3], cb[3];
        if (g0 < g1) { std::memcpy(ca, c0, sizeof ca); std::memcpy(cb, c1, sizeof cb); }
        else { std::memcpy(ca, c1, sizeof ca); std::memcpy(cb, c0, sizeof cb); }
        // computeOffsetDirections / computeCellWallValues / computeCellWallVertices
        int wn[3];
        for (int i = 0; i < 3; ++i) wn[i] = std::max(ca[i], cb[i]) - std::min(ca[i], cb[i]);
        int xo[3] = {0, 0, 0}, yo[3] = {0, 0, 0};
        if (wn[0] == 0) {
            xo[0] = 1;
            if (wn[1] == 0) yo[1] = 1;
            else yo[2] = 1;
        } else { xo[1] = 1; yo[2] = 1; }
        int mi[3];
        for (int i = 0; i < 3; ++i) mi[i] = std::max(ca[i], cb[i]);
        if (mi[0] < 0 || mi[1] < 0 || mi[2] < 0 || mi[0] >= g.nx - 1 || mi[1] >= g.ny - 1 || mi[2] >= g.nz - 1) { m.anomalies++; continue; }
        auto gv = [&](int dx, int dy, int dz) { return R(g.at(mi[0] + dx, mi[1] + dy, mi[2] + dz)); };
        R wall[4] = {gv(0, 0, 0), gv(xo[0], xo[1], xo[2]), gv(xo[0] + yo[0], xo[1] + yo[1], xo[2] + yo[2]), gv(yo[0], yo[1], yo[2])};
        V3<R> o = {R(g.ox), R(g.oy), R(g.oz)}, d = {R(g.dx), R(g.dy), R(g.dz)};
        auto corner = [&](int dx, int dy, int dz) {
            V3<R> iv = {R(mi[0] + dx), R(mi[1] + dy), R(mi[2] + dz)};
            return o + d * iv;
        };
        V3<R> pc0 = corner(0, 0, 0), pc2 = corner(xo[0] + yo[0], xo[1] + yo[1], xo[2] + yo[2]);
        const bool omitX = xo[0] == 0 && yo[0] == 0;
        const bool omitY = xo[1] == 0 && yo[1] == 0;
        auto down = [&](V3<R> p, R& u, R& v) {
            if (omitX) { u = p.y; v = p.z; }
            else if (omitY) { u = p.x; v = p.z; }
            else { u = p.x; v = p.y; }
        };
        R minU, minV, maxU, maxV;
        down(pc0, minU, minV);
        down(pc2, maxU, maxV);
        int sorted[4] = {-1, -1, -1, -1};
        for (int i = 0; i < 4; ++i) {
            const int hid = b.h[i];
            int hh = m.he[m.he[hid].n].n;
            V3<R> vp = m.pos[m.he[hh].o];
            R u, v;
            down(vp, u, v);
            const R mnu = u - minU, mnv = v - minV, mxu = u - maxU, mxv = v - maxV;
            if (mnv < R(0)) sorted[0] = hid;
            if (mxu > R(0)) sorted[1] = hid;
            if (mxv > R(0)) sorted[2] = hid;
            if (mnu < R(0)) sorted[3] = hid;
        }
        bool bad = false;
        for (int i = 0; i < 4; ++i) bad = bad || sorted[i] == -1;
        if (bad) { m.anomalies++; continue; }
        // AsymptoticDecider2D(wall[0], wall[1], wall[2], wall[3]).quadrant0BelongsTogether, AsymptoticDecider.h:107-164
        const R m00 = wall[0], m10 = wall[1], m11 = wall[2], m01 = wall[3];
        const R C = m00 - m10 - m01 + m11;
        const R delta = (m00 * m11 - m10 * m01) / C;
        const bool q0 = (m00 < iso && delta >= iso) || (m00 >= iso && delta < iso);
        const bool connect01 = !q0;
        if (connect01) {
            m.he[sorted[0]].t = sorted[1]; m.he[sorted[1]].t = sorted[0];
            m.he[sorted[2]].t = sorted[3]; m.he[sorted[3]].t = sorted[2];
        } else {
            m.he[sorted[0]].t = sorted[3]; m.he[sorted[3]].t = sorted[0];
            m.he[sorted[1]].t = sorted[2]; m.he[sorted[2]].t = sorted[1];
        }
    }
}

// --------------------------------------------------------------------------
// ManifoldSplitter, ManifoldSplitter.h / ManifoldSplitter.cu:25-1387
struct MergeList {
    std::vector<int> num;
    std::vector<int> ex, ey;  // glm::ivec2 storage
    explicit MergeList(int n) : num(n, 1), ex(n, -1), ey(n, -1) {}
    int elem(int a, int i) const { return i == 1 ? ey[a] : ex[a]; }  // glm operator[] default -> x
    void mark(int a, int b, int64_t& anomalies) {
        if (a < 0 || a >= (int)num.size()) { anomalies++; return; }
        const int idx = num[a]++;
        if (idx > 5) return;
        if (idx - 1 == 1) ey[a] = b;
        else ex[a] = b;
    }
    // DeviceMergeList::getAllPartnersOf
    int partners(int v, int out[12], int64_t& anomalies) const {
        for (int i = 0; i < 12; ++i) out[i] = -1;
        if (v < 0 || v >= (int)num.size()) { anomalies++; return 1; }
        int n = num[v];
        if (n <= 1) return n;
        for (int i = 0; i < n - 1; ++i) {
            if (i >= 12) { anomalies++; break; }
            out[i] = elem(v, i);
        }
        for (int i = 0; i < n - 1 && i < 12; ++i) {
            const int p = out[i];
            if (p < 0 || p >= (int)num.size()) { anomalies++; continue; }
            const int np = num[p];
            for (int k = 0; k < np - 1; ++k) {
                const int pp = elem(p, k);
                bool contained = v == pp;
                for (int z = 0; z < n - 1 && z < 12; ++z) contained = contained || out[z] == pp;
                if (!contained) {
                    if (n - 1 >= 12) { anomalies++; continue; }
                    out[n - 1] = pp;
                    ++n;
                }
            }
        }
        return n;
    }
};

template <class R>
struct WCVertices {
    std::vector<V3<R>> pos, nrm;
    std::vector<uint8_t> flags;
    std::vector<int> c2o, o2c;
    int cap = 0, size = 0;
    WCVertices(int capacity, int orig) : pos(capacity, V3<R>{0, 0, 0}), nrm(capacity, V3<R>{0, 0, 0}), flags(capacity, 0), c2o(capacity, -1), o2c((size_t)orig + capacity, -1), cap(capacity) {}
    int add(V3<R> p, V3<R> n) {
        if (size >= cap) return -1;
        pos[size] = p; nrm[size] = n;
        return size++;
    }
};
struct WCHalfedges {
    std::vector<HE> he;
    std::vector<uint8_t> flags;
    std::vector<int> c2o, o2c;
    int cap = 0, size = 0;
    WCHalfedges(int capacity, int orig) : he(capacity, HE{-1, -1, -1, -1}), flags(capacity, 0), c2o(capacity, -1), o2c((size_t)orig + capacity, -1), cap(capacity) {}
    int add(HE e) {
        if (size >= cap) return -1;
        he[size] = e;
        return size++;
    }
};
struct WCFaces {
    std::vector<int> first;
    std::vector<uint16_t> attr;
    std::vector<uint8_t> flags;
    std::vector<int> c2o, o2c;
    int cap = 0, size = 0;
    WCFaces(int capacity, int orig) : first(capacity, -1), attr(capacity, QA_COLOR), flags(capacity, 0), c2o(capacity, -1), o2c((size_t)orig + capacity, -1), cap(capacity) {}
    int push(int f, uint16_t a) {
        if (size >= cap) return -1;
        first[size] = f; attr[size] = a;
        return size++;
    }
};

template <class R>
static void moveToSurfaceInCell(const Grid& g, R iso, V3<R>& point) {
    V3<R> o = {R(g.ox), R(g.oy), R(g.oz)}, d = {R(g.dx), R(g.dy), R(g.dz)};
    int c[3];
    cellOf(g, point, c);
    R s[8];
    for (int k = 0; k < 8; ++k) s[k] = R(g.atc(c[0] + (k & 1), c[1] + ((k >> 1) & 1), c[2] + ((k >> 2) & 1)));
    SurfaceMover<R> mv{s, iso, R(0.05f)};
    V3<R> iv = {R(c[0]), R(c[1]), R(c[2])}, iv1 = {R(c[0] + 1), R(c[1] + 1), R(c[2] + 1)};
    V3<R> mn = o + iv * d, mx = o + iv1 * d;
    point = (point - mn) / (mx - mn);
    mv.move(point);
    point = o + (iv + point) * d;
}

template <class R>
static void splitNonManifold(Mesh<R>& m, const Grid& g, R iso, int step, int maxIters) {
    const bool copyDataBack = true;
    int iter = 0;
    while (true) {
        const int F = (int)m.face_he.size();
        int n = 0;
        for (int f = 0; f < F; ++f) n += (m.fattr[f] & QA_NM) ? 1 : 0;
        if (n == 0) break;
        // still non-manifold after the allowed iterations (the reference loops until clean)
        if (iter >= maxIters) { m.anomalies++; break; }
        const int H = (int)m.he.size();
        const int V = (int)m.pos.size();
        WCHalfedges wh(20 * n, H);
        WCVertices<R> wv(8 * n, V);
        WCFaces wf(5 * n, F);
        // copyVerticesToWorkingCopy + markDuplicateVertices
        for (int f = 0; f < F; ++f) {
            if (!(m.fattr[f] & QA_NM)) continue;
            int h = m.face_he[f];
            for (int i = 0; i < 4; ++i) {
                const int v = m.he[h].o;
                const int c = wv.add(m.pos[v], m.nrm[v]);
                wv.c2o[c] = v;
                wv.o2c[v] = c;
                h = m.he[h].n;
            }
        }
        for (size_t v = 0; v < wv.o2c.size(); ++v)
            if (wv.o2c[v] != -1) wv.flags[wv.o2c[v]] &= (uint8_t)~VF_REMOVE;
        // copyHalfedgesToWorkingCopy
        for (int f = 0; f < F; ++f) {
            if (!(m.fattr[f] & QA_NM)) continue;
            int h = m.face_he[f];
            for (int i = 0; i < 4; ++i) {
                const HE e = m.he[h];
                const int c = wh.add(e);
                wh.c2o[c] = h;
                wh.o2c[h] = c;
                h = e.n;
            }
        }
        // copyFacesToWorkingCopy
        for (int f = 0; f < F; ++f) {
            if (!(m.fattr[f] & QA_NM)) continue;
            const int first = wh.o2c[m.face_he[f]];
            const int c = wf.push(first, m.fattr[f]);
            wf.c2o[c] = f;
            wf.o2c[f] = c;
        }
        // updateWorkingCopyHalfedges
        for (int c = 0; c < wh.size; ++c) {
            HE& e = wh.he[c];
            e.o = wv.o2c[e.o];
            e.f = wf.o2c[e.f];
            e.n = wh.o2c[e.n];
            if (e.t >= 0) {
                const int mt = wh.o2c[e.t];
                e.t = (mt == -1) ? (-e.t - 2) : mt;
            }
        }
        // updateFlags: markHalfedgesAsNonManifold
        for (const Bucket& b : m.buckets) {
            if (!(b.h[0] != -1 && b.h[1] != -1 && b.h[2] != -1 && b.h[3] != -1)) continue;
            for (int i = 0; i < 4; ++i) {
                const int c = wh.o2c[b.h[i]];
                if (c < 0) { m.anomalies++; continue; }
                wh.flags[c] |= VF_NM;
                wv.flags[wh.he[c].o] |= VF_NM;
            }
        }
        hashClear(m);  // freeHashTable
        // subdivideQuads -> performSubdivideQuads
        if (step >= S_SUBDIVIDE) {
            const R uvs[4][2] = {{R(0.33f), R(0.33f)}, {R(0.67f), R(0.33f)}, {R(0.67f), R(0.67f)}, {R(0.33f), R(0.67f)}};
            for (int f = 0; f < F; ++f) {
                if (!(m.fattr[f] & QA_NM)) continue;
                const int faceID = wf.o2c[f];
                int fh[4];
                HE fe[4];
                int cur = wf.first[faceID];
                for (int i = 0; i < 4; ++i) { fh[i] = cur; fe[i] = wh.he[cur]; cur = fe[i].n; }
                int vid[4];
                V3<R> vp[4], vn[4];
                for (int i = 0; i < 4; ++i) { vid[i] = fe[i].o; vp[i] = wv.pos[vid[i]]; vn[i] = wv.nrm[vid[i]]; }
                int inner[4];
                for (int k = 0; k < 4; ++k) {
                    const R u = uvs[k][0], v = uvs[k][1];
                    V3<R> ap = mixv(vp[0], vp[1], u), bp = mixv(vp[3], vp[2], u);
                    V3<R> an = mixv(vn[0], vn[1], u), bn = mixv(vn[3], vn[2], u);
                    inner[k] = wv.add(mixv(ap, bp, v), mixv(an, bn, v));
                    wv.flags[inner[k]] &= (uint8_t)~VF_REMOVE;
                }
                int ih[4];
                for (int i = 0; i < 4; ++i) ih[i] = wh.add(HE{inner[i], -1, -1, -1});
                for (int i = 0; i < 4; ++i) wh.he[ih[i]].n = ih[(i + 1) & 3];
                for (int i = 0; i < 4; ++i) wh.he[ih[i]].f = faceID;
                wf.first[faceID] = ih[0];
                int of[4][4];
                int prev = 3;
                for (int i = 0; i < 4; ++i) {
                    const int qv[3] = {vid[i], inner[i], inner[prev]};
                    for (int k = 0; k < 3; ++k) of[i][k] = wh.add(HE{qv[k], -1, -1, -1});
                    of[i][3] = fh[prev];
                    for (int k = 0; k < 4; ++k) wh.he[of[i][k]].n = of[i][(k + 1) & 3];
                    const int nf = wf.push(of[i][0], QA_COLOR);
                    for (int k = 0; k < 4; ++k) wh.he[of[i][k]].f = nf;
                    wf.first[nf] = of[i][0];
                    prev = i;
                }
                prev = 3;
                for (int i = 0; i < 4; ++i) {
                    wh.he[of[i][2]].t = of[prev][0]; wh.he[of[prev][0]].t = of[i][2];
                    wh.he[of[i][1]].t = ih[prev]; wh.he[ih[prev]].t = of[i][1];
                    prev = i;
                }
                for (int i = 0; i < 4; ++i) {  // swapHalfedgeMappings(outer, inner)
                    const int a = fh[i], b = ih[i];
                    const int ao = wh.c2o[a], bo = wh.c2o[b];
                    wh.c2o[a] = bo;
                    wh.c2o[b] = ao;
                    const bool ae = ao != -1, be = bo != -1;
                    if (ae && !be) wh.o2c[ao] = b;
                    else m.anomalies++;
                }
            }
        }
        // collapseNonManifoldEdges
        if (step >= S_COLLAPSE) {
            MergeList mlv(wv.size), mlh(wh.size);
            for (int h = 0; h < wh.size; ++h) {
                if (!(wh.flags[h] & VF_NM)) continue;
                const int tw = wh.he[h].t;
                if (tw < 0) { m.anomalies++; continue; }
                if (tw < h) continue;
                int cf[2][4];
                HE cur = wh.he[h];
                cf[0][0] = h;
                HE ct = wh.he[cur.t];
                cf[1][0] = cur.t;
                for (int i = 1; i < 4; ++i) {
                    cf[0][i] = cur.n;
                    cf[1][4 - i] = ct.n;
                    cur = wh.he[cur.n];
                    ct = wh.he[ct.n];
                }
                for (int fc = 0; fc < 2; ++fc) {
                    for (int i = 0; i < 4; ++i) wh.flags[cf[fc][i]] |= VF_REMOVE;
                    wf.flags[wh.he[cf[fc][0]].f] |= VF_REMOVE;
                }
                {
                    const int a = wh.he[cf[0][2]].o, b = wh.he[cf[1][1]].o;
                    mlv.mark(a, b, m.anomalies);
                    mlv.mark(b, a, m.anomalies);
                }
                {
                    const int a = wh.he[cf[0][3]].o, b = wh.he[cf[1][2]].o;
                    mlv.mark(a, b, m.anomalies);
                    mlv.mark(b, a, m.anomalies);
                }
                for (int i = 1; i < 4; ++i) {
                    const int ta = wh.he[cf[0][i]].t, tb = wh.he[cf[1][i]].t;
                    mlh.mark(cf[0][i], tb, m.anomalies);
                    mlh.mark(tb, cf[0][i], m.anomalies);
                    mlh.mark(cf[1][i], ta, m.anomalies);
                    mlh.mark(ta, cf[1][i], m.anomalies);
                }
            }
            if (step >= S_MERGE) {
                // mergeVertices
                for (int v = 0; v < wv.size; ++v) {
                    int P[12];
                    const int np = mlv.partners(v, P, m.anomalies);
                    if (np == 1) continue;
                    if (np < 1 || np > 12) { m.anomalies++; continue; }
                    int smallest = v;
                    V3<R> p = wv.pos[v], nn = wv.nrm[v];
                    bool obsolete = false;
                    for (int i = 0; i < np - 1; ++i) {
                        const int q = P[i];
                        if (q < 0 || q >= wv.size) { m.anomalies++; continue; }
                        p = p + wv.pos[q];
                        nn = nn + wv.nrm[q];
                        if (q < smallest) { obsolete = true; smallest = q; }
                    }
                    if (obsolete) { wv.flags[v] |= VF_REMOVE; continue; }
                    p = divs(p, R(np));
                    nn = divs(nn, R(np));
                    wv.pos[smallest] = p;
                    wv.nrm[smallest] = p;  // the reference stores the position as the normal here
                }
                // updateHalfedgeVertexIDs
                for (int h = 0; h < wh.size; ++h) {
                    const int v = wh.he[h].o;
                    int P[12];
                    const int np = mlv.partners(v, P, m.anomalies);
                    if (np == 1) continue;
                    int smallest = v;
                    for (int i = 0; i < np - 1; ++i)
                        if (P[i] >= 0 && P[i] < smallest) smallest = P[i];
                    wh.he[h].o = smallest;
                }
            }
            if (step >= S_TWINS) {
                // mergeHalfedges
                for (int h = 0; h < wh.size; ++h) {
                    if (wh.flags[h] & VF_REMOVE) continue;
                    int P[12];
                    const int np = mlh.partners(h, P, m.anomalies);
                    if (np == 1) continue;
                    int valid = -1;
                    for (int i = 0; i < np - 1; ++i) {
                        if (P[i] < 0 || P[i] >= wh.size) continue;
                        const int tw = wh.he[P[i]].t;
                        if (tw < 0 || tw >= wh.size) { m.anomalies++; continue; }
                        if (!(wh.flags[tw] & VF_REMOVE)) { valid = tw; break; }
                    }
                    if (valid == -1) continue;
                    if (valid < h) continue;
                    wh.he[h].t = valid;
                    wh.he[valid].t = h;
                }
            }
        }
        // moveVerticesToSurface
        if (step >= S_MOVE) {
            for (int v = 0; v < wv.size; ++v) {
                if (wv.flags[v] & VF_REMOVE) continue;
                moveToSurfaceInCell(g, iso, wv.pos[v]);
            }
        }
        // copyWorkingCopiesBack
        if (copyDataBack || step >= S_ITERATIVE) {
            const int Fsz = (int)m.face_he.size(), Hsz = (int)m.he.size();
            m.vert_he.resize((size_t)m.pos.size() + wv.size, 0);
            m.vflags.resize((size_t)m.pos.size() + wv.size, 0);
            (void)Fsz; (void)Hsz;
            for (int v = 0; v < wv.size; ++v) {
                if (wv.flags[v] & VF_REMOVE) continue;
                int o = wv.c2o[v];
                if (o == -1) {
                    o = (int)m.pos.size();
                    m.pos.push_back(wv.pos[v]);
                    m.nrm.push_back(wv.nrm[v]);
                    wv.c2o[v] = o;
                } else {
                    m.pos[o] = wv.pos[v];
                    m.nrm[o] = wv.nrm[v];
                }
            }
            for (int h = 0; h < wh.size; ++h) {
                if (wh.flags[h] & VF_REMOVE) continue;
                int o = wh.c2o[h];
                if (o == -1) {
                    o = (int)m.he.size();
                    m.he.push_back(wh.he[h]);
                    wh.c2o[h] = o;
                } else m.he[o] = wh.he[h];
            }
            for (int f = 0; f < wf.size; ++f) {
                if (wf.flags[f] & VF_REMOVE) continue;
                int o = wf.c2o[f];
                const int first = wh.c2o[wf.first[f]];
                if (o == -1) {
                    o = (int)m.face_he.size();
                    m.face_he.push_back(first);
                    m.fattr.push_back(wf.attr[f]);
                    wf.c2o[f] = o;
                } else {
                    m.face_he[o] = first;
                    m.fattr[o] = wf.attr[f];
                }
            }
            for (int h = 0; h < wh.size; ++h) {
                if (wh.flags[h] & VF_REMOVE) continue;
                HE& e = m.he[wh.c2o[h]];
                e.n = wh.c2o[e.n];
                e.f = wf.c2o[e.f];
                e.o = wv.c2o[e.o];
                int tw = e.t;
                if (tw < 0) {
                    tw = -tw - 2;
                    e.t = tw;
                    if (tw != -1) m.he[tw].t = wh.c2o[h];
                } else e.t = wh.c2o[tw];
            }
        }
        refreshHashTable(m);
        ++iter;
        m.splitIterations = iter;
        if (!(step >= S_ITERATIVE) || !copyDataBack) break;
    }
}

// --------------------------------------------------------------------------
// ManifoldCreator::detectNonManifoldVertices, ManifoldCreator.cu:319-536
template <class R>
static int fanWalk(const Mesh<R>& m, int start, bool cw, int64_t& anomalies) {
    auto step = [&](int h) -> int {
        const HE& e = m.he[h];
        if (cw) {
            if (e.t != -1) return m.he[e.t].n;
            return -1;
        }
        const HE& s = m.he[m.he[m.he[e.n].n].n];
        if (s.t != -1) return s.t;
        return -1;
    };
    const int origin = m.he[start].o;
    int result = start;
    int cur = step(start);
    int guard = 0;
    while (cur != -1) {
        if (++guard > 4096) { anomalies++; break; }
        const HE& e = m.he[cur];
        if (cur == result) break;
        if (origin != e.o) break;
        if (cur < result) result = cur;
        cur = step(cur);
    }
    return result;
}

template <class R>
static void detectNonManifoldVertices(Mesh<R>& m) {
    const int H = (int)m.he.size();
    std::vector<std::pair<int, int>> fans(H);  // (origin, key)
    for (int h = 0; h < H; ++h) {
        const int a = fanWalk(m, h, true, m.anomalies);
        const int b = fanWalk(m, h, false, m.anomalies);
        const int key = std::min(h, std::min(a, b));
        fans[h] = {m.he[h].o, key};
    }
    std::sort(fans.begin(), fans.end());
    std::vector<int> num(m.vflags.size(), 0);
    for (int i = 0; i < H; ++i)
        if (i == 0 || fans[i] != fans[i - 1]) num[fans[i].first]++;
    for (size_t v = 0; v < m.vflags.size(); ++v) {
        if (num[v] > 1) m.vflags[v] |= VF_FANS;
        else m.vflags[v] &= (uint8_t)~VF_FANS;
    }
}

// --------------------------------------------------------------------------
struct Stats {
    int64_t n_vertices, n_halfedges, n_faces;
    int64_t n_dmc_vertices, n_dmc_quads;
    int64_t nm_edges4, nm_edges3, n_edges, nm_vertices;
    int64_t quads_with_nm[5];
    int64_t split_iterations;
    int64_t anomalies;
};

template <class R>
struct Result {
    Mesh<R> mesh;
    Stats st;
    Grid grid;
};

template <class R>
static void fillStats(const Mesh<R>& m, int flags, Stats& s);

template <class R>
static Result<R>* run(const float* data, const int dims[3], const float origin[3], const float delta[3], float isoF, int flags, int refinement, int step, int maxIters) {
    auto* res = new Result<R>();
    res->grid = prepareGrid(data, dims, origin, delta, refinement, (flags & F_PAD1) != 0);
    Mesh<R>& m = res->mesh;
    const R iso = R(isoF);
    runDMC(m, res->grid, iso, (flags & F_SURF_ITER) != 0);
    if (!m.pos.empty()) {
        halfedgeInitialize(m);
        if (flags & F_P3X3Y_COLOR) simplify3X3YColor(m);
        if (flags & F_P3X3Y_OLD) simplify3X3YOld(m);
        if (flags & F_P3333) simplify3333(m);
        fixManifoldTwins(m, res->grid, iso);
        if (flags & F_SPLIT_NM) splitNonManifold(m, res->grid, iso, step, maxIters);
        if (flags & F_DETECT_NM_VERTS) detectNonManifoldVertices(m);
        // the vertex flag/pointer arrays may be longer than the vertex list after a split
        m.vflags.resize(m.pos.size());
        m.vert_he.resize(m.pos.size());
    }
    fillStats(m, flags, res->st);
    return res;
}

template <class R>
static void fillStats(const Mesh<R>& m, int flags, Stats& s) {
    std::memset(&s, 0, sizeof s);
    s.n_vertices = (int64_t)m.pos.size();
    s.n_halfedges = (int64_t)m.he.size();
    s.n_faces = (int64_t)m.face_he.size();
    s.n_dmc_vertices = m.dmcVertices;
    s.n_dmc_quads = m.dmcQuads;
    for (const Bucket& b : m.buckets) {
        if (b.n == 4) s.nm_edges4++;
        if (b.n == 3) s.nm_edges3++;
        if (b.n > 0) s.n_edges++;
    }
    s.nm_vertices = -1;
    if (flags & F_DETECT_NM_VERTS) {
        s.nm_vertices = 0;
        for (uint8_t f : m.vflags) s.nm_vertices += (f & VF_FANS) ? 1 : 0;
    }
    for (size_t f = 0; f < m.face_he.size() && !m.pos.empty(); ++f) {
        int h = m.face_he[f];
        int k = 0;
        for (int i = 0; i < 4; ++i) {
            const int nh = m.he[h].n;
            if (hashNumFaces(m, m.he[h].o, m.he[nh].o) > 2) ++k;
            h = nh;
        }
        s.quads_with_nm[k]++;
    }
    s.split_iterations = m.splitIterations;
    s.anomalies = m.anomalies;
}

// The simplifications on a caller-built quad mesh (no grid): HalfedgeMesh::initialize,
// then simplify3X3Y / simplify3X3YOld / simplify3333 in the order runPostProcesses
// calls them (DualMarchingCubes.cu:335-380).  quads[4q+i] is corner i of quad q;
// attr[q] is its QuadrilateralAttribute word (colour in bits 0-4).
template <class R>
static Result<R>* runOnQuads(int nV, const float* pos, const float* nrm, int nQ, const int* quads, const uint16_t* attr, int flags) {
    auto* res = new Result<R>();
    Mesh<R>& m = res->mesh;
    m.pos.resize(nV);
    m.nrm.resize(nV);
    for (int v = 0; v < nV; ++v) {
        m.pos[v] = V3<R>{R(pos[3 * v]), R(pos[3 * v + 1]), R(pos[3 * v + 2])};
        m.nrm[v] = V3<R>{R(nrm[3 * v]), R(nrm[3 * v + 1]), R(nrm[3 * v + 2])};
    }
    m.quad.assign(quads, quads + 4 * (size_t)nQ);
    m.qattr.assign(attr, attr + nQ);
    m.dmcVertices = nV;
    m.dmcQuads = nQ;
    if (nV > 0 && nQ > 0) {
        halfedgeInitialize(m);
        if (flags & F_P3X3Y_COLOR) simplify3X3YColor(m);
        if (flags & F_P3X3Y_OLD) simplify3X3YOld(m);
        if (flags & F_P3333) simplify3333(m);
    }
    fillStats(m, flags & ~F_DETECT_NM_VERTS, res->st);
    return res;
}

}  // namespace orc

// --------------------------------------------------------------------------
// C interface for the tests (ctypes)
extern "C" {

typedef struct {
    int32_t flags;
    int32_t refinement;
    int32_t split_step;
    int32_t max_split_iters;
} orc_options;

struct orc_handle {
    int fp64;
    void* r;
};

// hand-built quad meshes through the simplifications (see runOnQuads)
void* orc_simplify_quads(int32_t nV, const float* pos, const float* nrm, int32_t nQ, const int32_t* quads, const uint16_t* attr, int32_t flags, orc::Stats* st) {
    auto* h = new orc_handle();
    h->fp64 = 0;
    auto* r = orc::runOnQuads<float>(nV, pos, nrm, nQ, quads, attr, flags);
    h->r = r;
    if (st) *st = r->st;
    return h;
}

// test hook: one planted mistake (orc::PLANT_*) or 0; returns the previous value
int32_t orc_set_plant(int32_t which) {
    const int old = orc::g_plant;
    orc::g_plant = which;
    return old;
}

void* orc_extract(const float* data, const int32_t dims[3], const float origin[3], const float delta[3], float iso, const orc_options* o, int32_t fp64, orc::Stats* st) {
    const int d[3] = {dims[0], dims[1], dims[2]};
    auto* h = new orc_handle();
    h->fp64 = fp64;
    if (fp64) {
        auto* r = orc::run<double>(data, d, origin, delta, iso, o->flags, o->refinement, o->split_step, o->max_split_iters);
        h->r = r;
        if (st) *st = r->st;
    } else {
        auto* r = orc::run<float>(data, d, origin, delta, iso, o->flags, o->refinement, o->split_step, o->max_split_iters);
        h->r = r;
        if (st) *st = r->st;
    }
    return h;
}

}  // extern "C"
template <class R>
static void getMesh(orc::Result<R>* r, float* pos, float* nrm, int32_t* he, int32_t* face_he, uint16_t* face_attr, uint8_t* vflags, int32_t* vhe) {
    auto& m = r->mesh;
    for (size_t i = 0; i < m.pos.size(); ++i) {
        if (pos) { pos[3 * i] = (float)m.pos[i].x; pos[3 * i + 1] = (float)m.pos[i].y; pos[3 * i + 2] = (float)m.pos[i].z; }
        if (nrm) { nrm[3 * i] = (float)m.nrm[i].x; nrm[3 * i + 1] = (float)m.nrm[i].y; nrm[3 * i + 2] = (float)m.nrm[i].z; }
    }
    for (size_t i = 0; i < m.he.size() && he; ++i) {
        he[4 * i] = m.he[i].o; he[4 * i + 1] = m.he[i].f; he[4 * i + 2] = m.he[i].n; he[4 * i + 3] = m.he[i].t;
    }
    for (size_t i = 0; i < m.face_he.size(); ++i) {
        if (face_he) face_he[i] = m.face_he[i];
        if (face_attr) face_attr[i] = m.fattr[i];
    }
    for (size_t i = 0; i < m.pos.size(); ++i) {
        if (vflags) vflags[i] = m.vflags.empty() ? 0 : m.vflags[i];
        if (vhe) vhe[i] = m.vert_he.empty() ? -1 : m.vert_he[i];
    }
}

extern "C" {
void orc_get_mesh(void* hv, float* pos, float* nrm, int32_t* he, int32_t* face_he, uint16_t* face_attr, uint8_t* vflags, int32_t* vhe) {
    auto* h = (orc_handle*)hv;
    if (h->fp64) getMesh((orc::Result<double>*)h->r, pos, nrm, he, face_he, face_attr, vflags, vhe);
    else getMesh((orc::Result<float>*)h->r, pos, nrm, he, face_he, face_attr, vflags, vhe);
}

void orc_get_positions_f64(void* hv, double* pos, double* nrm) {
    auto* h = (orc_handle*)hv;
    if (!h->fp64) return;
    auto& m = ((orc::Result<double>*)h->r)->mesh;
    for (size_t i = 0; i < m.pos.size(); ++i) {
        pos[3 * i] = m.pos[i].x; pos[3 * i + 1] = m.pos[i].y; pos[3 * i + 2] = m.pos[i].z;
        if (nrm) { nrm[3 * i] = m.nrm[i].x; nrm[3 * i + 1] = m.nrm[i].y; nrm[3 * i + 2] = m.nrm[i].z; }
    }
}

// the processed (padded / refined) grid of a result
void orc_get_grid(void* hv, int32_t dims[3], float origin[3], float delta[3], float* values) {
    auto* h = (orc_handle*)hv;
    const orc::Grid& g = h->fp64 ? ((orc::Result<double>*)h->r)->grid : ((orc::Result<float>*)h->r)->grid;
    dims[0] = g.nx; dims[1] = g.ny; dims[2] = g.nz;
    origin[0] = g.ox; origin[1] = g.oy; origin[2] = g.oz;
    delta[0] = g.dx; delta[1] = g.dy; delta[2] = g.dz;
    if (values) std::memcpy(values, g.v.data(), g.v.size() * sizeof(float));
}

void orc_free(void* hv) {
    auto* h = (orc_handle*)hv;
    if (!h) return;
    if (h->fp64) delete (orc::Result<double>*)h->r;
    else delete (orc::Result<float>*)h->r;
    delete h;
}

// one cell's polygon, for the table/decider cross-check: writes n, sizes[4], edges[12]
int32_t orc_cell_polygon(const float s[8], float iso, int32_t use_decider, int32_t sizes[4], int32_t edges[12]) {
    int mc = 0;
    for (int k = 0; k < 8; ++k) mc |= (s[k] >= iso ? 1 : 0) << k;
    orc::Polygon p = use_decider ? orc::polygonAsymptotic<float>(s, iso) : orc::cellPolygon<float>(mc, s, iso);
    int tot = 0;
    for (int c = 0; c < 4; ++c) { sizes[c] = p.size[c]; tot += p.size[c]; }
    for (int j = 0; j < 12; ++j) edges[j] = j < tot ? p.edge[j] : -1;
    return p.n;
}

}  // extern "C"
