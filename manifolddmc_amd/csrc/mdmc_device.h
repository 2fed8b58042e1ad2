// mdmc_device.h -- device-side arithmetic shared by the gfx950 kernels.
//
// Every float expression here keeps the operation order of the reference
// (MarchingCubes/VisualDMC/src/DMC) and the library is built with
// -ffp-contract=off and IEEE division / square root, so results are the same
// bits a correctly rounded scalar evaluation of that order gives.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mc_tables.inc"

namespace mdmc {

// ---------------------------------------------------------------- vectors (glm semantics)
struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ f3 scl(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 divs(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
// glm::max / glm::min / glm::clamp (func_common.inl: (x<y)?y:x, (y<x)?y:x)
__device__ __forceinline__ float gmax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float gmin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
__device__ __forceinline__ f3 vclamp(f3 v, float lo, float hi) { return f3{gclamp(v.x, lo, hi), gclamp(v.y, lo, hi), gclamp(v.z, lo, hi)}; }
// glm::normalize = v * inversesqrt(dot(v,v)), dot = (x*x + y*y) + z*z, inversesqrt = 1/sqrt
__device__ __forceinline__ f3 normalize3(f3 v) {
    const float d = v.x * v.x + v.y * v.y + v.z * v.z;
    const float inv = 1.0f / sqrtf(d);  // sqrtf lowers to the correctly rounded sequence (__fsqrt_rn does not here)
    return f3{v.x * inv, v.y * inv, v.z * inv};
}
// glm::mix(x, y, a) = x*(1-a) + y*a
__device__ __forceinline__ float mixs(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ f3 mixv(f3 a, f3 b, float t) { return f3{mixs(a.x, b.x, t), mixs(a.y, b.y, t), mixs(a.z, b.z, t)}; }

// Interpolation<float>::interpolateTrilinearly, Interpolation.inl:48-56
__device__ __forceinline__ float trilinear(const float s[8], f3 u) {
    return (1.0f - u.z) * ((1.0f - u.y) * (s[0] + u.x * (s[1] - s[0])) + u.y * (s[2] + u.x * (s[3] - s[2]))) +
           u.z * ((1.0f - u.y) * (s[4] + u.x * (s[5] - s[4])) + u.y * (s[6] + u.x * (s[7] - s[6])));
}

// ---------------------------------------------------------------- cube tables
// corner k = (k&1, (k>>1)&1, (k>>2)&1); edge e runs from corner EA(e) to EB(e) along axis AX(e)
// (EdgeVertices{16,49,50,32,84,117,118,100,64,81,115,98}, CellIntersection.inl:383).
// Packed as nibbles so the lookups stay in scalar registers.
__device__ __forceinline__ int edgeA(int e) { return (int)((0x231046540210ULL >> (4 * e)) & 0xF); }
__device__ __forceinline__ int edgeB(int e) { return (int)((0x675467752331ULL >> (4 * e)) & 0xF); }
__device__ __forceinline__ int edgeAxis(int e) { return (int)((0x222210101010ULL >> (4 * e)) & 0xF); }
// quad slot for (edge, edgy bit): CellIntersection::computeVertexPosition, CellIntersection.inl:362-366
__device__ __forceinline__ int edgeSlot(int e, int edgy) { return (int)((240177437832960ULL >> (4 * e + 2 * edgy)) & 3ULL); }

static __constant__ uint64_t c_polygon[256] = {MDMC_POLYGON_TABLE_DATA};
static __constant__ uint64_t c_ambiguous[4] = {MDMC_AMBIGUOUS_MASK_DATA};

// A polygon is packed like the table words: nibbles 0-3 contour sizes, nibbles 4.. the edges.
__device__ __forceinline__ int polyCount(uint64_t w) {
    int n = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) n += (((w >> (4 * c)) & 0xF) != 0) ? 1 : 0;
    return n;
}
__device__ __forceinline__ int polySize(uint64_t w, int c) { return (int)((w >> (4 * c)) & 0xF); }
__device__ __forceinline__ int polyEdge(uint64_t w, int j) { return (int)((w >> (16 + 4 * j)) & 0xF); }

// The contour that holds edge e (or -1).
__device__ __forceinline__ int polyContourOf(uint64_t w, int e) {
    int j = 0;
    for (int c = 0; c < 4; ++c) {
        const int s = polySize(w, c);
        for (int i = 0; i < s; ++i, ++j)
            if (polyEdge(w, j) == e) return c;
    }
    return -1;
}

// Faces of the cell: corners and edges in the walk order of faceVertexLookup /
// faceEdgeLookup (CellIntersection.inl:290-292), one nibble each.
#define MDMC_FACE_CORNERS {0x3120u, 0x6475u, 0x1504u, 0x7362u, 0x2064u, 0x7531u}
#define MDMC_FACE_EDGES {0x0123u, 0x4765u, 0x4908u, 0x2A6Bu, 0x83B7u, 0x95A1u}

// Face case of face f in MC case mc: bit i = corner i of the face is inside (>= iso).
__host__ __device__ __forceinline__ int faceCase(int mc, int f) {
    const uint32_t fcorner[6] = MDMC_FACE_CORNERS;
    int fc = 0;
    for (int i = 0; i < 4; ++i) fc |= ((mc >> ((fcorner[f] >> (4 * i)) & 0xF)) & 1) << i;
    return fc;
}

// CellIntersection::marchingCubesAsymptoticPolygon + handleAsymptoticFaceCase,
// CellIntersection.inl:203-338, with the asymptotic decider's outcomes given:
// bit f of `dec` = (decider value >= iso) on face f, read only for the faces whose
// case is 6 or 9.  Links are kept as nibbles: out[e] in `outw`; contours are walked
// from the smallest edge that has a link.
__host__ __device__ __forceinline__ uint64_t polygonFromFaceDecisions(int mc, uint32_t dec) {
    const uint32_t fedge[6] = MDMC_FACE_EDGES;
    uint64_t outw = 0xFFFFFFFFFFFFULL;  // 12 nibbles of 0xF
    uint32_t set = 0;
#define MDMC_CONNECT(a, b)                                                        \
    do {                                                                          \
        outw = (outw & ~(0xFULL << (4 * (a)))) | ((uint64_t)(b) << (4 * (a)));    \
        set |= (1u << (a)) | (1u << (b));                                         \
    } while (0)
    for (int f = 0; f < 6; ++f) {
        int e[4];
        for (int i = 0; i < 4; ++i) e[i] = (int)((fedge[f] >> (4 * i)) & 0xF);
        const bool up = (dec >> f) & 1u;
        switch (faceCase(mc, f)) {
            case 1: MDMC_CONNECT(e[0], e[3]); break;
            case 2: MDMC_CONNECT(e[1], e[0]); break;
            case 3: MDMC_CONNECT(e[1], e[3]); break;
            case 4: MDMC_CONNECT(e[3], e[2]); break;
            case 5: MDMC_CONNECT(e[0], e[2]); break;
            case 7: MDMC_CONNECT(e[1], e[2]); break;
            case 8: MDMC_CONNECT(e[2], e[1]); break;
            case 10: MDMC_CONNECT(e[2], e[0]); break;
            case 11: MDMC_CONNECT(e[2], e[3]); break;
            case 12: MDMC_CONNECT(e[3], e[1]); break;
            case 13: MDMC_CONNECT(e[0], e[1]); break;
            case 14: MDMC_CONNECT(e[3], e[0]); break;
            case 6:
                if (up) { MDMC_CONNECT(e[3], e[0]); MDMC_CONNECT(e[1], e[2]); }
                else { MDMC_CONNECT(e[1], e[0]); MDMC_CONNECT(e[3], e[2]); }
                break;
            case 9:
                if (up) { MDMC_CONNECT(e[0], e[1]); MDMC_CONNECT(e[2], e[3]); }
                else { MDMC_CONNECT(e[0], e[3]); MDMC_CONNECT(e[2], e[1]); }
                break;
            default: break;
        }
    }
#undef MDMC_CONNECT
    uint64_t w = 0;
    int pos = 0, n = 0;
    for (int e = 0; e < 12 && n < 4; ++e) {
        if (!((set >> e) & 1u)) continue;
        int out = (int)((outw >> (4 * e)) & 0xF);
        int cnt = 0;
        w |= (uint64_t)e << (16 + 4 * (pos + cnt));
        ++cnt;
        int guard = 0;
        while (out != e && guard++ < 12 && out < 12 && pos + cnt < 12) {
            w |= (uint64_t)out << (16 + 4 * (pos + cnt));
            ++cnt;
            const int in = out;
            out = (int)((outw >> (4 * in)) & 0xF);
            set &= ~(1u << in);
        }
        w |= (uint64_t)cnt << (4 * n);
        pos += cnt;
        ++n;
    }
    return w;
}

// asymptotic decider of face f (CellIntersection.inl:203-210): (ad - bc) / (a + d - b - c) >= iso
template <int F>
__device__ __forceinline__ bool faceDecider(const float s[8], float iso) {
    constexpr uint32_t fc[6] = MDMC_FACE_CORNERS;
    const float v0 = s[fc[F] & 0xF], v1 = s[(fc[F] >> 4) & 0xF], v2 = s[(fc[F] >> 8) & 0xF], v3 = s[(fc[F] >> 12) & 0xF];
    const float d = (v0 * v3 - v1 * v2) / (v0 + v3 - v1 - v2);
    return d >= iso;
}

// the decider outcome of every face whose case is 6 or 9 (bit f), for a cell of case mc
__device__ __forceinline__ uint32_t faceDecisions(int mc, const float s[8], float iso) {
    uint32_t dec = 0;
#define MDMC_FACE(F)                                                  \
    {                                                                 \
        const int fc = faceCase(mc, F);                               \
        if ((fc == 6 || fc == 9) && faceDecider<F>(s, iso)) dec |= 1u << F; \
    }
    MDMC_FACE(0) MDMC_FACE(1) MDMC_FACE(2) MDMC_FACE(3) MDMC_FACE(4) MDMC_FACE(5)
#undef MDMC_FACE
    return dec;
}

__device__ __forceinline__ uint64_t polygonAsymptotic(const float s[8], float iso) {
    int mc = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) mc |= (s[k] >= iso ? 1 : 0) << k;
    return polygonFromFaceDecisions(mc, faceDecisions(mc, s, iso));
}

__device__ __forceinline__ bool caseAmbiguous(int c) { return (c_ambiguous[c >> 6] >> (c & 63)) & 1ULL; }

__device__ __forceinline__ uint64_t cellPolygon(int mc, const float s[8], float iso) {
    return caseAmbiguous(mc) ? polygonAsymptotic(s, iso) : c_polygon[mc];
}

// VertexToSurfaceIterator::movePointToSurface / updateGradient / computeGradient,
// CellIntersection.inl:12-122 (30 iterations, step 0.05, final clamp to [1e-4, 1-1e-4]).
__device__ __forceinline__ f3 moverGradient(const float* s, f3 u) {
    f3 g;
    g.x = (1 - u.z) * ((1 - u.y) * (s[1] - s[0]) + u.y * (s[3] - s[2])) + u.z * ((1 - u.y) * (s[5] - s[4]) + u.y * (s[7] - s[6]));
    g.y = (1 - u.z) * ((1 - u.x) * (s[2] - s[0]) + u.x * (s[3] - s[1])) + u.z * ((1 - u.x) * (s[6] - s[4]) + u.x * (s[7] - s[5]));
    g.z = (1 - u.y) * ((1 - u.x) * (s[4] - s[0]) + u.x * (s[5] - s[1])) + u.y * ((1 - u.x) * (s[6] - s[2]) + u.x * (s[7] - s[3]));
    return g;
}
__device__ __forceinline__ int moverUpdate(const float* s, float iso, f3& point, f3& g, f3 i1, f3& i2, float v1, float& v2) {
    int r = 0;
    g = normalize3(g);
    i2 = i1 + scl(0.05f, g);
    if (i2.x <= 0.f || i2.x >= 1.f) r |= 2;
    if (i2.y <= 0.f || i2.y >= 1.f) r |= 4;
    if (i2.z <= 0.f || i2.z >= 1.f) r |= 8;
    i2 = vclamp(i2, 0.f, 1.f);
    v2 = trilinear(s, i2);
    if ((v1 <= iso && iso <= v2) || (v2 <= iso && iso <= v1)) {
        float e = v1;
        if (v1 != v2) e = (iso - v1) / (v2 - v1);
        point = i1 + scl(e, i2 - i1);
        r |= 1;
    }
    return r;
}
// The iteration as a resumable state (moverInit, then moverStep until done), so that a
// kernel can run a few steps on every lane and hand the unfinished ones to other lanes.
struct Mover {
    f3 i1, point;
    float v1;
    int it;
    bool done;
};
__device__ __forceinline__ void moverInit(Mover& m, const float* s, f3 point) {
    m.i1 = point;
    m.point = point;
    m.v1 = trilinear(s, point);
    m.it = 0;
    m.done = false;
}
// one pass of the 30-step loop: the crossing found (done), or the next point
__device__ __forceinline__ void moverStep(Mover& m, const float* s, float iso) {
    f3 g = moverGradient(s, m.i1);
    if (m.v1 > iso) g = f3{-g.x, -g.y, -g.z};
    f3 i2 = m.i1;
    float v2 = 0.f;
    const int r = moverUpdate(s, iso, m.point, g, m.i1, i2, m.v1, v2);
    if (r & 1) {
        m.done = true;
        return;
    }
    if (r & 14) {
        if (r & 2) g.x = 0.f;
        else if (r & 4) g.y = 0.f;
        else g.z = 0.f;
        const int r2 = moverUpdate(s, iso, m.point, g, m.i1, i2, m.v1, v2);
        if (r2 & 1) {
            m.done = true;
            return;
        }
    }
    m.i1 = i2;
    m.v1 = v2;
    if (++m.it == 30) m.done = true;
}
__device__ __forceinline__ f3 moverResult(const Mover& m) { return vclamp(m.point, 0.0001f, 1.f - 0.0001f); }
__device__ __forceinline__ void movePointToSurface(const float* s, float iso, f3& point) {
    Mover m;
    moverInit(m, s, point);
    while (!m.done) moverStep(m, s, iso);
    point = moverResult(m);
}

// QuadrilateralAttribute bits (DeviceVectors.h:334-450) and vertex Flags bits (Halfedge.h:14-111).
// QA_DEAD has no reference counterpart: a quad removed by an earlier simplification of a chain
// whose compaction is deferred to the chain's last step (never in an output array).
enum : uint32_t { QA_COLOR = 0x1F, QA_NM = 0x20, QA_PATTERN = 0x40, QA_REMOVE = 0x80, QA_BOUNDARY = 0x100, QA_DEAD = 0x200 };
enum : uint32_t { VF_NM = 1, VF_BOUNDARY = 2, VF_WEIRD = 4, VF_FANS = 8, VF_REMOVE = 16 };

}  // namespace mdmc
