// mdmc_faces.hip -- face-level exports of OriginalDMC over the resident halfedge mesh.
//
// Restates (MarchingCubes/OriginalDMC/src):
//   FaceColoring::colorFaces, color_faces, optimize_face_coloring   FaceColoring.cu:5-141
//     Faces carry one of 24 colours from the edge colouring of the grid (3 * cell parity +
//     axis); colours 5..23 are replaced, one colour class after the other, by the first of
//     0..4 that none of the face's (up to) four neighbours has; then every face of colour 4
//     takes the first of 0..3 that its neighbours leave free.  The colour-4 test of
//     color_faces compares the third neighbour twice and never the fourth (FaceColoring.cu:
//     61); that is kept.  A class is read from the state before its launch and written to a
//     second buffer, so the result does not depend on thread order even where two faces of
//     one class touch (never the case in a mesh straight from the grid pass, whose colouring
//     is proper); faces whose colour is outside 5..23 (31 = none) keep it.
//   quadrilateral_to_triangle (its MaxMin-angle variant), Triangles::minAngle,
//   mean_ratio_measure, QualityMeasure::mean_ratio / area
//                                     DualMarchingCubes.cu:51-91, :93-107, Triangles.h:80-90,
//                                     QualityMeasure.h:37-70
//     Each quad (v0, v1, v2, v3) is cut along the diagonal whose two triangles have the larger
//     smallest angle (angles by the cosine rule): (v1, v3, v0) + (v1, v2, v3) when the v0-v2
//     cut's smallest angle is smaller, or equal with a second angle no larger; else (v0, v1,
//     v2) + (v0, v2, v3).  Every triangle gets the mean ratio 2 sqrt(3) |e1 x e2| / (l0 + l1 +
//     l2) (1 for an equilateral triangle, 0 for a degenerate one).
#include "mdmc_device.h"
#include "mdmc_internal.h"

namespace mdmc {

namespace {

__device__ __forceinline__ int neighbourFace(const int4* he, int h) {
    const int t = he[h].w;
    return t >= 0 ? he[t].y : -1;
}

__device__ __forceinline__ int colourOf(const uint8_t* col, int f) { return f >= 0 ? (int)col[f] : -1; }

// one colour class: faces of colour c (read from `in`) take the first free colour of 0..4
__global__ void kColourClass(const int4* __restrict__ he, const int32_t* __restrict__ face_he, int64_t F, int c,
                             const uint8_t* __restrict__ in, uint8_t* __restrict__ out) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const int own = in[f];
    if (own != c) {
        out[f] = (uint8_t)own;
        return;
    }
    const int e0 = face_he[f];
    const int e1 = he[e0].z, e2 = he[e1].z, e3 = he[e2].z;
    const int c0 = colourOf(in, neighbourFace(he, e0)), c1 = colourOf(in, neighbourFace(he, e1));
    const int c2 = colourOf(in, neighbourFace(he, e2)), c3 = colourOf(in, neighbourFace(he, e3));
    int nc = own;
    if (c0 != 0 && c1 != 0 && c2 != 0 && c3 != 0) nc = 0;
    else if (c0 != 1 && c1 != 1 && c2 != 1 && c3 != 1) nc = 1;
    else if (c0 != 2 && c1 != 2 && c2 != 2 && c3 != 2) nc = 2;
    else if (c0 != 3 && c1 != 3 && c2 != 3 && c3 != 3) nc = 3;
    else if (c0 != 4 && c1 != 4 && c2 != 4 && c2 != 4) nc = 4;  // the fourth neighbour is not tested
    out[f] = (uint8_t)nc;
}

// faces of colour 4 move to the first free colour of 0..3, if any
__global__ void kColourOptimize(const int4* __restrict__ he, const int32_t* __restrict__ face_he, int64_t F, const uint8_t* __restrict__ in,
                                uint8_t* __restrict__ out) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const int own = in[f];
    if (own != 4) {
        out[f] = (uint8_t)own;
        return;
    }
    const int e0 = face_he[f];
    const int e1 = he[e0].z, e2 = he[e1].z, e3 = he[e2].z;
    const int c0 = colourOf(in, neighbourFace(he, e0)), c1 = colourOf(in, neighbourFace(he, e1));
    const int c2 = colourOf(in, neighbourFace(he, e2)), c3 = colourOf(in, neighbourFace(he, e3));
    int nc = own;
    if (c0 != 0 && c1 != 0 && c2 != 0 && c3 != 0) nc = 0;
    else if (c0 != 1 && c1 != 1 && c2 != 1 && c3 != 1) nc = 1;
    else if (c0 != 2 && c1 != 2 && c2 != 2 && c3 != 2) nc = 2;
    else if (c0 != 3 && c1 != 3 && c2 != 3 && c3 != 3) nc = 3;
    out[f] = (uint8_t)nc;
}

__global__ void kFaceColourBits(const uint32_t* __restrict__ fattr, int64_t F, uint8_t* __restrict__ col) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f < F) col[f] = (uint8_t)(fattr[f] & QA_COLOR);
}

__device__ __forceinline__ f3 vtx(const float* pos, int v) { return mk3(pos[3 * (int64_t)v], pos[3 * (int64_t)v + 1], pos[3 * (int64_t)v + 2]); }

// Triangles::minAngle: the three angles by the cosine rule, the smallest
__device__ __forceinline__ float minAngle(f3 v0, f3 v1, f3 v2) {
    const float a = sqrtf((v1.x - v0.x) * (v1.x - v0.x) + (v1.y - v0.y) * (v1.y - v0.y) + (v1.z - v0.z) * (v1.z - v0.z));
    const float b = sqrtf((v2.x - v1.x) * (v2.x - v1.x) + (v2.y - v1.y) * (v2.y - v1.y) + (v2.z - v1.z) * (v2.z - v1.z));
    const float c = sqrtf((v0.x - v2.x) * (v0.x - v2.x) + (v0.y - v2.y) * (v0.y - v2.y) + (v0.z - v2.z) * (v0.z - v2.z));
    const float A = acosf((b * b + c * c - a * a) / (2 * b * c));
    const float B = acosf((a * a + c * c - b * b) / (2 * a * c));
    const float C = acosf((b * b + a * a - c * c) / (2 * b * a));
    return fminf(fminf(A, B), C);
}

// QualityMeasure::mean_ratio with area = |(v1 - v0) x (v2 - v0)|
__device__ __forceinline__ float meanRatio(f3 v0, f3 v1, f3 v2) {
    const float l0 = (v1.x - v0.x) * (v1.x - v0.x) + (v1.y - v0.y) * (v1.y - v0.y) + (v1.z - v0.z) * (v1.z - v0.z);
    const float l1 = (v2.x - v1.x) * (v2.x - v1.x) + (v2.y - v1.y) * (v2.y - v1.y) + (v2.z - v1.z) * (v2.z - v1.z);
    const float l2 = (v0.x - v2.x) * (v0.x - v2.x) + (v0.y - v2.y) * (v0.y - v2.y) + (v0.z - v2.z) * (v0.z - v2.z);
    const float l = l0 + l1 + l2;
    const float x1 = v1.x - v0.x, x2 = v2.x - v0.x;
    const float y1 = v1.y - v0.y, y2 = v2.y - v0.y;
    const float z1 = v1.z - v0.z, z2 = v2.z - v0.z;
    const float a1 = (y1 * z2 - z1 * y2) * (y1 * z2 - z1 * y2);
    const float a2 = (x1 * z2 - z1 * x2) * (x1 * z2 - z1 * x2);
    const float a3 = (x1 * y2 - y1 * x2) * (x1 * y2 - y1 * x2);
    const float area = sqrtf(a1 + a2 + a3);
    return 2.0f * sqrtf(3.0f) * area / l;
}

__global__ void kTriangulateMaxMin(const int4* __restrict__ he, const int32_t* __restrict__ face_he, int64_t F, const float* __restrict__ pos,
                                   uint32_t* __restrict__ tri, float* __restrict__ quality) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    int h = face_he[f];
    int v[4];
    for (int i = 0; i < 4; ++i) {
        const int4 e = he[h];
        v[i] = e.x;
        h = e.z;
    }
    const f3 p0 = vtx(pos, v[0]), p1 = vtx(pos, v[1]), p2 = vtx(pos, v[2]), p3 = vtx(pos, v[3]);
    float a1 = minAngle(p0, p1, p2), a2 = minAngle(p0, p2, p3);
    const float b1 = fminf(a1, a2), b2 = fmaxf(a1, a2);
    a1 = minAngle(p1, p3, p0);
    a2 = minAngle(p1, p2, p3);
    const float c1 = fminf(a1, a2), c2 = fmaxf(a1, a2);
    int t[6];
    if (b1 < c1 || (b1 == c1 && b2 <= c2)) {
        t[0] = v[1]; t[1] = v[3]; t[2] = v[0];
        t[3] = v[1]; t[4] = v[2]; t[5] = v[3];
    } else {
        t[0] = v[0]; t[1] = v[1]; t[2] = v[2];
        t[3] = v[0]; t[4] = v[2]; t[5] = v[3];
    }
    for (int k = 0; k < 6; ++k) tri[6 * f + k] = (uint32_t)t[k];
    if (quality) {
        quality[2 * f] = meanRatio(vtx(pos, t[0]), vtx(pos, t[1]), vtx(pos, t[2]));
        quality[2 * f + 1] = meanRatio(vtx(pos, t[3]), vtx(pos, t[4]), vtx(pos, t[5]));
    }
}

}  // namespace

void launchFaceColouring(const int4* he, const int32_t* face_he, const uint32_t* fattr, int64_t F, uint8_t* a, uint8_t* b, uint8_t** result,
                         hipStream_t s) {
    *result = a;
    if (F <= 0) return;
    kFaceColourBits<<<gridFor(F), kBlock, 0, s>>>(fattr, F, a);
    uint8_t* cur = a;
    uint8_t* nxt = b;
    for (int c = 5; c < 24; ++c) {
        kColourClass<<<gridFor(F), kBlock, 0, s>>>(he, face_he, F, c, cur, nxt);
        std::swap(cur, nxt);
    }
    kColourOptimize<<<gridFor(F), kBlock, 0, s>>>(he, face_he, F, cur, nxt);
    *result = nxt;
}

void launchTriangulateMaxMin(const int4* he, const int32_t* face_he, int64_t F, const float* pos, uint32_t* tri, float* quality, hipStream_t s) {
    if (F > 0) kTriangulateMaxMin<<<gridFor(F), kBlock, 0, s>>>(he, face_he, F, pos, tri, quality);
}

}  // namespace mdmc
