// mdmc_generate.hip -- the reference's synthetic grids, generated on the device.
//
// Restates (MarchingCubes/VisualDMC/src/DMC):
//   implicit surfaces          UniformGrid.inl:25-127 (sphere, tori, genus two, iWP, pwHybrid,
//                              neovius, Goursat, Steiner's Roman surface)
//   fillVolume / generateVolume       UniformGrid.inl:137-189, :463-486: dims points on
//                              [-1, 1]^3, delta = 2 / (dims - 1), origin -1, value at
//                              origin + index * delta
//   fillEdgeCase / generateEdgeCase   UniformGrid.inl:191-269, :489-509: two 4x4x3 grids of
//                              850 / 950 values with a non-manifold vertex at iso 900
// Each point is one thread writing one float, x fastest.
#include "mdmc_device.h"
#include "mdmc_internal.h"

namespace mdmc {

namespace {

// LazyEngine::Constants::PI is a double: a float times it is evaluated in double and rounded
// when stored into the float coordinate (UniformGrid.inl:61-63, :71-73, :81-83)
constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ float sq(float x) { return x * x; }
__device__ __forceinline__ float qu(float x) { return x * x * x * x; }

__device__ __forceinline__ float torusH(f3 p, f3 c, float px, float py) {
    const float cc = sq(px), a = sq(py);
    const float x = sq(p.x - c.x), y = sq(p.y - c.y), z = sq(p.z - c.z);
    return sq(x + y + z + cc - a) - 4 * cc * (x + y);
}
__device__ __forceinline__ float torusV(f3 p, f3 c, float px, float py) {
    const float cc = sq(px), a = sq(py);
    const float x = sq(p.x - c.x), y = sq(p.y - c.y), z = sq(p.z - c.z);
    return sq(x + y + z + cc - a) - 4 * cc * (x + z);
}

__device__ __forceinline__ float surfaceValue(int sc, f3 p) {
    switch (sc) {
        case 0:  // sphere of radius 0.4
            return p.x * p.x + p.y * p.y + p.z * p.z - 0.16f;
        case 1:  // torus, radii 0.3 / 0.15 around the z axis
            return torusH(p, mk3(0.f, 0.f, 0.f), 0.3f, 0.15f);
        case 2: {  // two-holed torus
            const float v1 = torusH(p, mk3(-0.38f, 0.f, 0.f), 0.3f, 0.15f);
            const float v2 = torusH(p, mk3(0.2f, 0.38f, 0.f), 0.3f, 0.15f);
            return fminf(v1, v2);
        }
        case 3: {  // four-holed torus
            const float v1 = torusH(p, mk3(-0.38f, 0.f, 0.f), 0.3f, 0.15f);
            const float v2 = torusH(p, mk3(0.38f, 0.f, 0.f), 0.3f, 0.15f);
            const float v3 = torusV(p, mk3(0.f, 0.f, -0.38f), 0.3f, 0.15f);
            const float v4 = torusV(p, mk3(0.f, 0.f, 0.38f), 0.3f, 0.15f);
            return fminf(fminf(fminf(v1, v2), v3), v4);
        }
        case 5: {  // iWP, alpha 5.01
            const float al = 5.01f;
            const float x = (float)((double)(al * (p.x + 1)) * kPi), y = (float)((double)(al * (p.y + 1)) * kPi),
                        z = (float)((double)(al * (p.z + 1)) * kPi);
            return cosf(x) * cosf(y) + cosf(y) * cosf(z) + cosf(z) * cosf(x) - cosf(x) * cosf(y) * cosf(z);
        }
        case 6: {  // pwHybrid, alpha 1.01
            const float al = 1.01f;
            const float x = (float)((double)(al * (p.x + 1)) * kPi), y = (float)((double)(al * (p.y + 1)) * kPi),
                        z = (float)((double)(al * (p.z + 1)) * kPi);
            return 4.0f * (cosf(x) * cosf(y) + cosf(y) * cosf(z) + cosf(z) * cosf(x)) - 3 * cosf(x) * cosf(y) * cosf(z) + 0.8f;
        }
        case 7: {  // neovius, alpha 1
            const float al = 1.f;
            const float x = (float)((double)(al * (p.x + 1)) * kPi), y = (float)((double)(al * (p.y + 1)) * kPi),
                        z = (float)((double)(al * (p.z + 1)) * kPi);
            return 3 * (cosf(x) + cosf(y) + cosf(z)) + 4 * cosf(x) * cosf(y) * cosf(z);
        }
        case 8:  // Goursat surface, a = -1, b = 0, c = 0.5
            return qu(p.x) + qu(p.y) + qu(p.z) + -1.0f * (sq(p.x) + sq(p.y) + sq(p.z)) + 0.0f * (sq(p.x) + sq(p.y) + sq(p.z)) + 0.5f;
        case 9: {  // Steiner's Roman surface, scaled by 1.5
            const float x = 1.5f * p.x, y = 1.5f * p.y, z = 1.5f * p.z;
            return sq(x * x + y * y + z * z - 1.0f) - (sq(z - 1) - 2.0f * x * x) * (sq(z + 1) - 2 * y * y);
        }
        default: {  // 4 and anything else: genus two
            float x = (p.x + 1.0f) / 2.0f, y = (p.y + 1.0f) / 2.0f, z = (p.z + 1.0f) / 2.0f;
            x = 1.0f * (4.0f * x - 2.0f);
            y = 1.0f * (4.0f * y - 2.0f);
            z = 1.0f * (4.0f * z - 2.0f);
            float t = 2 * y * (y * y - 3 * x * x) * (1 - z * z);
            t += (x * x + y * y) * (x * x + y * y);
            t -= (9 * z * z - 1) * (1 - z * z);
            return t;
        }
    }
}

__global__ void kFillSurface(float* __restrict__ out, int nx, int ny, int nz, int sc, f3 origin, f3 delta) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)nx * ny * nz;
    if (i >= n) return;
    const int x = (int)(i % nx);
    const int64_t r = i / nx;
    const int y = (int)(r % ny), z = (int)(r / ny);
    const f3 p = origin + mk3((float)x, (float)y, (float)z) * delta;
    out[i] = surfaceValue(sc, p);
}

// the two 4x4x3 edge cases, [z][y][x]: 950 at the marked points, 850 elsewhere (bit
// z * 16 + y * 4 + x of `high` marks point (x, y, z))
__global__ void kFillEdgeCase(float* __restrict__ out, int ec) {
    const int i = threadIdx.x;
    if (i >= 48) return;
    // case 0: (z0, y1, x2), (z0, y2, x1); case 1 adds (z1, y2, x1)
    const uint64_t high = ec == 0 ? ((1ULL << 6) | (1ULL << 9)) : ((1ULL << 6) | (1ULL << 9) | (1ULL << 25));
    out[i] = ((high >> i) & 1ULL) ? 950.f : 850.f;
}

}  // namespace

void launchFillSurface(float* out, const int dims[3], int surfaceCase, const float origin[3], const float delta[3], hipStream_t s) {
    const int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
    if (n > 0)
        kFillSurface<<<gridFor(n), kBlock, 0, s>>>(out, dims[0], dims[1], dims[2], surfaceCase, f3{origin[0], origin[1], origin[2]},
                                                   f3{delta[0], delta[1], delta[2]});
}

void launchFillEdgeCase(float* out, int edgeCase, hipStream_t s) { kFillEdgeCase<<<1, 64, 0, s>>>(out, edgeCase); }

}  // namespace mdmc
