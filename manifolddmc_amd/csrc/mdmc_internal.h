// mdmc_internal.h -- host-side declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace mdmc {

constexpr int kBlock = 256;

struct GridDev {
    const float* v;
    int nx, ny, nz;
    float ox, oy, oz;
    float dx, dy, dz;
};

// --- scans (mdmc_scan.hip): exclusive prefix of n int32 values into int64,
// returns the total through *total_dev (device int64).  scratch is sized by scanScratchBytes.
// The result is exact in every case; a tile whose look-back wait passed its spin limit
// summed its prefix directly and adds 1 to *directOut (when given; device memory), a
// diagnostic only.
size_t scanScratchBytes(int64_t n);
// total_dev may point into mapped pinned host memory
void scanExclusiveI32(const int32_t* in, int64_t* out, int64_t n, int64_t* total_dev, void* scratch, hipStream_t s, int64_t* directOut = nullptr);
// stream compaction in one pass: the i < n with (attr[i] & mask) != 0, ascending, into
// idx (capacity n); the count through *total_dev; scratch sized by scanScratchBytes(n)
void compactFlagged(const uint32_t* attr, int64_t n, uint32_t mask, int32_t* idx, int64_t* total_dev, void* scratch, hipStream_t s,
                    int64_t* directOut = nullptr);

inline unsigned gridFor(int64_t n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

// --- the reference's synthetic grids (mdmc_generate.hip): implicit surface sc (SurfaceCase
// numbering) at origin + index * delta, and the 4x4x3 edge case ec
void launchFillSurface(float* out, const int dims[3], int surfaceCase, const float origin[3], const float delta[3], hipStream_t s);
void launchFillEdgeCase(float* out, int edgeCase, hipStream_t s);

// --- face exports of OriginalDMC (mdmc_faces.hip): the 5-colouring of the faces' grid-edge
// colours (a, b: F bytes each; *result points at the one holding the colours), and the
// MaxMin-angle triangulation with per-triangle mean ratio (quality may be null)
void launchFaceColouring(const int4* he, const int32_t* face_he, const uint32_t* fattr, int64_t F, uint8_t* a, uint8_t* b, uint8_t** result,
                         hipStream_t s);
void launchTriangulateMaxMin(const int4* he, const int32_t* face_he, int64_t F, const float* pos, uint32_t* tri, float* quality, hipStream_t s);

// --- DMC grid pass (mdmc_dmc.hip)
// padding / refinement (refineGrid) of several grids in one launch: grid g reads in[inx*iny*inz]
// and writes out[onx*ony*onz] with the blocks from base on (prepareBlocks(dims) of them)
struct PrepareDesc {
    const float* in;
    float* out;
    int inx, iny, inz, onx, ony, onz;
    int64_t base;
};
// blocks of the padding / refinement pass over one processed grid of od points: one per row
int64_t prepareBlocks(const int od[3]);
// special: a device word that is zero on entry (set when an input value needs the exact blend)
void launchPrepareGrids(const PrepareDesc* descDev, int G, int64_t blocks, int refinement, int pad, int64_t* special, hipStream_t s);
void launchDmcTrianglesFromFaces(const int4* he, const int32_t* face_he, int64_t F, uint32_t* idx, hipStream_t s);
// render vertices {position, normal, rgba by vertex flags} of every vertex with an outgoing
// halfedge into out[10 * V] (the caller zero-fills out first)
void launchRenderVertices(const int4* he, int64_t H, const float* pos, const float* nrm, const uint32_t* vflags, float* out, hipStream_t s);

}  // namespace mdmc
