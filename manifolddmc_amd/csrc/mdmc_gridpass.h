// mdmc_gridpass.h -- launch interface of the DMC grid pass (classify, scan, emit).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mdmc_internal.h"

namespace mdmc {

// Shapes of the grid pass.  Every shape gives the same mesh; the autotuner picks the fastest
// per workload class.
//   1..3: tiles of 1024 / 2048 / 4096 consecutive cells, dealt to the 8 XCDs in contiguous
//         chunks (classify, scan, emit)
//   4..6: the same tiles in launch order (block b -> tile b)
//   7, 8: row slabs of 4 / 2 rows: a block owns a band of up to 254 cells of its rows and
//         marches through a chunk of slices (count, scan over row segments, emit)
enum : int { GP_SHAPE_AUTO = 0, GP_SHAPE_FIRST = 1, GP_SHAPE_LAST_TILE = 6, GP_SHAPE_SLAB4 = 7, GP_SHAPE_SLAB2 = 8, GP_SHAPE_QUEUE = 9,
             GP_SHAPE_LAST = 9, GP_SHAPE_DEFAULT = 2 };
int gridPassTile(int shape);
inline bool gridPassIsSlab(int shape) { return shape == GP_SHAPE_SLAB4 || shape == GP_SHAPE_SLAB2; }
// 2048-cell tiles taken from a work queue by persistent blocks
inline bool gridPassIsQueue(int shape) { return shape == GP_SHAPE_QUEUE; }
inline int gridPassSlabRows(int shape) { return shape == GP_SHAPE_SLAB4 ? 4 : 2; }
constexpr int SLAB_COLS = 256;            // threads of a slab block, one LDS column each
constexpr int SLAB_XMAX = SLAB_COLS - 2;  // cells of a band row (plus a halo cell and the x + 1 points)
const char* gridPassShapeName(int shape);

// One (grid, iso-value) problem of an extraction.  A batch concatenates the problems'
// cells (cellBase) and tiles (tileBase) in order; a single extraction is a batch of one.
struct ProblemDev {
    GridDev g;  // the processed (padded / refined) grid
    float iso;
    int CX, CY;        // cells per row / per slice
    int64_t cells;     // CX * CY * CZ
    int64_t cellBase;  // first cell in the concatenation
    int64_t tileBase;  // first grid-pass tile
    // row slabs: bands of bandW cells per row, rowBlocks blocks of rows, zChunks chunks of zChunk
    // slices; the problem's blocks start at blockBase, its row segments (slice, row, band) at
    // segBase, segment (z, y, b) = segBase + (z * CY + y) * bands + b
    int bands, bandW, rowBlocks, zChunk, zChunks;
    int64_t blockBase, segBase;
};

struct GridPassArgs {
    const ProblemDev* probs;
    int nProblems;
    int surfIter;
    int64_t numTiles;
    // the work-queue counters of a queue-shape pass (classify, emit; zeroed by its classify launch)
    int64_t* queue;
    // classify -> emit: per tile its active cells' 32-bit records (polygon table entry, cell in
    // the tile, owned edges, contour count) at t * tile + active index; per tile its vertex /
    // quad / active-cell counts
    uint32_t* rec;
    int32_t* tileV;   // [numTiles + 1], last entry 0 (scanned in place of a total)
    int32_t* tileQ;   // [numTiles + 1]
    int32_t* tileNA;  // [numTiles]
    // per active cell (concatenated cell index): vertex index inside its tile (14 bits) and the
    // contour of each edge a later cell's quads read (2 bits each for edges 1,2,4,5,6,7,9,10,11)
    uint32_t* cellMap;
    // exclusive prefix of tileV / tileQ: each tile's first vertex / quad
    const int64_t* tileVOff;
    const int64_t* tileQOff;
    float* pos;
    float* nrm;
    int4* quads;
    uint32_t* qattr;
    // per vertex: its cell inside its problem and its contour (emit -> vertices): the cell's
    // polygon table entry | contour index << 10
    int32_t* vcell;
    uint16_t* vct;
    int64_t V;
    // per vertex its out-degree in the halfedge mesh (interior edges of its contour), when not null
    int32_t* outdeg;
    // recorded for the adjoint when not null: per cell (first vertex << 2 | contours - 1) or -1
    // (single-problem extractions only)
    int32_t* adjMap;
    // per vertex its problem, when not null
    int32_t* vpid;
    // row slabs: blocks of the launch, vertex / quad count per row segment ([S + 1], last 0)
    // and their exclusive prefixes
    int64_t numBlocks, numSegs;
    int32_t* segV;
    int32_t* segQ;
    const int64_t* segVOff;
    const int64_t* segQOff;
};

// uploads the polygon table of the classify kernel; once per device, with that device current
void initGridPassTables();
void launchGridClassify(int shape, const GridPassArgs& a, hipStream_t s);
// emit (quads, vertex cells) and the dual vertices
void launchGridEmit(int shape, const GridPassArgs& a, hipStream_t s);
// row slabs: per row segment vertex / quad counts; then (after the scan of the counts) quads,
// vertex work items and the dual vertices
void launchSlabCount(int shape, const GridPassArgs& a, hipStream_t s);
void launchSlabEmit(int shape, const GridPassArgs& a, hipStream_t s);
// the halfedge initialize's outgoing lists from the grid pass's quads: entry off[v] + rank (the
// rank from v's contour) of halfedge 4q + i from corner v = i holds 4q + i (out) and corners
// i + 1 and i - 1 (ends); also the face -> halfedge array and the face attributes
void launchQuadLists(const int4* quads, const uint32_t* qattr, int64_t Q, int64_t V, const int32_t* vcell, const uint16_t* vct, const int32_t* vpid,
                     const ProblemDev* probs, const int64_t* off, int32_t* out, int2* ends, int32_t* face_he, uint32_t* fattr, int64_t* invalid,
                     hipStream_t s);

}  // namespace mdmc
