// mdmc_gridpass.h -- launch interface of the single-kernel DMC grid pass.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mdmc_internal.h"

namespace mdmc {

enum : int { GP_ERR_OVERFLOW = 1, GP_ERR_TIMEOUT = 2, GP_ERR_RANGE = 4 };

// One (grid, iso-value) problem of an extraction.  A batch concatenates the problems'
// cells (cellBase) and tiles (tileBase) in order; a single extraction is a batch of one.
struct ProblemDev {
    GridDev g;  // the processed (padded / refined) grid
    float iso;
    int CX, CY;        // cells per row / per slice
    int64_t cells;     // CX * CY * CZ
    int64_t cellBase;  // first cell in the concatenation
    int64_t tileBase;  // first grid-pass tile
};

struct GridPassArgs {
    const ProblemDev* probs;
    int nProblems;
    int surfIter;
    int64_t totalCells;
    int64_t reach;  // max over problems of CX * CY + CX: how far back a quad's cells lie
    int64_t numTiles;
    unsigned long long* status;  // per tile: flag | quads | vertices (look-back)
    int* done;                   // per tile: finished reading the ring
    unsigned int* ticket;        // tile claim counter
    unsigned long long* ring;    // tag | edge->contour map | first vertex id of cell c at c & (R-1)
    int64_t R;                   // power of two, 1 << ringShift
    int ringShift;
    float* pos;
    float* nrm;
    int64_t capV;
    int4* quads;
    uint32_t* qattr;
    int64_t capQ;
    int* err;
    // recorded for the adjoint when not null: per cell (first vertex << 2 | contours - 1) or -1,
    // per vertex its cell (single-problem extractions only)
    int32_t* adjMap;
    int32_t* adjCell;
    // per vertex its problem, when not null
    int32_t* vpid;
};

int gridPassTile();
// ring size (power of two) for totalCells cells whose quads reach back `reach` cells
void gridPassRing(int64_t totalCells, int64_t reach, int64_t* R, int* shift);
void launchGridPass(const GridPassArgs& a, hipStream_t s);
void launchGridPassTotals(const unsigned long long* status, int64_t numTiles, int64_t* out, hipStream_t s);

}  // namespace mdmc
