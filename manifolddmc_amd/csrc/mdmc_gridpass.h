// mdmc_gridpass.h -- launch interface of the single-kernel DMC grid pass.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mdmc_internal.h"

namespace mdmc {

enum : int { GP_ERR_OVERFLOW = 1, GP_ERR_TIMEOUT = 2, GP_ERR_RANGE = 4 };

struct GridPassArgs {
    GridDev g;
    float iso;
    int surfIter;
    int64_t C;
    int CX, CY;
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
    // per vertex its cell
    int32_t* adjMap;
    int32_t* adjCell;
};

int gridPassTile();
// ring size (power of two) for a grid of C cells, CX x CY per z-slice
void gridPassRing(int64_t C, int CX, int CY, int64_t* R, int* shift);
void launchGridPass(const GridPassArgs& a, hipStream_t s);
void launchGridPassTotals(const unsigned long long* status, int64_t numTiles, int64_t* out, hipStream_t s);

}  // namespace mdmc
