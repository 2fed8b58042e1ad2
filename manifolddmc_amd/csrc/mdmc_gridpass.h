// mdmc_gridpass.h -- launch interface of the single-kernel DMC grid pass.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mdmc_internal.h"

namespace mdmc {

enum : int { GP_ERR_OVERFLOW = 1, GP_ERR_TIMEOUT = 2 };

struct GridPassArgs {
    GridDev g;
    float iso;
    int surfIter;
    int64_t C;
    int CX, CY;
    int64_t numTiles;
    unsigned long long* status;  // per tile: flag | quads | vertices (look-back)
    int* ready;                  // per tile: ring entries published
    int* done;                   // per tile: finished reading the ring
    unsigned int* ticket;        // tile claim counter
    int2* ring;                  // (first vertex id, edge->contour map) of cell c at c % R
    int64_t R;
    float* pos;
    float* nrm;
    int64_t capV;
    int4* quads;
    uint32_t* qattr;
    int64_t capQ;
    int* err;
};

int gridPassTile();
void launchGridPass(const GridPassArgs& a, hipStream_t s);
void launchGridPassTotals(const unsigned long long* status, int64_t numTiles, int64_t* out, hipStream_t s);

}  // namespace mdmc
