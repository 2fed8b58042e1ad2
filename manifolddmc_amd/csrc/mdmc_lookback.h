// mdmc_lookback.h -- single-pass ordered compaction: a block selects elements of its tile of
// R * kLbBlock consecutive elements (element base + r * kLbBlock + threadIdx.x in round r) and
// learns the tile's first output slot by decoupled look-back over the tiles before it (the
// scheme of the prefix sum in mdmc_scan.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdmc {

constexpr int kLbBlock = 256;
constexpr int kLbSpinLimit = 1 << 22;

// look-back words (one per tile, zero before the launch): flag (2 bits: 1 aggregate, 2 inclusive)
// and value (62 bits), read and written with relaxed device-scope atomics
__device__ __forceinline__ uint64_t lbLoad(const unsigned long long* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void lbStore(unsigned long long* p, uint64_t v) {
    __hip_atomic_store(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Run by the block's first wave: publishes the tile's count, walks back over the earlier
// tiles' words and returns (through *s_off, read after a barrier) the number of selected
// elements before the tile.  A wait past the spin limit does not trust the dispatch order:
// countBefore(lane) (the selected elements among base .. in steps of 64 from lane) gives the
// exact prefix instead, and *directOut (device memory) counts the tile as a diagnostic.
template <class CountBefore>
__device__ __forceinline__ void lbTileOffset(unsigned long long* status, int64_t t, int64_t tileSum, CountBefore countBefore, int64_t* directOut,
                                             long long* s_off) {
    constexpr uint64_t FLAG_AGG = 1ULL, FLAG_INC = 2ULL, VAL_MASK = (1ULL << 62) - 1;
    const int lane = threadIdx.x & 63;
    if (t == 0) {
        if (lane == 0) {
            lbStore(&status[0], (FLAG_INC << 62) | ((uint64_t)tileSum & VAL_MASK));
            *s_off = 0;
        }
        return;
    }
    if (lane == 0) lbStore(&status[t], (FLAG_AGG << 62) | ((uint64_t)tileSum & VAL_MASK));
    int64_t accum = 0;
    int64_t i = t - 1;
    int spins = 0;
    bool done = false, direct = false;
    while (!done) {
        const int64_t j = i - lane;
        const uint64_t st = j >= 0 ? lbLoad(&status[j]) : (FLAG_INC << 62);
        const uint64_t f = st >> 62;
        const uint64_t incMask = __ballot(f == FLAG_INC);
        const uint64_t zeroMask = __ballot(f == 0);
        const int firstInc = incMask ? __ffsll((unsigned long long)incMask) - 1 : 64;
        const uint64_t upto = firstInc >= 63 ? ~0ULL : ((1ULL << (firstInc + 1)) - 1ULL);
        if (zeroMask & upto) {
            // the ballots make this branch and the counter uniform over the wave
            if (++spins > kLbSpinLimit) {
                direct = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        int64_t val = lane <= firstInc ? (int64_t)(st & VAL_MASK) : 0;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) val += __shfl_down(val, d, 64);
        accum += __shfl(val, 0, 64);
        if (firstInc < 64) done = true;
        else i -= 64;
    }
    if (direct) {
        int64_t sum = countBefore(lane);
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) sum += __shfl_down(sum, d, 64);
        accum = __shfl(sum, 0, 64);
        if (lane == 0 && directOut) atomicAdd((unsigned long long*)directOut, 1ULL);
    }
    if (lane == 0) {
        lbStore(&status[t], (FLAG_INC << 62) | ((uint64_t)(accum + tileSum) & VAL_MASK));
        *s_off = accum;
    }
}

// Per round r the wave's selection ballot m[r]; on return s_pre[r * 4 + wave] holds the
// number of selected elements of the tile before that wave's 64 elements of round r (index
// order: round, then wave), and the tile's total is returned.  Call from every thread;
// s_cnt holds R * 4 ints, s_pre R * 4 + 1.
template <int R>
__device__ __forceinline__ int lbBlockRanks(const uint64_t m[R], int* s_cnt, int* s_pre) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int r = 0; r < R; ++r) s_cnt[r * (kLbBlock / 64) + wid] = __popcll(m[r]);
    __syncthreads();
    constexpr int N = R * (kLbBlock / 64);  // at most 64 counts, one wave scans them
    static_assert(N <= 64, "one wave scans the per-round wave counts");
    if (threadIdx.x < 64) {
        const int v = threadIdx.x < N ? s_cnt[threadIdx.x] : 0;
        int x = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (threadIdx.x < N) s_pre[threadIdx.x] = x - v;
        if (threadIdx.x == N - 1) s_pre[N] = x;
    }
    __syncthreads();
    return s_pre[N];
}

}  // namespace mdmc
