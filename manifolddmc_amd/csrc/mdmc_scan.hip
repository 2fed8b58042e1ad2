// mdmc_scan.hip -- device-wide exclusive prefix sum of int32 counts into int64 offsets,
// in one pass: block b scans tile b of 4096 counts and gets its offset by decoupled
// look-back over the tiles before it.
//
// Tile = block index, not an atomic ticket: the blocks of a launch go to the XCDs
// round-robin and each XCD starts its blocks in increasing order, so every tile below the
// lowest resident one has finished (a lower tile not yet started would sit on an XCD whose
// resident blocks are all lower still), and the lowest resident tile never waits on an
// unstarted one.  The ticket cost one same-address atomic per block: 100 us against 67 us
// for 17 M counts.  That dispatch order is observed behaviour, not a documented
// guarantee, so correctness does not rest on it: a tile whose look-back wait passes its
// spin limit sums every count before it straight from the input instead (exact, only
// slower), then publishes as usual, so a tile that waits on an unstarted one still ends
// and frees its slot.
//
// Look-back words pack flag (2 bits) and value (62 bits), so relaxed device-scope
// atomics suffice: a reader needs nothing beyond the word it reads.
#include "mdmc_internal.h"

namespace mdmc {

namespace {
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kTile = kScanBlock * kScanItems;
constexpr int kSpinLimit = 1 << 22;

constexpr uint64_t FLAG_AGG = 1ULL, FLAG_INC = 2ULL;
constexpr uint64_t VAL_MASK = (1ULL << 62) - 1;

__device__ __forceinline__ uint64_t loadRelaxed(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void storeRelaxed(unsigned long long* p, uint64_t v) {
    __hip_atomic_store(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// scratch layout: [0], [1] unused, [2 ..] one look-back word per tile
__global__ void kScanReset(unsigned long long* scratch, int64_t words) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < words) scratch[i] = 0;
}

__global__ void __launch_bounds__(kScanBlock) kScan(const int32_t* __restrict__ in, int64_t* __restrict__ out, int64_t n, int64_t numTiles,
                                                    unsigned long long* __restrict__ scratch, int64_t* __restrict__ total,
                                                    int64_t* __restrict__ directOut) {
    __shared__ __align__(16) int32_t s_in[kTile];
    __shared__ int64_t s_wave[kScanBlock / 64];
    __shared__ long long s_off;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int64_t t = blockIdx.x;
    if (t >= numTiles) return;
    unsigned long long* status = scratch + 2;
    const int64_t base = t * kTile;

    // 16 consecutive counts per thread: four 16-byte loads each when the tile is full and
    // the input 16-byte aligned, else a coalesced load through LDS
    int32_t v[kScanItems];
    const bool vec = base + kTile <= n && (((uintptr_t)in) & 15) == 0;
    if (vec) {
        const int4* p = reinterpret_cast<const int4*>(in + base) + tid * (kScanItems / 4);
#pragma unroll
        for (int k = 0; k < kScanItems / 4; ++k) {
            const int4 q = p[k];
            v[4 * k] = q.x;
            v[4 * k + 1] = q.y;
            v[4 * k + 2] = q.z;
            v[4 * k + 3] = q.w;
        }
    } else {
        for (int i = tid; i < kTile; i += kScanBlock) {
            const int64_t j = base + i;
            s_in[i] = j < n ? in[j] : 0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) v[k] = s_in[tid * kScanItems + k];
    }
    int64_t acc = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) acc += v[k];
    int64_t x = acc;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_wave[wid] = x;
    __syncthreads();
    int64_t add = 0, tileSum = 0;
#pragma unroll
    for (int w = 0; w < kScanBlock / 64; ++w) {
        const int64_t s = s_wave[w];
        if (w < wid) add += s;
        tileSum += s;
    }
    int64_t run = x - acc + add;  // exclusive prefix of this thread inside the tile

    // look-back (wave 0)
    if (tid < 64) {
        if (t == 0) {
            if (lane == 0) {
                storeRelaxed(&status[0], (FLAG_INC << 62) | ((uint64_t)tileSum & VAL_MASK));
                s_off = 0;
            }
        } else {
            if (lane == 0) storeRelaxed(&status[t], (FLAG_AGG << 62) | ((uint64_t)tileSum & VAL_MASK));
            int64_t accum = 0;
            int64_t i = t - 1;
            int spins = 0;
            bool done = false, direct = false;
            while (!done) {
                const int64_t j = i - lane;
                const uint64_t st = j >= 0 ? loadRelaxed(&status[j]) : (FLAG_INC << 62);
                const uint64_t f = st >> 62;
                const uint64_t incMask = __ballot(f == FLAG_INC);
                const uint64_t zeroMask = __ballot(f == 0);
                const int firstInc = incMask ? __ffsll((unsigned long long)incMask) - 1 : 64;
                const uint64_t upto = firstInc >= 63 ? ~0ULL : ((1ULL << (firstInc + 1)) - 1ULL);
                if (zeroMask & upto) {
                    // the ballots make this branch and the counter uniform over the wave
                    if (++spins > kSpinLimit) {
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
                // the exclusive prefix straight from the counts before this tile
                int64_t sum = 0;
                for (int64_t k = lane; k < base; k += 64) sum += in[k];
#pragma unroll
                for (int d = 32; d > 0; d >>= 1) sum += __shfl_down(sum, d, 64);
                accum = __shfl(sum, 0, 64);
                if (lane == 0 && directOut) atomicAdd((unsigned long long*)directOut, 1ULL);
            }
            if (lane == 0) {
                storeRelaxed(&status[t], (FLAG_INC << 62) | ((uint64_t)(accum + tileSum) & VAL_MASK));
                s_off = accum;
            }
        }
    }
    __syncthreads();
    run += s_off;

    // exclusive offsets back through LDS, half a tile at a time, for coalesced int64 stores
    int64_t* s_out = reinterpret_cast<int64_t*>(s_in);  // 4096 int32 slots hold 2048 int64
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        __syncthreads();
        if ((tid >> 7) == half) {
            int64_t r = run;
#pragma unroll
            for (int k = 0; k < kScanItems; ++k) {
                s_out[(tid & 127) * kScanItems + k] = r;
                r += v[k];
            }
        }
        __syncthreads();
        const int64_t o = base + (int64_t)half * (kTile / 2);
        if (vec && (((uintptr_t)out) & 15) == 0) {
            // pairs of offsets as 16-byte stores
            const longlong2* s2 = reinterpret_cast<const longlong2*>(s_out);
            longlong2* o2 = reinterpret_cast<longlong2*>(out + o);
            for (int i = tid; i < kTile / 4; i += kScanBlock) o2[i] = s2[i];
        } else {
            for (int i = tid; i < kTile / 2; i += kScanBlock)
                if (o + i < n) out[o + i] = s_out[i];
        }
    }
    if (t == numTiles - 1 && tid == 0) *total = s_off + tileSum;
}
// Stream compaction in the same single pass: the indices i < n with (attr[i] & mask) != 0, in
// increasing order, into idx[0 .. total).  Each tile counts its selected elements (16 per
// thread), scans the counts in the block, gets the tile's first output slot by the same
// look-back as kScan (and the same exact fallback: a direct count of the selected elements
// before the tile) and writes its indices; nothing is written per unselected element.
__global__ void __launch_bounds__(kScanBlock) kCompactFlagged(const uint32_t* __restrict__ attr, int64_t n, uint32_t mask, int32_t* __restrict__ idx,
                                                              int64_t numTiles, unsigned long long* __restrict__ scratch,
                                                              int64_t* __restrict__ total, int64_t* __restrict__ directOut) {
    __shared__ int s_wave[kScanBlock / 64];
    __shared__ long long s_off;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int64_t t = blockIdx.x;
    if (t >= numTiles) return;
    unsigned long long* status = scratch + 2;
    const int64_t base = t * kTile;
    const int64_t first = base + (int64_t)tid * kScanItems;  // this thread's 16 consecutive elements

    uint32_t bits = 0;
    if (first + kScanItems <= n && (((uintptr_t)attr) & 15) == 0) {
        const uint4* p = reinterpret_cast<const uint4*>(attr + first);
#pragma unroll
        for (int k = 0; k < kScanItems / 4; ++k) {
            const uint4 q = p[k];
            bits |= ((q.x & mask) ? 1u : 0u) << (4 * k);
            bits |= ((q.y & mask) ? 1u : 0u) << (4 * k + 1);
            bits |= ((q.z & mask) ? 1u : 0u) << (4 * k + 2);
            bits |= ((q.w & mask) ? 1u : 0u) << (4 * k + 3);
        }
    } else {
        for (int k = 0; k < kScanItems; ++k)
            if (first + k < n && (attr[first + k] & mask)) bits |= 1u << k;
    }
    const int cnt = __popc(bits);
    int x = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_wave[wid] = x;
    __syncthreads();
    int add = 0, tileSum = 0;
#pragma unroll
    for (int w = 0; w < kScanBlock / 64; ++w) {
        const int sw = s_wave[w];
        if (w < wid) add += sw;
        tileSum += sw;
    }
    const int run = x - cnt + add;  // exclusive prefix of this thread inside the tile

    if (tid < 64) {
        if (t == 0) {
            if (lane == 0) {
                storeRelaxed(&status[0], (FLAG_INC << 62) | ((uint64_t)tileSum & VAL_MASK));
                s_off = 0;
            }
        } else {
            if (lane == 0) storeRelaxed(&status[t], (FLAG_AGG << 62) | ((uint64_t)tileSum & VAL_MASK));
            int64_t accum = 0;
            int64_t i = t - 1;
            int spins = 0;
            bool done = false, direct = false;
            while (!done) {
                const int64_t j = i - lane;
                const uint64_t st = j >= 0 ? loadRelaxed(&status[j]) : (FLAG_INC << 62);
                const uint64_t f = st >> 62;
                const uint64_t incMask = __ballot(f == FLAG_INC);
                const uint64_t zeroMask = __ballot(f == 0);
                const int firstInc = incMask ? __ffsll((unsigned long long)incMask) - 1 : 64;
                const uint64_t upto = firstInc >= 63 ? ~0ULL : ((1ULL << (firstInc + 1)) - 1ULL);
                if (zeroMask & upto) {
                    // the ballots make this branch and the counter uniform over the wave
                    if (++spins > kSpinLimit) {
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
                // the selected elements before this tile, counted straight from the input
                int64_t sum = 0;
                for (int64_t k = lane; k < base; k += 64) sum += (attr[k] & mask) ? 1 : 0;
#pragma unroll
                for (int d = 32; d > 0; d >>= 1) sum += __shfl_down(sum, d, 64);
                accum = __shfl(sum, 0, 64);
                if (lane == 0 && directOut) atomicAdd((unsigned long long*)directOut, 1ULL);
            }
            if (lane == 0) {
                storeRelaxed(&status[t], (FLAG_INC << 62) | ((uint64_t)(accum + tileSum) & VAL_MASK));
                s_off = accum;
            }
        }
    }
    __syncthreads();
    int64_t o = s_off + run;
    for (uint32_t b = bits; b; b &= b - 1) idx[o++] = (int32_t)(first + __ffs(b) - 1);
    if (t == numTiles - 1 && tid == 0) *total = s_off + tileSum;
}
}  // namespace

size_t scanScratchBytes(int64_t n) {
    const int64_t tiles = (n + kTile - 1) / kTile;
    return (size_t)(tiles + 2) * sizeof(unsigned long long) + 256;
}

void scanExclusiveI32(const int32_t* in, int64_t* out, int64_t n, int64_t* total, void* scratch, hipStream_t s, int64_t* directOut) {
    unsigned long long* w = (unsigned long long*)scratch;
    const int64_t tiles = n > 0 ? (n + kTile - 1) / kTile : 0;
    const int64_t words = tiles + 2;
    kScanReset<<<(unsigned)((words + 255) / 256), 256, 0, s>>>(w, words);
    if (tiles == 0) {
        kScanReset<<<1, 1, 0, s>>>((unsigned long long*)total, 1);
        return;
    }
    kScan<<<(unsigned)tiles, kScanBlock, 0, s>>>(in, out, n, tiles, w, total, directOut);
}

void compactFlagged(const uint32_t* attr, int64_t n, uint32_t mask, int32_t* idx, int64_t* total, void* scratch, hipStream_t s, int64_t* directOut) {
    unsigned long long* w = (unsigned long long*)scratch;
    const int64_t tiles = n > 0 ? (n + kTile - 1) / kTile : 0;
    const int64_t words = tiles + 2;
    kScanReset<<<(unsigned)((words + 255) / 256), 256, 0, s>>>(w, words);
    if (tiles == 0) {
        kScanReset<<<1, 1, 0, s>>>((unsigned long long*)total, 1);
        return;
    }
    kCompactFlagged<<<(unsigned)tiles, kScanBlock, 0, s>>>(attr, n, mask, idx, tiles, w, total, directOut);
}

}  // namespace mdmc
