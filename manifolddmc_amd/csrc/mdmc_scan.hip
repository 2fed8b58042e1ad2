// mdmc_scan.hip -- device-wide exclusive prefix sums (reduce, scan the block sums, add back).
#include "mdmc_internal.h"

namespace mdmc {

namespace {
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kTile = kScanBlock * kScanItems;

// inclusive scan of one int64 per thread across a 256-thread block (4 waves of 64)
__device__ __forceinline__ int64_t blockInclusive(int64_t v, int64_t* lds) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    if (lane == 63) lds[wid] = v;
    __syncthreads();
    int64_t add = 0;
    for (int w = 0; w < wid; ++w) add += lds[w];
    __syncthreads();
    return v + add;
}

template <class T>
__global__ void __launch_bounds__(kScanBlock) kScanLocal(const T* __restrict__ in, int64_t* __restrict__ out, int64_t n, int64_t* __restrict__ sums) {
    __shared__ int64_t lds[4];
    const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kScanItems;
    int64_t v[kScanItems];
    int64_t acc = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t j = base + i;
        v[i] = (j < n) ? (int64_t)in[j] : 0;
        acc += v[i];
    }
    const int64_t incl = blockInclusive(acc, lds);
    int64_t run = incl - acc;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t j = base + i;
        if (j < n) out[j] = run;
        run += v[i];
    }
    if (threadIdx.x == kScanBlock - 1) sums[blockIdx.x] = incl;
}

__global__ void __launch_bounds__(kScanBlock) kScanAdd(int64_t* __restrict__ out, int64_t n, const int64_t* __restrict__ offs) {
    const int64_t base = (int64_t)blockIdx.x * kTile;
    const int64_t o = offs[blockIdx.x];
    for (int i = threadIdx.x; i < kTile; i += kScanBlock) {
        const int64_t j = base + i;
        if (j < n) out[j] += o;
    }
}

__global__ void kScanTotal(const int64_t* sums, const int64_t* offs, int64_t nb, int64_t* total) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *total = (nb > 0) ? offs[nb - 1] + sums[nb - 1] : 0;
}

int64_t tiles(int64_t n) { return (n + kTile - 1) / kTile; }

__global__ void kZero64(int64_t* p) { *p = 0; }

template <class T>
void scanImpl(const T* in, int64_t* out, int64_t n, int64_t* total, char* scratch, hipStream_t s) {
    const int64_t nb = tiles(n);
    if (n <= 0) {
        kZero64<<<1, 1, 0, s>>>(total);
        return;
    }
    int64_t* sums = (int64_t*)scratch;
    int64_t* offs = sums + nb;
    char* rest = (char*)(offs + nb);
    kScanLocal<T><<<(unsigned)nb, kScanBlock, 0, s>>>(in, out, n, sums);
    if (nb > 1) {
        int64_t* t2 = (int64_t*)rest;
        scanImpl<int64_t>(sums, offs, nb, t2, rest + 64, s);
        kScanAdd<<<(unsigned)nb, kScanBlock, 0, s>>>(out, n, offs);
    } else {
        kZero64<<<1, 1, 0, s>>>(offs);
    }
    kScanTotal<<<1, 1, 0, s>>>(sums, offs, nb, total);
}
}  // namespace

size_t scanScratchBytes(int64_t n) {
    size_t bytes = 0;
    int64_t m = n;
    do {
        const int64_t nb = tiles(m);
        bytes += (size_t)(2 * nb) * sizeof(int64_t) + 64;
        m = nb;
    } while (m > 1);
    return bytes + 256;
}

void scanExclusiveI32(const int32_t* in, int64_t* out, int64_t n, int64_t* total, void* scratch, hipStream_t s) {
    scanImpl<int32_t>(in, out, n, total, (char*)scratch, s);
}
void scanExclusiveI64(const int64_t* in, int64_t* out, int64_t n, int64_t* total, void* scratch, hipStream_t s) {
    scanImpl<int64_t>(in, out, n, total, (char*)scratch, s);
}

}  // namespace mdmc
