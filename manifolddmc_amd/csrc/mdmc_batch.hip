// mdmc_batch.hip -- the batched form: many (grid, iso) problems in one extraction.
//
// No reference counterpart (the reference extracts one grid per call).  The
// problems' meshes are disjoint, so every stage of the pipeline runs once over
// their union: the grid pass walks all problems' tiles in one launch, and the
// halfedge, simplification, twin-repair, split and detection kernels work on
// global ids as they do for one mesh.  Element ids stay grouped by problem
// except where the splitter appends; at the end a stable sort by problem makes
// every problem's vertices, halfedges and faces contiguous (CSR-packed), in the
// order a single extraction of that problem produces.
//
// Until the splitter runs, every array is ordered by problem (the grid pass emits the
// problems in order and every compaction is stable).  The splitter only appends, so
// each array is a sorted prefix of n0 elements plus a tail; the tail is sorted stably
// by problem (least-significant-bit-first binary splits: per bit a prefix sum of the
// "bit is 0" flags and a scatter) and each problem's tail goes after its prefix part.
#include <algorithm>
#include <vector>

#include "mdmc_device.h"
#include "mdmc_internal.h"
#include "mdmc_pipeline.h"

namespace mdmc {

namespace {

__device__ __forceinline__ int heO(const int4* he, int h) { return ((const int*)he)[4 * (int64_t)h]; }

__global__ void kIota(int32_t* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}

// problem of each halfedge (its origin's) and of each face (its first halfedge's origin's)
__global__ void kHalfedgeKeys(const int4* __restrict__ he, int64_t H, const int32_t* __restrict__ pid, int32_t* __restrict__ key) {
    const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h < H) key[h] = pid[heO(he, (int)h)];
}
__global__ void kFaceKeys2(const int4* __restrict__ he, const int32_t* __restrict__ face_he, int64_t F, const int32_t* __restrict__ pid,
                           int32_t* __restrict__ key) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f < F) key[f] = pid[heO(he, face_he[f])];
}

__global__ void kSplitFlags(const int32_t* __restrict__ perm, const int32_t* __restrict__ key, int64_t n, int bit, int32_t* __restrict__ zero) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) zero[i] = ((key[perm[i]] >> bit) & 1) ? 0 : 1;
}
__global__ void kSplitScatter(const int32_t* __restrict__ perm, const int32_t* __restrict__ zero, const int64_t* __restrict__ pre, int64_t n,
                              const int64_t* __restrict__ zeros, int32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t p = pre[i];
    out[zero[i] ? p : *zeros + (i - p)] = perm[i];
}

// final position of every element: prefix element i of problem b -> out[b] + i - pre[b];
// tail element at sorted tail position j -> out[b] + preCount[b] + j - tail[b]
// (offsets: [0, B] prefix starts, [B+1, 2B+1] tail starts, [2B+2, 3B+2] output starts)
__global__ void kPlace(const int32_t* __restrict__ key, int64_t n0, int64_t n, const int32_t* __restrict__ tailPerm, const int64_t* __restrict__ offs,
                       int B, int32_t* __restrict__ perm, int32_t* __restrict__ inv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t* pre = offs;
    const int64_t* tail = offs + (B + 1);
    const int64_t* out = offs + 2 * (B + 1);
    int64_t old, dst;
    if (i < n0) {
        old = i;
        const int b = key[i];
        dst = out[b] + (i - pre[b]);
    } else {
        const int64_t j = i - n0;
        old = n0 + tailPerm[j];
        const int b = key[old];
        dst = out[b] + (pre[b + 1] - pre[b]) + (j - tail[b]);
    }
    perm[dst] = (int32_t)old;
    inv[old] = (int32_t)dst;
}

// Per-key counts of keys that arrive sorted (key(i) = key[perm ? perm[i] : i], values in
// [0, B)): thread i in [0, n] writes first[k] = i for every k in (key(i-1), key(i)], with
// key(-1) = -1 and key(n) = B, so first[k] is the first index whose key is >= k and the
// count of key k is first[k + 1] - first[k].  No atomics; a key out of order or out of
// range raises *bad.
__global__ void kKeyStarts(const int32_t* __restrict__ key, const int32_t* __restrict__ perm, int64_t n, int B, int64_t* __restrict__ first,
                           int* __restrict__ bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    const int k = i < n ? key[perm ? perm[i] : i] : B;
    const int kp = i > 0 ? key[perm ? perm[i - 1] : i - 1] : -1;
    if (k < kp || k > B || (i < n && (k == B || k < 0))) {
        atomicOr(bad, 1);
        return;
    }
    for (int kk = kp + 1; kk <= k; ++kk) first[kk] = i;
}
__global__ void kCountsFromStarts(const int64_t* __restrict__ first, int B, unsigned long long* __restrict__ count, int stride, int slot) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < B) count[(int64_t)k * stride + slot] = (unsigned long long)(first[k + 1] - first[k]);
}

__global__ void kPermuteVertices(const int32_t* __restrict__ perm, int64_t V, const float* __restrict__ pos, const float* __restrict__ nrm,
                                 const uint32_t* __restrict__ vflags, const int32_t* __restrict__ pid, float* __restrict__ npos,
                                 float* __restrict__ nnrm, uint32_t* __restrict__ nvflags, int32_t* __restrict__ npid) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const int64_t o = perm[i];
    for (int k = 0; k < 3; ++k) {
        npos[3 * i + k] = pos[3 * o + k];
        nnrm[3 * i + k] = nrm[3 * o + k];
    }
    nvflags[i] = vflags[o];
    npid[i] = pid[o];
}

__global__ void kPermuteHalfedges(const int32_t* __restrict__ perm, int64_t H, const int4* __restrict__ he, const int32_t* __restrict__ bsize,
                                  const int32_t* __restrict__ brank, const int32_t* __restrict__ vNew, const int32_t* __restrict__ hNew,
                                  const int32_t* __restrict__ fNew, int4* __restrict__ nhe, int32_t* __restrict__ nbsize,
                                  int32_t* __restrict__ nbrank) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= H) return;
    const int o = perm[i];
    const int4 e = he[o];
    nhe[i] = make_int4(vNew[e.x], fNew[e.y], hNew[e.z], e.w >= 0 ? hNew[e.w] : e.w);
    nbsize[i] = bsize[o];
    nbrank[i] = brank[o];
}

__global__ void kPermuteFaces(const int32_t* __restrict__ perm, int64_t F, const int32_t* __restrict__ face_he, const uint32_t* __restrict__ fattr,
                              const int32_t* __restrict__ hNew, int32_t* __restrict__ nface_he, uint32_t* __restrict__ nfattr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F) return;
    const int o = perm[i];
    nface_he[i] = hNew[face_he[o]];
    nfattr[i] = fattr[o];
}

// the vertex -> halfedge array of each problem as its own extraction leaves it: highest
// initialize-time halfedge 4q+i starting at the vertex, q counted inside the problem;
// -1 where none; 0 for vertices the splitter added
__global__ void kVertexHalfedgeInitBatch(int32_t* __restrict__ vh, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) vh[i] = -1;
}
__global__ void kVertexHalfedgeMax(const int4* __restrict__ quads, int64_t Q, int32_t* __restrict__ vh) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Q) return;
    const int4 v = quads[q];
    const int b = (int)(4 * q);
    atomicMax(&vh[v.x], b + 0);
    atomicMax(&vh[v.y], b + 1);
    atomicMax(&vh[v.z], b + 2);
    atomicMax(&vh[v.w], b + 3);
}
__global__ void kVertexHalfedgeLocal(const int32_t* __restrict__ vhInit, int64_t Vinit, const int32_t* __restrict__ vNew, const int32_t* __restrict__ pidNew,
                                     const int64_t* __restrict__ qBase, int64_t V, int32_t* __restrict__ out) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    if (v >= Vinit) {
        out[vNew[v]] = 0;
        return;
    }
    const int h = vhInit[v];
    const int64_t nv = vNew[v];
    out[nv] = h < 0 ? -1 : (int32_t)(h - 4 * qBase[pidNew[nv]]);
}

// bucket / face / vertex counters of HalfedgeMesh.cu:1164-1239 per problem.  Elements are
// sorted by problem, so a wave usually holds one problem: it then adds its sums once.
// Per-problem statistics of one element kind over its regrouped range [0, n), where every
// problem's elements are contiguous and problems ascend: each block reduces a chunk of
// kStatChunk elements in LDS, one row per problem the chunk touches (up to kStatRows; the
// elements of further ones add to global memory directly), and adds every non-zero row
// entry once.  KIND 0: halfedges (slots 0-2: NM-4 / NM-3 / all edges, counted at the
// bucket's first member), 1: faces (slots 3-7: quads with k NM edges), 2: vertices (slot 8:
// non-manifold vertices).  out[9 * problem + slot].
constexpr int kStatChunk = 4096, kStatRows = 32;

__device__ __forceinline__ void statAdd(unsigned* acc, unsigned long long* out, bool live, int p, int pbase, int slot, int accSlot, unsigned v) {
    // one LDS atomic per wave when all its live lanes share the problem
    const int p0 = __shfl(p, 0, 64);
    const uint64_t alive = __ballot(live);
    const uint64_t same = __ballot(live && p == p0);
    if (alive == 0) return;
    if (same == alive) {
        unsigned sum = live ? v : 0u;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) sum += __shfl_down(sum, d, 64);
        const int pk = __shfl(p, __ffsll((unsigned long long)alive) - 1, 64);
        if ((threadIdx.x & 63) == 0 && sum) {
            const int r = pk - pbase;
            if (r >= 0 && r < kStatRows) atomicAdd(&acc[r * 5 + accSlot], sum);
            else atomicAdd(&out[(int64_t)pk * 9 + slot], (unsigned long long)sum);
        }
    } else if (live && v) {
        const int r = p - pbase;
        if (r >= 0 && r < kStatRows) atomicAdd(&acc[r * 5 + accSlot], v);
        else atomicAdd(&out[(int64_t)p * 9 + slot], (unsigned long long)v);
    }
}

template <int KIND>
__global__ void kStatsBatch(const int4* __restrict__ he, const int32_t* __restrict__ face_he, const int32_t* __restrict__ bsize,
                            const int32_t* __restrict__ brank, const uint32_t* __restrict__ vflags, const int32_t* __restrict__ pid, int64_t n,
                            unsigned long long* __restrict__ out) {
    constexpr int nslots = KIND == 0 ? 3 : (KIND == 1 ? 5 : 1);
    constexpr int base = KIND == 0 ? 0 : (KIND == 1 ? 3 : 8);
    __shared__ unsigned acc[kStatRows * 5];
    const int64_t c0 = (int64_t)blockIdx.x * kStatChunk;
    const int64_t c1 = c0 + kStatChunk < n ? c0 + kStatChunk : n;
    for (int t = threadIdx.x; t < kStatRows * 5; t += blockDim.x) acc[t] = 0;
    auto problemOf = [&](int64_t i) -> int {
        if (KIND == 0) return pid[heO(he, (int)i)];
        if (KIND == 1) return pid[heO(he, face_he[i])];
        return pid[i];
    };
    const int pbase = problemOf(c0);
    __syncthreads();
    // uniform trip count so that every lane takes part in the wave reductions
    for (int64_t i0 = c0; i0 < c1; i0 += blockDim.x) {
        const int64_t i = i0 + threadIdx.x;
        const bool inRange = i < c1;
        if (KIND == 0) {
            const bool live = inRange && brank[i] == 0;
            const int p = live ? problemOf(i) : 0;
            const int nb = live ? bsize[i] : 0;
            statAdd(acc, out, live, p, pbase, 0, 0, nb == 4);
            statAdd(acc, out, live, p, pbase, 1, 1, nb == 3);
            statAdd(acc, out, live, p, pbase, 2, 2, nb > 0);
        } else if (KIND == 1) {
            int p = 0, k = 0;
            if (inRange) {
                int h = face_he[i];
                p = pid[heO(he, h)];
                for (int j = 0; j < 4; ++j) {
                    if (bsize[h] > 2) ++k;
                    h = ((const int*)he)[4 * (int64_t)h + 2];
                }
            }
            for (int kk = 0; kk < 5; ++kk) statAdd(acc, out, inRange, p, pbase, 3 + kk, kk, k == kk);
        } else {
            const int p = inRange ? pid[i] : 0;
            statAdd(acc, out, inRange, p, pbase, 8, 0, inRange && (vflags[i] & VF_FANS) ? 1u : 0u);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kStatRows * nslots; t += blockDim.x) {
        const int r = t / nslots, s = t % nslots;
        const unsigned v = acc[r * 5 + s];
        if (v) atomicAdd(&out[(int64_t)(pbase + r) * 9 + base + s], (unsigned long long)v);
    }
}

}  // namespace

// stable permutation of [0, n) sorting key[] in [0, B)
int32_t* Pipeline::stableSortByKey(const int32_t* key, int64_t n, int B, const std::string& tag) {
    int32_t* a = arena.get<int32_t>(tag + "_perm", (size_t)std::max<int64_t>(n, 1));
    int32_t* b = arena.get<int32_t>(tag + "_perm2", (size_t)std::max<int64_t>(n, 1));
    int32_t* flag = arena.get<int32_t>("sort_flag", (size_t)std::max<int64_t>(n, 1));
    int64_t* pre = arena.get<int64_t>("sort_pre", (size_t)n + 1);
    int64_t* zeros = arena.get<int64_t>("sort_zeros", 1);
    void* scratch = arena.get<char>("sort_scan", scanScratchBytes(n + 1));
    if (n == 0) return a;
    kIota<<<gridFor(n), kBlock, 0, stream>>>(a, n);
    int bits = 0;
    while ((1 << bits) < B) ++bits;
    for (int bit = 0; bit < bits; ++bit) {
        kSplitFlags<<<gridFor(n), kBlock, 0, stream>>>(a, key, n, bit, flag);
        scanExclusiveI32(flag, pre, n, zeros, scratch, stream, lookbackDiag());
        kSplitScatter<<<gridFor(n), kBlock, 0, stream>>>(a, flag, pre, n, zeros, b);
        std::swap(a, b);
    }
    return a;
}

// counts of keys that arrive sorted, into count[k * stride + slot] (k in [0, B))
void Pipeline::countSorted(const int32_t* key, const int32_t* perm, int64_t n, int B, unsigned long long* count, int stride, int slot) {
    int64_t* first = arena.get<int64_t>("key_first", (size_t)B + 1);
    int* bad = arena.get<int>("batch_bad", 1);
    kKeyStarts<<<gridFor(n + 1), kBlock, 0, stream>>>(key, perm, n, B, first, bad);
    kCountsFromStarts<<<gridFor(B), kBlock, 0, stream>>>(first, B, count, stride, slot);
}

void Pipeline::checkBatchKeys() {
    int bad = 0;
    MDMC_HIP(hipMemcpyAsync(&bad, arena.get<int>("batch_bad", 1), sizeof(int), hipMemcpyDeviceToHost, stream));
    MDMC_HIP(hipStreamSynchronize(stream));
    if (bad) throw HipError("batched form: elements out of problem order", -4);
}

void Pipeline::problemDmcCounts(int B) {
    zero(arena.get<int>("batch_bad", 1), sizeof(int));
    batchDmc.assign(2 * (size_t)B, 0);
    unsigned long long* cnt = arena.get<unsigned long long>("batch_dmc", 2 * (size_t)B);
    zero(cnt, 2 * (size_t)B * sizeof(unsigned long long));
    if (V > 0) countSorted(pid(), nullptr, V, B, cnt, 2, 0);
    if (Q > 0) {
        int32_t* qk = arena.get<int32_t>("batch_qkey", (size_t)Q);
        // a quad's problem is its first vertex's: reuse the halfedge-key kernel on the quads (x = first vertex)
        kHalfedgeKeys<<<gridFor(Q), kBlock, 0, stream>>>(arena.ptr<int4>("quads"), Q, pid(), qk);
        countSorted(qk, nullptr, Q, B, cnt, 2, 1);
    }
    MDMC_HIP(hipMemcpyAsync(batchDmc.data(), cnt, 2 * (size_t)B * sizeof(int64_t), hipMemcpyDeviceToHost, stream));
    checkBatchKeys();
}

// permutation (new -> old) and its inverse for one array of n elements whose first n0
// are ordered by key; per-key counts go to cnt[4 * b + slot]
void Pipeline::placeByProblem(const int32_t* key, int64_t n0, int64_t n, int B, const std::string& tag, int32_t** perm, int32_t** inv,
                              std::vector<int64_t>& off) {
    *perm = arena.get<int32_t>(tag + "_perm", (size_t)std::max<int64_t>(n, 1));
    *inv = arena.get<int32_t>(tag + "_inv", (size_t)std::max<int64_t>(n, 1));
    const int64_t nt = n - n0;
    const int32_t* tp = stableSortByKey(key + n0, nt, B, tag + "_tail");
    unsigned long long* cnt = arena.get<unsigned long long>(tag + "_cnt", 2 * (size_t)B);
    zero(cnt, 2 * (size_t)B * sizeof(unsigned long long));
    if (n0 > 0) countSorted(key, nullptr, n0, B, cnt, 2, 0);
    if (nt > 0) countSorted(key + n0, tp, nt, B, cnt, 2, 1);
    std::vector<unsigned long long> c(2 * (size_t)B);
    MDMC_HIP(hipMemcpyAsync(c.data(), cnt, c.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    MDMC_HIP(hipStreamSynchronize(stream));
    std::vector<int64_t> offs(3 * ((size_t)B + 1), 0);
    for (int b = 0; b < B; ++b) {
        offs[b + 1] = offs[b] + (int64_t)c[2 * b];
        offs[(B + 1) + b + 1] = offs[(B + 1) + b] + (int64_t)c[2 * b + 1];
        offs[2 * (B + 1) + b + 1] = offs[2 * (B + 1) + b] + (int64_t)(c[2 * b] + c[2 * b + 1]);
    }
    off.assign(offs.begin() + 2 * (B + 1), offs.end());
    int64_t* od = arena.get<int64_t>(tag + "_offs", offs.size());
    upload(od, offs.data(), offs.size() * sizeof(int64_t), 0);
    if (n > 0) kPlace<<<gridFor(n), kBlock, 0, stream>>>(key, n0, n, tp, od, B, *perm, *inv);
}

void Pipeline::regroupBatch(int B) {
    zero(arena.get<int>("batch_bad", 1), sizeof(int));
    int32_t* vp = pid();
    int4* he = arena.ptr<int4>("he");
    int32_t* face_he = arena.ptr<int32_t>("face_he");
    // quads per problem at the last initialize (for the vertex -> halfedge values)
    unsigned long long* qc = arena.get<unsigned long long>("batch_qcnt", (size_t)B);
    zero(qc, (size_t)B * sizeof(unsigned long long));
    if (vertHeQ > 0) {
        int32_t* qk = arena.get<int32_t>("batch_qkey", (size_t)vertHeQ);
        kHalfedgeKeys<<<gridFor(vertHeQ), kBlock, 0, stream>>>(arena.ptr<int4>("quads"), vertHeQ, vp, qk);
        countSorted(qk, nullptr, vertHeQ, B, qc, 1, 0);
    }
    int32_t* vhInit = arena.get<int32_t>("batch_vh_init", (size_t)std::max<int64_t>(vertHeV, 1));
    if (vertHeV > 0) kVertexHalfedgeInitBatch<<<gridFor(vertHeV), kBlock, 0, stream>>>(vhInit, vertHeV);
    if (vertHeQ > 0) kVertexHalfedgeMax<<<gridFor(vertHeQ), kBlock, 0, stream>>>(arena.ptr<int4>("quads"), vertHeQ, vhInit);
    std::vector<unsigned long long> qcount((size_t)B);
    MDMC_HIP(hipMemcpyAsync(qcount.data(), qc, (size_t)B * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    // keys; the elements that existed at the last initialize come first, in problem order
    int32_t* hk = arena.get<int32_t>("batch_hkey", (size_t)std::max<int64_t>(H, 1));
    int32_t* fk = arena.get<int32_t>("batch_fkey", (size_t)std::max<int64_t>(F, 1));
    if (H > 0) kHalfedgeKeys<<<gridFor(H), kBlock, 0, stream>>>(he, H, vp, hk);
    if (F > 0) kFaceKeys2<<<gridFor(F), kBlock, 0, stream>>>(he, face_he, F, vp, fk);
    int32_t *pv, *ph, *pf, *vNew, *hNew, *fNew;
    placeByProblem(vp, std::min(vertHeV, V), V, B, "sv", &pv, &vNew, batchVOff);
    placeByProblem(hk, std::min(4 * vertHeQ, H), H, B, "sh", &ph, &hNew, batchHOff);
    placeByProblem(fk, std::min(vertHeQ, F), F, B, "sf", &pf, &fNew, batchFOff);
    std::vector<int64_t> qBase((size_t)B + 1, 0);
    for (int b = 0; b < B; ++b) qBase[b + 1] = qBase[b] + (int64_t)qcount[b];
    int64_t* qBaseDev = arena.get<int64_t>("batch_qbase", (size_t)B + 1);
    upload(qBaseDev, qBase.data(), qBase.size() * sizeof(int64_t), 1);
    float* npos = arena.get<float>("pos2", 3 * (size_t)std::max<int64_t>(V, 1));
    float* nnrm = arena.get<float>("nrm2", 3 * (size_t)std::max<int64_t>(V, 1));
    uint32_t* nvf = arena.get<uint32_t>("vflags2", (size_t)std::max<int64_t>(V, 1));
    int32_t* npid = arena.get<int32_t>("pid2", (size_t)std::max<int64_t>(V, 1));
    int4* nhe = arena.get<int4>("he2", (size_t)std::max<int64_t>(H, 1));
    int32_t* nbs = arena.get<int32_t>("bsize2", (size_t)std::max<int64_t>(H, 1));
    int32_t* nbr = arena.get<int32_t>("brank2", (size_t)std::max<int64_t>(H, 1));
    int32_t* nfh = arena.get<int32_t>("face_he2", (size_t)std::max<int64_t>(F, 1));
    uint32_t* nfa = arena.get<uint32_t>("fattr2", (size_t)std::max<int64_t>(F, 1));
    if (V > 0)
        kPermuteVertices<<<gridFor(V), kBlock, 0, stream>>>(pv, V, arena.ptr<float>("pos"), arena.ptr<float>("nrm"), arena.ptr<uint32_t>("vflags"), vp, npos,
                                                            nnrm, nvf, npid);
    if (H > 0)
        kPermuteHalfedges<<<gridFor(H), kBlock, 0, stream>>>(ph, H, he, arena.ptr<int32_t>("bsize"), arena.ptr<int32_t>("brank"), vNew, hNew, fNew, nhe,
                                                             nbs, nbr);
    if (F > 0) kPermuteFaces<<<gridFor(F), kBlock, 0, stream>>>(pf, F, face_he, arena.ptr<uint32_t>("fattr"), hNew, nfh, nfa);
    int32_t* vh = arena.get<int32_t>("batch_vert_he", (size_t)std::max<int64_t>(V, 1));
    if (V > 0) kVertexHalfedgeLocal<<<gridFor(V), kBlock, 0, stream>>>(vhInit, vertHeV, vNew, npid, qBaseDev, V, vh);
    auto back = [&](const char* name, const void* src, size_t bytes) {
        if (bytes) MDMC_HIP(hipMemcpyAsync(arena.ptr<char>(name), src, bytes, hipMemcpyDeviceToDevice, stream));
    };
    back("pos", npos, 3 * (size_t)V * sizeof(float));
    back("nrm", nnrm, 3 * (size_t)V * sizeof(float));
    back("vflags", nvf, (size_t)V * sizeof(uint32_t));
    back("pid", npid, (size_t)V * sizeof(int32_t));
    back("he", nhe, (size_t)H * sizeof(int4));
    back("bsize", nbs, (size_t)H * sizeof(int32_t));
    back("brank", nbr, (size_t)H * sizeof(int32_t));
    back("face_he", nfh, (size_t)F * sizeof(int32_t));
    back("fattr", nfa, (size_t)F * sizeof(uint32_t));
    batchVertHe = true;
    // per-problem statistics
    unsigned long long* st = arena.get<unsigned long long>("batch_stats", 9 * (size_t)B);
    zero(st, 9 * (size_t)B * sizeof(unsigned long long));
    {
        const int4* h4 = arena.ptr<int4>("he");
        const int32_t* fh = arena.ptr<int32_t>("face_he");
        const int32_t* bs = arena.ptr<int32_t>("bsize");
        const int32_t* br = arena.ptr<int32_t>("brank");
        const uint32_t* vf = arena.ptr<uint32_t>("vflags");
        auto chunks = [](int64_t n) { return (unsigned)((n + kStatChunk - 1) / kStatChunk); };
        if (H > 0) kStatsBatch<0><<<chunks(H), kBlock, 0, stream>>>(h4, fh, bs, br, vf, pid(), H, st);
        if (F > 0) kStatsBatch<1><<<chunks(F), kBlock, 0, stream>>>(h4, fh, bs, br, vf, pid(), F, st);
        if (V > 0) kStatsBatch<2><<<chunks(V), kBlock, 0, stream>>>(h4, fh, bs, br, vf, pid(), V, st);
    }
    batchStats.assign(9 * (size_t)B, 0);
    MDMC_HIP(hipMemcpyAsync(batchStats.data(), st, batchStats.size() * sizeof(int64_t), hipMemcpyDeviceToHost, stream));
    checkBatchKeys();
}

}  // namespace mdmc
