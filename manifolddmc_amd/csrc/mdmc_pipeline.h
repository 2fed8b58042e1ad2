// mdmc_pipeline.h -- the stage driver behind mdmc_extract.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

#include "mdmc_arena.h"
#include "mdmc_gridpass.h"
#include "mdmc_internal.h"

namespace mdmc {

// up to kFillSetMax byte fills for one launch (kFillMany): per range its pointer, the unaligned
// head and tail bytes, the 16-byte aligned body and the range's first block
constexpr int kFillSetMax = 12;
struct FillSet {
    uint8_t* p[kFillSetMax];
    size_t head[kFillSetMax], body[kFillSetMax], n[kFillSetMax];
    unsigned start[kFillSetMax];
    uint8_t byte[kFillSetMax];
    int count = 0;
};

class Pipeline {
   public:
    Pipeline();
    ~Pipeline();

    hipStream_t stream = nullptr;
    Arena arena;

    // live sizes of the resident mesh
    int64_t V = 0, Q = 0, H = 0, F = 0;
    int64_t dmcV = 0, dmcQ = 0;
    // vertices and quads of the last halfedge initialize (the vertex -> halfedge array)
    int64_t vertHeV = 0, vertHeQ = 0;
    int64_t splitIterations = 0;
    int64_t extraAnomalies = 0;

    // simplifications still to run: while positive, the halfedge initialize computes the
    // vertex valence they start from
    int simplifyQueue = 0;
    bool valenceReady = false;
    // simplifications run since the initialize without compaction (the last of a chain compacts)
    int deferredSteps = 0;
    // vertices the outgoing lists of the last halfedge build describe (-1: none), and
    // whether a split since then keeps the per-vertex "touched" flags
    int64_t listsV = -1;
    bool splitTouched = false;
    // per halfedge whether its record still has the build's layout (next = quadNext, face =
    // id / 4): tracked once a split step writes records back; before that every one has
    bool canonTracked = false;
    const uint8_t* canonFlags();
    // the halfedge records (he) are written from the quads and the twins after the
    // simplifications, which read only those
    bool heMaterialized = false;
    void materializeHalfedges();
    // while the simplifications run, vertex v's position and normal stay at slot vslot[v] of
    // pos / nrm (not moved by each compaction); materializeHalfedges gathers them back
    bool posSlotted = false;
    const int32_t* vertexSlots();

    void resetCounters();
    // the counters an extraction reports, in one read-back: anomalies, split iterations that
    // took the sequential merge pass, look-back tiles that summed their prefix directly
    void counterTotals(int64_t* anomalyCount, int64_t* mergeSeq, int64_t* lookbackDirect);
    // device word the single-pass scans and compactions count their direct-sum tiles in
    int64_t* lookbackDiag();
    // the padding pass's flag word (one of the counters, zeroed by resetCounters)
    int64_t* prepSpecial();

    // the grid pass over one or more (grid, iso) problems (their cellBase / tileBase and
    // cell counts are filled in); recordAdjoint keeps the cell -> vertex map ("adj_map") that
    // the adjoint reads with the vertex -> cell array ("gp_vcell") (one problem only)
    void dmc(std::vector<ProblemDev>& problems, bool surfIter, bool recordAdjoint = false);
    // the grid pass in parts, for the batched form's graph: problem geometry (no GPU work), the
    // tuned or fixed shape (GP_SHAPE_AUTO while a tuning is due), and the shape-determined front
    // (padding, classification, tile-count scans) captured once and replayed for later calls
    // with the same shapes and buffers (false: not replayable yet, run it eagerly), then the rest
    int64_t dmcSetup(std::vector<ProblemDev>& problems, bool recordAdjoint);
    int64_t assignTiles(std::vector<ProblemDev>& problems, int shape);
    void gridPassFront(std::vector<ProblemDev>& problems, int64_t C, int shape, bool surfIter, void* stagedProbs, bool events);
    void gridPassBack(int64_t C, int shape, bool recordAdjoint);
    int shapeClass(int64_t C, bool surfIter) const;
    int resolveShape(int64_t C, bool surfIter);
    bool runFrontGraph(const std::vector<int64_t>& shapeSig, const void* prep, size_t prepBytes, void* prepDev, int G, int64_t prepBlocks,
                       int refinement, int pad, std::vector<ProblemDev>& problems, int64_t C, int shape, bool surfIter);
    void noteEagerFront(const std::vector<int64_t>& shapeSig) {
        fgLastEager_ = shapeSig;
        const std::vector<int64_t> b = frontBuffers();
        fgLastEager_.insert(fgLastEager_.end(), b.begin(), b.end());
    }
    void releaseGraphs();
    bool useGraphs = true;
    // P3X3Y merges of a deferred chain step with a plain neighbourhood update their twins
    // directly (else every merge's region is rebuilt)
    bool localMerges = true;
    // the splitter's merge phase always in the one-thread ascending pass (a verification switch)
    bool sequentialMerges = false;
    int64_t graphReplays = 0;
    int nProblems = 1;
    ProblemDev* probsDev = nullptr;
    // per vertex its problem (several problems), else nullptr
    int32_t* pid();

    // batched form (mdmc_batch.hip): per-problem DMC vertex / quad counts right after the
    // grid pass, then the final regroup into contiguous per-problem ranges with offsets,
    // per-problem vertex -> halfedge values and statistics
    void problemDmcCounts(int B);
    void regroupBatch(int B);
    std::vector<int64_t> batchDmc;                     // [2B]: vertices, quads
    std::vector<int64_t> batchVOff, batchHOff, batchFOff;  // [B + 1]
    std::vector<int64_t> batchStats;                   // [9B], as computeStats
    bool batchVertHe = false;
    void halfedgeInitialize();
    void simplify3X3Y(bool colour);
    void simplify3333();
    void fixManifoldTwins();
    void split(int step, int maxIters);
    void detectNonManifoldVertices();
    void computeStats(int64_t out[9]);

    void copyOut(float* positions, float* normals, int32_t* halfedges, int32_t* face_halfedge, uint16_t* face_attr, uint8_t* vertex_flags,
                 int32_t* vertex_halfedge, bool dstDevice);
    void triangles(uint32_t* idx, bool dstDevice);
    void renderVertices(float* out, bool dstDevice);
    void faceColours(uint8_t* out, bool dstDevice);
    void triangulateMaxMin(uint32_t* tri, float* quality, bool dstDevice);

    int64_t readI64(const int64_t* dev);
    // mapped pinned host words for small read-backs: kernels and copies write them
    // directly, so a read-back is one stream synchronisation (slots 0..kPinSlots-1)
    static constexpr int kPinSlots = 64;
    int64_t* pinHost();
    int64_t* pinDev();
    // host -> device copy of a small host array through pinned staging (region 0 or 1, so
    // that two uploads of one stage do not share a buffer); returns after queueing the copy
    void upload(void* dst, const void* src, size_t bytes, int region);
    // duration of the last grid pass (classify + scan, then emit), from events on the
    // launch stream; the host read-back of the sizes between them is not counted
    void kernelTimes(double* msGridPass);
    hipEvent_t kev[4] = {nullptr, nullptr, nullptr, nullptr};
    // grid-pass tile shape: GP_SHAPE_AUTO (autotuned per workload class) or a fixed shape;
    // lastShape is the one the last extraction used, tuneMs the last tuning's timings
    int tileShape = GP_SHAPE_AUTO;
    int lastShape = GP_SHAPE_DEFAULT;
    std::map<int, int> tuned;
    double tuneMs[GP_SHAPE_LAST + 1] = {};

   private:
    int64_t* pinH_ = nullptr;
    int64_t* pinD_ = nullptr;
    void* up_[2] = {nullptr, nullptr};
    hipEvent_t upEv_[2] = {nullptr, nullptr};
    size_t upCap_[2] = {0, 0};
    int64_t* counters();
    int64_t* nmSlotParts();
    int64_t* nmFoundParts();
    int64_t scan32(const int32_t* in, int64_t n, const std::string& outName, int64_t** out);
    int64_t compactNM(int32_t* nm);
    void zero(void* p, size_t bytes);
    void fillNeg(void* p, size_t bytes);
    void launchFill(void* p, uint8_t byte, size_t bytes);
    // fills collected and launched together (at flush or at the end of the scope)
    class FillBatch {
       public:
        explicit FillBatch(Pipeline& p) : pl(p) {}
        ~FillBatch() { flush(); }
        void add(void* p, uint8_t byte, size_t bytes);
        void flush();

       private:
        Pipeline& pl;
        FillSet fs;
        unsigned blocks = 0;
    };
    // the splitter's old -> copy maps: -1 on [0, clean) between steps (each step resets the
    // entries it set); a map the arena reallocated, or one a failed step left dirty, is filled
    struct WcMap {
        void* ptr = nullptr;
        int64_t clean = 0, cap = 0;
    } wcMaps[3];
    bool wcMapsDirty = true;
    void gridPass(std::vector<ProblemDev>& problems, int64_t C, int shape, bool surfIter, bool recordAdjoint);
    void slabPass(std::vector<ProblemDev>& problems, int64_t C, int shape, bool surfIter, bool recordAdjoint);
    GridPassArgs ga_{};
    // the captured front of a batched extraction: the graph, the shapes / options / buffer
    // addresses it was captured with, its pinned descriptor staging, its kernel arguments
    struct FrontGraph {
        hipGraphExec_t exec = nullptr;
        std::vector<int64_t> sig;
        void* stage = nullptr;
        size_t stageCap = 0;
    } fg_;
    GridPassArgs fgArgs_{};
    std::vector<int64_t> fgLastEager_;
    std::vector<int64_t> frontBuffers();
    void buildBuckets(const int32_t* list, int64_t n);
    int32_t* bucketPassValence(const int64_t* off, int64_t nV);
    void refreshHashTable(const uint8_t* dirtyV, int64_t oldH);
    // a refresh ran whose face-walk check (the invalid-record counter) is not read back yet
    bool refreshUnchecked = false;
    void checkValid(const char* what);
    void simplifyInit(int32_t** valence, int32_t** count, bool needCount);
    void simplifyCompactAndRebuild(int dropType, bool p3333);
    void simplifyDeferredStep(int dropType, bool p3333);
    void buildMergeList(const int2* rec, int64_t R, int count, int32_t* num, int32_t* ex, int32_t* ey, const std::string& tag);
    int32_t* stableSortByKey(const int32_t* key, int64_t n, int B, const std::string& tag);
    void countSorted(const int32_t* key, const int32_t* perm, int64_t n, int B, unsigned long long* count, int stride, int slot);
    void checkBatchKeys();
    void placeByProblem(const int32_t* key, int64_t n0, int64_t n, int B, const std::string& tag, int32_t** perm, int32_t** inv,
                        std::vector<int64_t>& off);
};

}  // namespace mdmc
