// mdmc_api.hip -- the C ABI (include/mdmc.h) over the stage driver.
//
// mdmc_extract follows DualMarchingCubes::run + runPostProcesses
// (DualMarchingCubes.cu:220-289, :335-380) with the grid preparation of
// UniformGridHost::readFromFile(file, refinement, addZeroBoundary)
// (UniformGrid.inl:573-608):
//   pad/refine -> DMC -> HalfedgeMesh::initialize -> [P3X3Y colour] -> [P3X3Y old]
//   -> [P3333] -> fixManifoldTwins (always) -> [split] -> [detect NM vertices]
#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/mdmc.h"
#include "mdmc_adjoint.h"
#include "mdmc_pipeline.h"

// what mdmc_backward needs from the last mdmc_extract
struct AdjointState {
    bool valid = false;
    mdmc::GridDev g{};  // the processed grid (arena buffer "grid")
    float iso = 0.f;
    int inDims[3] = {0, 0, 0}, outDims[3] = {0, 0, 0};
    int refinement = 0, pad = 0;
    int64_t V = 0;
};

struct mdmc_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    mdmc::Pipeline p;
    std::string err;
    bool hasResult = false;
    hipEvent_t ev[9] = {};
    AdjointState adj;
    int nProblems = 0;  // problems of the resident result
    int flags = 0;
    std::vector<mdmc::GridDev> procGrids;  // the processed grids of the resident result
};

namespace {

// Scratch of mdmc_backward_noalloc (the 8 corner contributions of every dual vertex, then the
// gradient on the padded / refined grid): at most kNoAllocFloats floats (128 MiB) per device,
// held from the library's load until the process ends.  Host code never passes `scratch`
// itself to a kernel: its address on each device comes from hipGetSymbolAddress in
// mdmc_create (with that device current) and is kept in g_noallocScratch.
constexpr size_t kNoAllocFloats = (size_t)32 << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) float scratch[kNoAllocFloats];
float* g_noallocScratch[kMaxDevices] = {};

mdmc_status fail(mdmc_ctx* ctx, mdmc_status st, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return st;
}

mdmc_status fromException(mdmc_ctx* ctx, const std::exception& e) {
    if (auto* h = dynamic_cast<const mdmc::HipError*>(&e)) {
        if (h->code == -2) return fail(ctx, MDMC_ERR_OOM, e.what());
        if (h->code == -3) return fail(ctx, MDMC_ERR_CAPACITY, e.what());
        if (h->code == -4) return fail(ctx, MDMC_ERR_INTERNAL, e.what());
        return fail(ctx, MDMC_ERR_HIP, e.what());
    }
    return fail(ctx, MDMC_ERR_HIP, e.what());
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        (void)hipGetDevice(&prev);
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        (void)hipGetDevice(&cur);
        if (prev >= 0 && cur != prev) (void)hipSetDevice(prev);
    }
};

}  // namespace

extern "C" {

mdmc_status mdmc_create(mdmc_ctx** out, int32_t device) {
    if (!out) return MDMC_ERR_INVALID_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return MDMC_ERR_HIP;
    int dev = device;
    if (dev < 0) (void)hipGetDevice(&dev);
    if (dev >= count || dev >= kMaxDevices) return MDMC_ERR_INVALID_ARG;
    DeviceGuard g(dev);
    {
        // per device, once per process: the grid pass's constant tables, and the address of
        // this device's copy of the no-allocation scratch array
        static std::mutex mu;
        static std::vector<char> ready;
        std::lock_guard<std::mutex> lock(mu);
        if ((int)ready.size() <= dev) ready.resize(dev + 1, 0);
        if (!ready[dev]) {
            try {
                mdmc::initGridPassTables();
            } catch (const std::exception&) {
                return MDMC_ERR_HIP;
            }
            float* ptr = nullptr;
            if (hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return MDMC_ERR_HIP;
            g_noallocScratch[dev] = ptr;
            ready[dev] = 1;
        }
    }
    mdmc_ctx* ctx = new (std::nothrow) mdmc_ctx();
    if (!ctx) return MDMC_ERR_OOM;
    ctx->device = dev;
    if (hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return MDMC_ERR_HIP;
    }
    for (auto& e : ctx->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete ctx;
            return MDMC_ERR_HIP;
        }
    ctx->p.stream = ctx->own;
    *out = ctx;
    return MDMC_OK;
}

void mdmc_destroy(mdmc_ctx* ctx) {
    if (!ctx) return;
    {
        DeviceGuard g(ctx->device);
        if (ctx->p.stream) (void)hipStreamSynchronize(ctx->p.stream);
        ctx->p.arena.release();
        for (auto& e : ctx->ev)
            if (e) (void)hipEventDestroy(e);
        if (ctx->own) (void)hipStreamDestroy(ctx->own);
    }
    delete ctx;
}

const char* mdmc_last_error(const mdmc_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

mdmc_status mdmc_set_stream(mdmc_ctx* ctx, void* s) {
    if (!ctx) return MDMC_ERR_INVALID_ARG;
    ctx->p.stream = s ? (hipStream_t)s : ctx->own;
    return MDMC_OK;
}

static_assert(MDMC_TILE_AUTO == mdmc::GP_SHAPE_AUTO && MDMC_TILE_1024 == mdmc::GP_SHAPE_FIRST &&
                  MDMC_TILE_4096_LINEAR == mdmc::GP_SHAPE_LAST_TILE && MDMC_TILE_SLAB4 == mdmc::GP_SHAPE_SLAB4 &&
                  MDMC_TILE_SLAB2 == mdmc::GP_SHAPE_SLAB2 && MDMC_TILE_2048_QUEUE == mdmc::GP_SHAPE_QUEUE &&
                  MDMC_TILE_SHAPES == mdmc::GP_SHAPE_LAST,
              "tile shape numbering");

mdmc_status mdmc_set_tile_shape(mdmc_ctx* ctx, int32_t shape) {
    if (!ctx) return MDMC_ERR_INVALID_ARG;
    if (shape < MDMC_TILE_AUTO || shape > MDMC_TILE_SHAPES) return fail(ctx, MDMC_ERR_INVALID_ARG, "unknown tile shape");
    ctx->p.tileShape = shape;
    return MDMC_OK;
}

mdmc_status mdmc_get_tile_shape(const mdmc_ctx* ctx, int32_t* last_shape, double* tune_ms) {
    if (!ctx || !last_shape) return MDMC_ERR_INVALID_ARG;
    *last_shape = ctx->p.lastShape;
    if (tune_ms)
        for (int i = MDMC_TILE_1024; i <= MDMC_TILE_SHAPES; ++i) tune_ms[i - MDMC_TILE_1024] = ctx->p.tuneMs[i];
    return MDMC_OK;
}

}  // extern "C"

namespace {

// processed-grid size of one input grid
void processedDims(const mdmc_grid* g, int r, int pad, int od[3]) {
    for (int i = 0; i < 3; ++i) od[i] = g->dims[i] + (g->dims[i] - 1) * r + 2 * pad;
}

// mdmc_extract (G = I = 1) and mdmc_extract_batch: pad / refine every grid, then one
// pipeline over all (grid, iso) problems
mdmc_status runExtraction(mdmc_ctx* ctx, const mdmc_grid* grids, int G, const float* isos, int I, const mdmc_options* opt, mdmc_stats* stats) {
    ctx->hasResult = false;
    ctx->adj.valid = false;
    ctx->procGrids.clear();
    ctx->nProblems = 0;
    if (!grids || !isos || !opt || G < 1 || I < 1) return fail(ctx, MDMC_ERR_INVALID_ARG, "null grids, iso-values or options, or an empty batch");
    if ((int64_t)G * I > (int64_t)1 << 20) return fail(ctx, MDMC_ERR_INVALID_ARG, "more than 2^20 problems");
    if (opt->refinement < 0 || opt->refinement > 64) return fail(ctx, MDMC_ERR_INVALID_ARG, "refinement must be in [0, 64]");
    const int flags = opt->flags;
    const int pad = (flags & MDMC_PAD1) ? 1 : 0;
    const int r = opt->refinement;
    const bool record = (flags & MDMC_RECORD_ADJOINT) != 0;
    if (record && (flags & (MDMC_SURF_ITER | MDMC_P3X3Y_COLOR | MDMC_P3X3Y_OLD | MDMC_P3333 | MDMC_SPLIT_NM)))
        return fail(ctx, MDMC_ERR_INVALID_ARG,
                    "MDMC_RECORD_ADJOINT needs MDMC_SURF_ITER and the vertex-changing post-processes (P3X3Y, P3333, SPLIT_NM) off");
    if (record && G * I != 1) return fail(ctx, MDMC_ERR_INVALID_ARG, "MDMC_RECORD_ADJOINT is for single extractions");
    int64_t Ptot = 0, PinTot = 0, cells = 0;
    std::vector<int64_t> pOff((size_t)G + 1, 0), inOff((size_t)G + 1, 0);
    for (int gi = 0; gi < G; ++gi) {
        const mdmc_grid* grid = &grids[gi];
        if (!grid->data) return fail(ctx, MDMC_ERR_INVALID_ARG, "null grid data");
        for (int i = 0; i < 3; ++i)
            if (grid->dims[i] < 1) return fail(ctx, MDMC_ERR_INVALID_ARG, "grid dims must be >= 1");
        int od[3];
        processedDims(grid, r, pad, od);
        if (od[0] < 2 || od[1] < 2 || od[2] < 2) return fail(ctx, MDMC_ERR_INVALID_ARG, "processed grid needs at least 2 points per axis");
        const int64_t P = (int64_t)od[0] * od[1] * od[2];
        Ptot += P;
        if ((int64_t)3 * Ptot >= ((int64_t)1 << 40)) return fail(ctx, MDMC_ERR_CAPACITY, "grid too large");
        pOff[gi + 1] = Ptot;
        PinTot += grid->data_on_device ? 0 : (int64_t)grid->dims[0] * grid->dims[1] * grid->dims[2];
        inOff[gi + 1] = PinTot;
        cells += (int64_t)I * (od[0] - 1) * (od[1] - 1) * (od[2] - 1);
    }
    DeviceGuard dg(ctx->device);
    mdmc::Pipeline& p = ctx->p;
    const int B = G * I;
    try {
        hipStream_t s = p.stream;
        MDMC_HIP(hipEventRecord(ctx->ev[0], s));
        p.resetCounters();
        p.splitIterations = 0;
        p.extraAnomalies = 0;
        // ---- grids: upload if needed, then pad / refine (refineGrid)
        float* gp = p.arena.get<float>("grid", (size_t)Ptot);
        float* up = PinTot ? p.arena.get<float>("input", (size_t)PinTot) : nullptr;
        std::vector<mdmc::ProblemDev> problems;
        problems.reserve((size_t)B);
        std::vector<mdmc::PrepareDesc> prep((size_t)G);
        int64_t prepBlocks = 0;
        mdmc::GridDev g0{};
        for (int gi = 0; gi < G; ++gi) {
            const mdmc_grid* grid = &grids[gi];
            const float* in = grid->data;
            if (!grid->data_on_device) {
                const int64_t Pin = (int64_t)grid->dims[0] * grid->dims[1] * grid->dims[2];
                MDMC_HIP(hipMemcpyAsync(up + inOff[gi], grid->data, (size_t)Pin * sizeof(float), hipMemcpyHostToDevice, s));
                in = up + inOff[gi];
            }
            int od[3];
            processedDims(grid, r, pad, od);
            prep[gi] = mdmc::PrepareDesc{in, gp + pOff[gi], grid->dims[0], grid->dims[1], grid->dims[2], od[0], od[1], od[2], prepBlocks};
            prepBlocks += mdmc::prepareBlocks(od);
            mdmc::GridDev g;
            g.v = gp + pOff[gi];
            g.nx = od[0];
            g.ny = od[1];
            g.nz = od[2];
            const float f = 1.f / (static_cast<float>(r) + 1.f);
            g.dx = grid->delta[0] * f;
            g.dy = grid->delta[1] * f;
            g.dz = grid->delta[2] * f;
            g.ox = grid->origin[0];
            g.oy = grid->origin[1];
            g.oz = grid->origin[2];
            if (pad) {
                g.ox -= g.dx;
                g.oy -= g.dy;
                g.oz -= g.dz;
            }
            if (gi == 0) g0 = g;
            ctx->procGrids.push_back(g);
            for (int k = 0; k < I; ++k) {
                mdmc::ProblemDev pr{};
                pr.g = g;
                pr.iso = isos[(int64_t)gi * I + k];
                problems.push_back(pr);
            }
        }
        const bool surf = (flags & MDMC_SURF_ITER) != 0;
        mdmc::PrepareDesc* pd = p.arena.get<mdmc::PrepareDesc>("prep_desc", (size_t)G);
        // A batch of grids already on the device, with a tile shape chosen: the shape-determined
        // front (padding, classification, tile-count scans) goes through a graph captured once
        // for these shapes and replayed (stage times then count it under the grid pass)
        const int64_t C = p.dmcSetup(problems, record);
        const int shapeNow = p.resolveShape(C, surf);
        const bool graphable = p.useGraphs && B > 1 && PinTot == 0 && !record && shapeNow != mdmc::GP_SHAPE_AUTO && !mdmc::gridPassIsSlab(shapeNow);
        std::vector<int64_t> shapeSig;
        if (graphable) {
            shapeSig = {G, I, r, pad, surf ? 1 : 0, shapeNow, prepBlocks};
            for (int gi = 0; gi < G; ++gi)
                for (int k = 0; k < 3; ++k) shapeSig.push_back(grids[gi].dims[k]);
        }
        if (graphable && p.runFrontGraph(shapeSig, prep.data(), (size_t)G * sizeof(mdmc::PrepareDesc), pd, G, prepBlocks, r, pad, problems, C,
                                         shapeNow, surf)) {
            MDMC_HIP(hipEventRecord(ctx->ev[1], s));
            p.gridPassBack(C, shapeNow, record);
        } else {
            p.upload(pd, prep.data(), (size_t)G * sizeof(mdmc::PrepareDesc), 1);
            mdmc::launchPrepareGrids(pd, G, prepBlocks, r, pad, p.prepSpecial(), s);
            MDMC_HIP(hipEventRecord(ctx->ev[1], s));
            // ---- DMC
            p.dmc(problems, surf, record);
            if (graphable) p.noteEagerFront(shapeSig);
        }
        if (B > 1) p.problemDmcCounts(B);
        MDMC_HIP(hipEventRecord(ctx->ev[2], s));
        int64_t st9[9] = {0};
        if (p.V > 0) {
            p.simplifyQueue = ((flags & MDMC_P3X3Y_COLOR) ? 1 : 0) + ((flags & MDMC_P3X3Y_OLD) ? 1 : 0) + ((flags & MDMC_P3333) ? 1 : 0);
            p.halfedgeInitialize();
            MDMC_HIP(hipEventRecord(ctx->ev[3], s));
            if (flags & MDMC_P3X3Y_COLOR) p.simplify3X3Y(true);
            if (flags & MDMC_P3X3Y_OLD) p.simplify3X3Y(false);
            if (flags & MDMC_P3333) p.simplify3333();
            MDMC_HIP(hipEventRecord(ctx->ev[4], s));
            p.fixManifoldTwins();
            MDMC_HIP(hipEventRecord(ctx->ev[5], s));
            if (flags & MDMC_SPLIT_NM) p.split(opt->split_step, opt->max_split_iters > 0 ? opt->max_split_iters : 64);
            MDMC_HIP(hipEventRecord(ctx->ev[6], s));
            if (flags & MDMC_DETECT_NM_VERTS) p.detectNonManifoldVertices();
            MDMC_HIP(hipEventRecord(ctx->ev[7], s));
            if (B == 1) {
                p.computeStats(st9);
            } else {
                p.regroupBatch(B);
                for (int b = 0; b < B; ++b)
                    for (int k = 0; k < 9; ++k) st9[k] += p.batchStats[9 * (size_t)b + k];
            }
        } else {
            // DualMarchingCubes::run returns early when no vertex was produced (DualMarchingCubes.cu:257-261)
            p.Q = p.H = p.F = 0;
            if (B > 1) {
                p.batchVOff.assign((size_t)B + 1, 0);
                p.batchHOff.assign((size_t)B + 1, 0);
                p.batchFOff.assign((size_t)B + 1, 0);
                p.batchStats.assign(9 * (size_t)B, 0);
            }
            for (int i = 3; i < 8; ++i) MDMC_HIP(hipEventRecord(ctx->ev[i], s));
        }
        MDMC_HIP(hipEventRecord(ctx->ev[8], s));
        MDMC_HIP(hipEventSynchronize(ctx->ev[8]));
        ctx->hasResult = true;
        ctx->nProblems = B;
        ctx->flags = flags;
        if (record) {
            AdjointState& a = ctx->adj;
            a.valid = true;
            a.g = g0;
            a.iso = isos[0];
            int od[3];
            processedDims(&grids[0], r, pad, od);
            for (int i = 0; i < 3; ++i) {
                a.inDims[i] = grids[0].dims[i];
                a.outDims[i] = od[i];
            }
            a.refinement = r;
            a.pad = pad;
            a.V = p.dmcV;
        }
        if (stats) {
            std::memset(stats, 0, sizeof *stats);
            stats->n_vertices = p.V;
            stats->n_halfedges = p.H;
            stats->n_faces = p.F;
            stats->n_dmc_vertices = p.dmcV;
            stats->n_dmc_quads = p.dmcQ;
            stats->nm_edges4 = st9[0];
            stats->nm_edges3 = st9[1];
            stats->n_edges = st9[2];
            for (int k = 0; k < 5; ++k) stats->quads_with_nm[k] = st9[3 + k];
            stats->nm_vertices = (flags & MDMC_DETECT_NM_VERTS) ? st9[8] : -1;
            stats->split_iterations = p.splitIterations;
            int64_t anomalyCount = 0;
            p.counterTotals(&anomalyCount, &stats->split_sequential_merges, &stats->lookback_fallbacks);
            stats->anomalies = anomalyCount + p.extraAnomalies;
            stats->cells = cells;
            float ms = 0.f;
            double* outs[8] = {&stats->ms_total, &stats->ms_grid, &stats->ms_dmc, &stats->ms_halfedge,
                               &stats->ms_simplify, &stats->ms_fix_twins, &stats->ms_split, &stats->ms_detect};
            MDMC_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[8]));
            *outs[0] = ms;
            for (int i = 1; i < 8; ++i) {
                MDMC_HIP(hipEventElapsedTime(&ms, ctx->ev[i - 1], ctx->ev[i]));
                *outs[i] = ms;
            }
            p.kernelTimes(&stats->ms_grid_pass_kernel);
        }
        return MDMC_OK;
    } catch (const std::exception& e) {
        (void)hipStreamSynchronize(p.stream);
        return fromException(ctx, e);
    }
}

}  // namespace

extern "C" {

mdmc_status mdmc_extract(mdmc_ctx* ctx, const mdmc_grid* grid, float iso, const mdmc_options* opt, mdmc_stats* stats) {
    if (!ctx) return MDMC_ERR_INVALID_ARG;
    return runExtraction(ctx, grid, 1, &iso, 1, opt, stats);
}

mdmc_status mdmc_extract_batch(mdmc_ctx* ctx, const mdmc_grid* grids, int32_t n_grids, const float* isos, int32_t isos_per_grid,
                               const mdmc_options* opt, mdmc_stats* stats) {
    if (!ctx) return MDMC_ERR_INVALID_ARG;
    return runExtraction(ctx, grids, n_grids, isos, isos_per_grid, opt, stats);
}

mdmc_status mdmc_batch_offsets(const mdmc_ctx* ctx, int64_t* vertex_offsets, int64_t* halfedge_offsets, int64_t* face_offsets) {
    if (!ctx) return MDMC_ERR_INVALID_ARG;
    if (!ctx->hasResult) return MDMC_ERR_NO_RESULT;
    const mdmc::Pipeline& p = ctx->p;
    const int B = ctx->nProblems;
    for (int b = 0; b <= B; ++b) {
        if (B == 1) {
            if (vertex_offsets) vertex_offsets[b] = b ? p.V : 0;
            if (halfedge_offsets) halfedge_offsets[b] = b ? p.H : 0;
            if (face_offsets) face_offsets[b] = b ? p.F : 0;
        } else {
            if (vertex_offsets) vertex_offsets[b] = p.batchVOff[b];
            if (halfedge_offsets) halfedge_offsets[b] = p.batchHOff[b];
            if (face_offsets) face_offsets[b] = p.batchFOff[b];
        }
    }
    return MDMC_OK;
}

mdmc_status mdmc_batch_stats(const mdmc_ctx* ctx, mdmc_stats* per_problem) {
    if (!ctx || !per_problem) return MDMC_ERR_INVALID_ARG;
    if (!ctx->hasResult) return MDMC_ERR_NO_RESULT;
    const mdmc::Pipeline& p = ctx->p;
    const int B = ctx->nProblems;
    if (B == 1) return MDMC_ERR_INVALID_ARG;
    for (int b = 0; b < B; ++b) {
        mdmc_stats& st = per_problem[b];
        std::memset(&st, 0, sizeof st);
        st.n_vertices = p.batchVOff[b + 1] - p.batchVOff[b];
        st.n_halfedges = p.batchHOff[b + 1] - p.batchHOff[b];
        st.n_faces = p.batchFOff[b + 1] - p.batchFOff[b];
        st.n_dmc_vertices = p.batchDmc[2 * b];
        st.n_dmc_quads = p.batchDmc[2 * b + 1];
        const int64_t* s9 = &p.batchStats[9 * (size_t)b];
        st.nm_edges4 = s9[0];
        st.nm_edges3 = s9[1];
        st.n_edges = s9[2];
        for (int k = 0; k < 5; ++k) st.quads_with_nm[k] = s9[3 + k];
        st.nm_vertices = (ctx->flags & MDMC_DETECT_NM_VERTS) ? s9[8] : -1;
        st.split_iterations = p.splitIterations;
        st.anomalies = -1;
        st.cells = -1;
        st.split_sequential_merges = -1;
        st.lookback_fallbacks = -1;
    }
    return MDMC_OK;
}

mdmc_status mdmc_backward_noalloc(mdmc_ctx* ctx, const float* dL_dpos, float* dL_dfield, int32_t accumulate) {
    if (!ctx) return MDMC_ERR_INVALID_ARG;
    if (!ctx->hasResult || !ctx->adj.valid)
        return fail(ctx, MDMC_ERR_NO_RESULT, "mdmc_backward_noalloc needs a preceding mdmc_extract with MDMC_RECORD_ADJOINT");
    const AdjointState& a = ctx->adj;
    if (a.V > 0 && !dL_dpos) return fail(ctx, MDMC_ERR_INVALID_ARG, "null dL_dpos");
    if (!dL_dfield) return fail(ctx, MDMC_ERR_INVALID_ARG, "null dL_dfield");
    float* base = g_noallocScratch[ctx->device];
    if (!base) return fail(ctx, MDMC_ERR_INTERNAL, "no-allocation scratch not set up for this device");
    const bool direct = a.pad == 0 && a.refinement == 0;
    const int64_t P = (int64_t)a.outDims[0] * a.outDims[1] * a.outDims[2];
    const int64_t need = 8 * a.V + (direct ? 0 : P);
    if (need > (int64_t)kNoAllocFloats)
        return fail(ctx, MDMC_ERR_CAPACITY, "mdmc_backward_noalloc: 8 floats per dual vertex plus one per padded / refined grid point exceed "
                                            "its scratch of 32 Mi floats (128 MiB); use mdmc_backward");
    DeviceGuard dg(ctx->device);
    mdmc::Pipeline& p = ctx->p;
    hipStream_t s = p.stream;
    const int32_t* map = p.arena.ptr<int32_t>("adj_map");
    const int32_t* cell = p.arena.ptr<int32_t>("gp_vcell");
    // contributions at the start of the scratch, the processed-grid gradient after them (both
    // written in full by their kernels, so nothing from an earlier call needs clearing)
    float* contrib = base;
    mdmc::launchAdjointVertex(a.g, a.iso, map, cell, a.V, dL_dpos, contrib, s);
    if (direct) {
        mdmc::launchAdjointGather(a.g, map, contrib, dL_dfield, accumulate ? 1 : 0, s);
    } else {
        float* dproc = base + 8 * a.V;
        mdmc::launchAdjointGather(a.g, map, contrib, dproc, 0, s);
        mdmc::launchPrepareAdjoint(a.inDims, a.outDims, a.refinement, a.pad, dproc, dL_dfield, accumulate ? 1 : 0, s);
    }
    if (hipGetLastError() != hipSuccess) return fail(ctx, MDMC_ERR_HIP, "mdmc_backward_noalloc: kernel launch failed");
    return MDMC_OK;
}

mdmc_status mdmc_backward(mdmc_ctx* ctx, const float* dL_dpos, float* dL_dfield, int32_t accumulate) {
    if (!ctx) return MDMC_ERR_INVALID_ARG;
    if (!ctx->hasResult || !ctx->adj.valid)
        return fail(ctx, MDMC_ERR_NO_RESULT, "mdmc_backward needs a preceding mdmc_extract with MDMC_RECORD_ADJOINT");
    const AdjointState& a = ctx->adj;
    if (a.V > 0 && !dL_dpos) return fail(ctx, MDMC_ERR_INVALID_ARG, "null dL_dpos");
    if (!dL_dfield) return fail(ctx, MDMC_ERR_INVALID_ARG, "null dL_dfield");
    DeviceGuard dg(ctx->device);
    mdmc::Pipeline& p = ctx->p;
    try {
        hipStream_t s = p.stream;
        const int32_t* map = p.arena.ptr<int32_t>("adj_map");
        const int32_t* cell = p.arena.ptr<int32_t>("gp_vcell");
        float* contrib = p.arena.get<float>("adj_contrib", 8 * (size_t)std::max<int64_t>(a.V, 1));
        mdmc::launchAdjointVertex(a.g, a.iso, map, cell, a.V, dL_dpos, contrib, s);
        if (a.pad == 0 && a.refinement == 0) {
            mdmc::launchAdjointGather(a.g, map, contrib, dL_dfield, accumulate ? 1 : 0, s);
        } else {
            const int64_t P = (int64_t)a.outDims[0] * a.outDims[1] * a.outDims[2];
            float* dproc = p.arena.get<float>("adj_dproc", (size_t)P);
            mdmc::launchAdjointGather(a.g, map, contrib, dproc, 0, s);
            mdmc::launchPrepareAdjoint(a.inDims, a.outDims, a.refinement, a.pad, dproc, dL_dfield, accumulate ? 1 : 0, s);
        }
        MDMC_HIP(hipGetLastError());
        return MDMC_OK;
    } catch (const std::exception& e) {
        (void)hipStreamSynchronize(p.stream);
        return fromException(ctx, e);
    }
}

mdmc_status mdmc_mesh_get_info(const mdmc_ctx* ctx, mdmc_mesh_info* info) {
    if (!ctx || !info) return MDMC_ERR_INVALID_ARG;
    if (!ctx->hasResult) return MDMC_ERR_NO_RESULT;
    info->n_vertices = ctx->p.V;
    info->n_halfedges = ctx->p.H;
    info->n_faces = ctx->p.F;
    return MDMC_OK;
}

mdmc_status mdmc_mesh_copy(const mdmc_ctx* cctx, float* positions, float* normals, int32_t* halfedges, int32_t* face_halfedge, uint16_t* face_attr,
                           uint8_t* vertex_flags, int32_t* vertex_halfedge, int32_t dst_on_device) {
    mdmc_ctx* ctx = const_cast<mdmc_ctx*>(cctx);
    if (!ctx) return MDMC_ERR_INVALID_ARG;
    if (!ctx->hasResult) return fail(ctx, MDMC_ERR_NO_RESULT, "no extraction result");
    DeviceGuard dg(ctx->device);
    try {
        ctx->p.copyOut(positions, normals, halfedges, face_halfedge, face_attr, vertex_flags, vertex_halfedge, dst_on_device != 0);
        return MDMC_OK;
    } catch (const std::exception& e) {
        return fromException(ctx, e);
    }
}

mdmc_status mdmc_mesh_triangles(const mdmc_ctx* cctx, uint32_t* indices, int32_t dst_on_device) {
    mdmc_ctx* ctx = const_cast<mdmc_ctx*>(cctx);
    if (!ctx || !indices) return MDMC_ERR_INVALID_ARG;
    if (!ctx->hasResult) return fail(ctx, MDMC_ERR_NO_RESULT, "no extraction result");
    DeviceGuard dg(ctx->device);
    try {
        ctx->p.triangles(indices, dst_on_device != 0);
        return MDMC_OK;
    } catch (const std::exception& e) {
        return fromException(ctx, e);
    }
}

mdmc_status mdmc_mesh_render_vertices(const mdmc_ctx* cctx, float* vertices, int32_t dst_on_device) {
    mdmc_ctx* ctx = const_cast<mdmc_ctx*>(cctx);
    if (!ctx || !vertices) return MDMC_ERR_INVALID_ARG;
    if (!ctx->hasResult) return fail(ctx, MDMC_ERR_NO_RESULT, "no extraction result");
    DeviceGuard dg(ctx->device);
    try {
        ctx->p.renderVertices(vertices, dst_on_device != 0);
        return MDMC_OK;
    } catch (const std::exception& e) {
        return fromException(ctx, e);
    }
}

mdmc_status mdmc_generate_surface(mdmc_ctx* ctx, int32_t surface_case, const int32_t dims[3], float* values, int32_t dst_on_device,
                                  float origin[3], float delta[3]) {
    if (!ctx || !dims || !values) return MDMC_ERR_INVALID_ARG;
    if (surface_case < MDMC_SURFACE_SPHERE || surface_case > MDMC_SURFACE_STEINER_ROMAN)
        return fail(ctx, MDMC_ERR_INVALID_ARG, "unknown surface case");
    if (dims[0] < 2 || dims[1] < 2 || dims[2] < 2) return fail(ctx, MDMC_ERR_INVALID_ARG, "grid dimensions below 2");
    DeviceGuard dg(ctx->device);
    try {
        // generateVolume: the domain [-1, 1]^3, delta = (max - min) / (dims - 1), origin = min
        float o[3], d[3];
        for (int k = 0; k < 3; ++k) {
            o[k] = -1.f;
            d[k] = (1.f - (-1.f)) / ((float)dims[k] - 1.f);
        }
        const size_t n = (size_t)dims[0] * dims[1] * dims[2];
        float* dst = dst_on_device ? values : ctx->p.arena.get<float>("gen_values", n);
        mdmc::launchFillSurface(dst, dims, surface_case, o, d, ctx->p.stream);
        MDMC_HIP(hipGetLastError());
        if (!dst_on_device) MDMC_HIP(hipMemcpyAsync(values, dst, n * sizeof(float), hipMemcpyDeviceToHost, ctx->p.stream));
        MDMC_HIP(hipStreamSynchronize(ctx->p.stream));
        for (int k = 0; k < 3; ++k) {
            if (origin) origin[k] = o[k];
            if (delta) delta[k] = d[k];
        }
        return MDMC_OK;
    } catch (const std::exception& e) {
        return fromException(ctx, e);
    }
}

mdmc_status mdmc_generate_edge_case(mdmc_ctx* ctx, int32_t edge_case, float* values, int32_t dst_on_device, int32_t dims[3], float origin[3],
                                    float delta[3]) {
    if (!ctx || !values) return MDMC_ERR_INVALID_ARG;
    if (edge_case != MDMC_EDGE_NM_VERTEX0 && edge_case != MDMC_EDGE_NM_VERTEX1) return fail(ctx, MDMC_ERR_INVALID_ARG, "unknown edge case");
    DeviceGuard dg(ctx->device);
    try {
        const int dm[3] = {4, 4, 3};
        float* dst = dst_on_device ? values : ctx->p.arena.get<float>("gen_values", 48);
        mdmc::launchFillEdgeCase(dst, edge_case, ctx->p.stream);
        MDMC_HIP(hipGetLastError());
        if (!dst_on_device) MDMC_HIP(hipMemcpyAsync(values, dst, 48 * sizeof(float), hipMemcpyDeviceToHost, ctx->p.stream));
        MDMC_HIP(hipStreamSynchronize(ctx->p.stream));
        for (int k = 0; k < 3; ++k) {
            if (dims) dims[k] = dm[k];
            if (origin) origin[k] = -1.f;
            if (delta) delta[k] = (1.f - (-1.f)) / ((float)dm[k] - 1.f);
        }
        return MDMC_OK;
    } catch (const std::exception& e) {
        return fromException(ctx, e);
    }
}

mdmc_status mdmc_mesh_face_colors(const mdmc_ctx* cctx, uint8_t* colors, int32_t dst_on_device) {
    mdmc_ctx* ctx = const_cast<mdmc_ctx*>(cctx);
    if (!ctx || !colors) return MDMC_ERR_INVALID_ARG;
    if (!ctx->hasResult) return fail(ctx, MDMC_ERR_NO_RESULT, "no extraction result");
    DeviceGuard dg(ctx->device);
    try {
        ctx->p.faceColours(colors, dst_on_device != 0);
        return MDMC_OK;
    } catch (const std::exception& e) {
        return fromException(ctx, e);
    }
}

mdmc_status mdmc_mesh_triangulate_maxmin(const mdmc_ctx* cctx, uint32_t* indices, float* quality, int32_t dst_on_device) {
    mdmc_ctx* ctx = const_cast<mdmc_ctx*>(cctx);
    if (!ctx || !indices) return MDMC_ERR_INVALID_ARG;
    if (!ctx->hasResult) return fail(ctx, MDMC_ERR_NO_RESULT, "no extraction result");
    DeviceGuard dg(ctx->device);
    try {
        ctx->p.triangulateMaxMin(indices, quality, dst_on_device != 0);
        return MDMC_OK;
    } catch (const std::exception& e) {
        return fromException(ctx, e);
    }
}

mdmc_status mdmc_set_graphs(mdmc_ctx* ctx, int32_t enable) {
    if (!ctx) return MDMC_ERR_INVALID_ARG;
    ctx->p.useGraphs = enable != 0;
    if (!enable) {
        DeviceGuard dg(ctx->device);
        ctx->p.releaseGraphs();
    }
    return MDMC_OK;
}

mdmc_status mdmc_set_local_merges(mdmc_ctx* ctx, int32_t enable) {
    if (!ctx) return MDMC_ERR_INVALID_ARG;
    ctx->p.localMerges = enable != 0;
    return MDMC_OK;
}

mdmc_status mdmc_set_sequential_merges(mdmc_ctx* ctx, int32_t enable) {
    if (!ctx) return MDMC_ERR_INVALID_ARG;
    ctx->p.sequentialMerges = enable != 0;
    return MDMC_OK;
}

mdmc_status mdmc_get_graph_replays(const mdmc_ctx* ctx, int64_t* replays) {
    if (!ctx || !replays) return MDMC_ERR_INVALID_ARG;
    *replays = ctx->p.graphReplays;
    return MDMC_OK;
}

mdmc_status mdmc_processed_grid(mdmc_ctx* ctx, int32_t grid_index, float* values, int32_t dst_on_device, int32_t dims[3], float origin[3],
                                float delta[3]) {
    if (!ctx) return MDMC_ERR_INVALID_ARG;
    if (!ctx->hasResult) return fail(ctx, MDMC_ERR_NO_RESULT, "no extraction result");
    if (grid_index < 0 || grid_index >= (int32_t)ctx->procGrids.size()) return fail(ctx, MDMC_ERR_INVALID_ARG, "grid index out of range");
    const mdmc::GridDev& g = ctx->procGrids[(size_t)grid_index];
    DeviceGuard dg(ctx->device);
    try {
        if (values) {
            const size_t n = (size_t)g.nx * g.ny * g.nz;
            MDMC_HIP(hipMemcpyAsync(values, g.v, n * sizeof(float), dst_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->p.stream));
            MDMC_HIP(hipStreamSynchronize(ctx->p.stream));
        }
        const int dm[3] = {g.nx, g.ny, g.nz};
        const float o[3] = {g.ox, g.oy, g.oz}, d[3] = {g.dx, g.dy, g.dz};
        for (int k = 0; k < 3; ++k) {
            if (dims) dims[k] = dm[k];
            if (origin) origin[k] = o[k];
            if (delta) delta[k] = d[k];
        }
        return MDMC_OK;
    } catch (const std::exception& e) {
        return fromException(ctx, e);
    }
}

}  // extern "C"
