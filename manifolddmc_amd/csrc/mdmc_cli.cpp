// mdmc_cli -- command-line front end of libmdmc.  Reads a grid in the reference's .bin
// layout, runs the DMC pipeline on the GPU through the C ABI (include/mdmc.h) and writes
// the mesh in the reference's .obj layout.
//
//   .bin  int32 dims[3], float32 deltas[3], then uint16 values, x fastest
//         (UniformGridHost::readFromFile / writeSelectionToFile, UniformGrid.inl:512-571,
//         :639-695); the origin is -0.5 * (dims - 1) * deltas, as readFromFile sets it
//   .obj  "v x y z" lines, "vn x y z" lines, then "f a//a b//b c//c d//d" per quad,
//         1-based, the quad's vertices in triangle-buffer order (Mesh::writeToFile,
//         Mesh.cu:446-506; HalfedgeMesh.cu:690-714)
//
// Host code only: no HIP in this file.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mdmc.h"

namespace {

struct Grid {
    int32_t dims[3] = {0, 0, 0};
    float deltas[3] = {0.f, 0.f, 0.f};
    std::vector<uint16_t> raw;
};

bool readBin(const std::string& path, Grid& g, std::string& err) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    bool ok = std::fread(g.dims, sizeof(int32_t), 3, f) == 3 && std::fread(g.deltas, sizeof(float), 3, f) == 3;
    if (ok && (g.dims[0] < 2 || g.dims[1] < 2 || g.dims[2] < 2)) {
        err = "grid dimensions below 2 in " + path;
        ok = false;
    } else if (ok) {
        const size_t n = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
        g.raw.resize(n);
        ok = std::fread(g.raw.data(), sizeof(uint16_t), n, f) == n;
        if (!ok) err = "short read of the values in " + path;
    } else {
        err = "short header in " + path;
    }
    std::fclose(f);
    return ok;
}

bool writeBin(const std::string& path, const Grid& g) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = std::fwrite(g.dims, sizeof(int32_t), 3, f) == 3 && std::fwrite(g.deltas, sizeof(float), 3, f) == 3 &&
              std::fwrite(g.raw.data(), sizeof(uint16_t), g.raw.size(), f) == g.raw.size();
    return std::fclose(f) == 0 && ok;
}

// Mesh::writeToFile (Mesh.cu:446-506): the render vertex buffer (10 floats per vertex:
// position, normal, colour; all zero for a vertex no halfedge starts at), then one quad per
// six triangle indices (x, y, z, w = indices 0, 1, 2, 4)
bool writeObj(const std::string& path, const std::vector<float>& vtx, const std::vector<uint32_t>& tri) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    const size_t nv = vtx.size() / 10;
    for (size_t i = 0; i < nv; ++i) std::fprintf(f, "v %g %g %g\n", vtx[10 * i], vtx[10 * i + 1], vtx[10 * i + 2]);
    for (size_t i = 0; i < nv; ++i) std::fprintf(f, "vn %g %g %g\n", vtx[10 * i + 3], vtx[10 * i + 4], vtx[10 * i + 5]);
    for (size_t i = 0; i + 6 <= tri.size(); i += 6) {
        const uint32_t a = tri[i] + 1, b = tri[i + 1] + 1, c = tri[i + 2] + 1, d = tri[i + 4] + 1;
        std::fprintf(f, "f %u//%u %u//%u %u//%u %u//%u\n", a, a, b, b, c, c, d, d);
    }
    return std::fclose(f) == 0;
}

void usage() {
    std::fprintf(stderr,
                 "usage: mdmc_cli GRID.bin [--iso V --out MESH.obj] [options]\n"
                 "  --iso V                 iso-value (extraction runs only when given)\n"
                 "  --out MESH.obj          mesh output (.obj appended when missing)\n"
                 "  --pad                   add the clamped one-cell boundary (readFromFile addZeroBoundary)\n"
                 "  --refine R              trilinear refinement (readFromFile refinement)\n"
                 "  --no-surface-iteration  m_enableSurfaceToVertexIteration = false\n"
                 "  --p3x3y-color --p3x3y-old --p3333   simplifications\n"
                 "  --no-split              m_splitNonManifoldHalfedges = false\n"
                 "  --no-detect             m_detectNonManifoldVertices = false\n"
                 "  --split-step N          ManifoldSplitter step (0-6, 100 = all)\n"
                 "  --max-split-iters N     bound of the split loop (default 64)\n"
                 "  --tile-shape S          grid-pass shape 1-9 (default: autotuned)\n"
                 "  --device D              GPU (default: current)\n"
                 "  --write-grid OUT.bin    write the grid back in the .bin layout\n"
                 "  --quiet                 no statistics on stdout\n");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2 || !std::strcmp(argv[1], "-h") || !std::strcmp(argv[1], "--help")) {
        usage();
        return argc < 2 ? 2 : 0;
    }
    const std::string in = argv[1];
    std::string out, gridOut;
    bool haveIso = false, quiet = false, pad = false;
    float iso = 0.f;
    int refine = 0, device = -1, tileShape = MDMC_TILE_AUTO;
    int flags = MDMC_SURF_ITER | MDMC_SPLIT_NM | MDMC_DETECT_NM_VERTS;
    int splitStep = MDMC_SPLIT_STEP_ALL, maxIters = 64;
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&](const char* what) -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "mdmc_cli: %s needs a value\n", what);
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--iso") {
            iso = std::strtof(next("--iso"), nullptr);
            haveIso = true;
        } else if (a == "--out") out = next("--out");
        else if (a == "--pad") pad = true;
        else if (a == "--refine") refine = std::atoi(next("--refine"));
        else if (a == "--no-surface-iteration") flags &= ~MDMC_SURF_ITER;
        else if (a == "--p3x3y-color") flags |= MDMC_P3X3Y_COLOR;
        else if (a == "--p3x3y-old") flags |= MDMC_P3X3Y_OLD;
        else if (a == "--p3333") flags |= MDMC_P3333;
        else if (a == "--no-split") flags &= ~MDMC_SPLIT_NM;
        else if (a == "--no-detect") flags &= ~MDMC_DETECT_NM_VERTS;
        else if (a == "--split-step") splitStep = std::atoi(next("--split-step"));
        else if (a == "--max-split-iters") maxIters = std::atoi(next("--max-split-iters"));
        else if (a == "--tile-shape") tileShape = std::atoi(next("--tile-shape"));
        else if (a == "--device") device = std::atoi(next("--device"));
        else if (a == "--write-grid") gridOut = next("--write-grid");
        else if (a == "--quiet") quiet = true;
        else {
            std::fprintf(stderr, "mdmc_cli: unknown option %s\n", a.c_str());
            usage();
            return 2;
        }
    }
    if (pad) flags |= MDMC_PAD1;
    Grid g;
    std::string err;
    if (!readBin(in, g, err)) {
        std::fprintf(stderr, "mdmc_cli: %s\n", err.c_str());
        return 1;
    }
    if (!gridOut.empty()) {
        if (gridOut.size() < 4 || gridOut.compare(gridOut.size() - 4, 4, ".bin") != 0) gridOut += ".bin";
        if (!writeBin(gridOut, g)) {
            std::fprintf(stderr, "mdmc_cli: cannot write %s\n", gridOut.c_str());
            return 1;
        }
    }
    if (!haveIso) return 0;
    if (out.empty()) {
        std::fprintf(stderr, "mdmc_cli: --iso needs --out\n");
        return 2;
    }
    if (out.size() < 4 || out.compare(out.size() - 4, 4, ".obj") != 0) out += ".obj";

    // the values as float (static_cast<T>(K), UniformGrid.inl:553-560) and the grid geometry
    std::vector<float> vals(g.raw.size());
    for (size_t i = 0; i < vals.size(); ++i) vals[i] = (float)g.raw[i];
    mdmc_grid grid;
    for (int k = 0; k < 3; ++k) {
        grid.dims[k] = g.dims[k];
        grid.delta[k] = g.deltas[k];
        grid.origin[k] = (-0.5f * (float)(g.dims[k] - 1)) * g.deltas[k];
    }
    grid.data = vals.data();
    grid.data_on_device = 0;

    mdmc_ctx* ctx = nullptr;
    if (mdmc_create(&ctx, device) != MDMC_OK) {
        std::fprintf(stderr, "mdmc_cli: no usable GPU (mdmc_create failed)\n");
        return 1;
    }
    int rc = 0;
    mdmc_options opt;
    std::memset(&opt, 0, sizeof opt);
    opt.flags = flags;
    opt.refinement = refine;
    opt.split_step = splitStep;
    opt.max_split_iters = maxIters;
    mdmc_stats st;
    mdmc_mesh_info info;
    if (mdmc_set_tile_shape(ctx, tileShape) != MDMC_OK || mdmc_extract(ctx, &grid, iso, &opt, &st) != MDMC_OK ||
        mdmc_mesh_get_info(ctx, &info) != MDMC_OK) {
        std::fprintf(stderr, "mdmc_cli: %s\n", mdmc_last_error(ctx));
        rc = 1;
    } else {
        std::vector<float> vtx(10 * (size_t)info.n_vertices);
        std::vector<uint32_t> tri(6 * (size_t)info.n_faces);
        if ((info.n_vertices > 0 && mdmc_mesh_render_vertices(ctx, vtx.data(), 0) != MDMC_OK) ||
            (info.n_faces > 0 && mdmc_mesh_triangles(ctx, tri.data(), 0) != MDMC_OK)) {
            std::fprintf(stderr, "mdmc_cli: %s\n", mdmc_last_error(ctx));
            rc = 1;
        } else if (!writeObj(out, vtx, tri)) {
            std::fprintf(stderr, "mdmc_cli: cannot write %s\n", out.c_str());
            rc = 1;
        } else if (!quiet) {
            std::printf("{\"vertices\": %lld, \"faces\": %lld, \"halfedges\": %lld, \"dmc_vertices\": %lld, \"dmc_quads\": %lld, "
                        "\"edges\": %lld, \"nm_edges4\": %lld, \"nm_edges3\": %lld, \"nm_vertices\": %lld, "
                        "\"quads_with_nm\": [%lld, %lld, %lld, %lld, %lld], \"split_iterations\": %lld, \"ms_total\": %.3f}\n",
                        (long long)st.n_vertices, (long long)st.n_faces, (long long)st.n_halfedges, (long long)st.n_dmc_vertices,
                        (long long)st.n_dmc_quads, (long long)st.n_edges, (long long)st.nm_edges4, (long long)st.nm_edges3,
                        (long long)st.nm_vertices, (long long)st.quads_with_nm[0], (long long)st.quads_with_nm[1],
                        (long long)st.quads_with_nm[2], (long long)st.quads_with_nm[3], (long long)st.quads_with_nm[4],
                        (long long)st.split_iterations, st.ms_total);
        }
    }
    mdmc_destroy(ctx);
    return rc;
}
