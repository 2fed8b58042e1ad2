"""Throughput of the Manifold DMC pipeline on one MI355X.

Workload (default): the 512^3 float32 gyroid sin(x)cos(y)+sin(y)cos(z)+sin(z)cos(x)
with 8 periods per axis plus 0.1*N(0,1) seeded noise, iso 0, pad by one, all
post-processes (P3X3Y colour + old, P3333, non-manifold split, non-manifold
vertex detection).  One step = one full extraction (DualMarchingCubes::run +
runPostProcesses) of the field, which is resident in HBM before timing starts.
Metric: Mcells/s = cells of the processed (padded) grid / seconds per step.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--cpu-planes", type=int, default=16, help="z-planes per slab in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--tile-shape", type=int, default=0, help="grid-pass shape 1-9 (0: autotuned)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES between two device buffers 5 times (reference for the HBM traffic counters) and exit")
    return ap.parse_args()


def profiled_traffic(n, shape):
    """HBM bytes of one grid pass (FETCH_SIZE and WRITE_SIZE, copy-calibrated) for the tile
    shape used, from the committed rocprofv3 --pmc summary of the 512^3 workload; None
    for other sizes or when the summary is missing."""
    path = os.path.join(ROOT, "profiles", "gridpass_gyroid512_traffic.md")
    if n != 512 or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes" and parts[1] == "shape=%d" % shape:
                return int(parts[2])
    return None


def stage_roofline(s, n_in, stats):
    """Per stage: minimal algorithmic HBM bytes (each input read once, each output written once)
    from the extraction's own counts, the stage's mean time, and the fraction of the 8 TB/s
    peak that those bytes over that time reach.  P = points of the padded grid, V/Q = DMC
    vertices / quads, H/F = final halfedges / faces; the halfedge mesh is 16 B per halfedge
    record + 8 B per face (first halfedge, attribute) + 1 B per vertex (flags)."""
    P = (n_in + 2) ** 3
    V, Q = s["n_dmc_vertices"], s["n_dmc_quads"]
    H, F, Vf = s["n_halfedges"], s["n_faces"], s["n_vertices"]
    he_mesh = lambda q, v: 16 * 4 * q + 8 * q + v
    model = {
        "ms_grid": ("pad-by-1: read the field, write the padded grid", 4 * n_in ** 3 + 4 * P),
        "ms_dmc": ("grid pass: read the padded grid, write vertices (pos + normal) and quads", 4 * P + 24 * V + 16 * Q),
        "ms_halfedge": ("halfedge initialize: read quads, write the halfedge mesh", 16 * Q + he_mesh(Q, V)),
        "ms_simplify": ("chain of 3 simplifications compacted once: each step reads the quads, their attributes and the "
                        "corner valences; the chain writes the compacted vertices, quads and their halfedge mesh",
                        3 * (20 * Q + 4 * V) + 16 * Q + he_mesh(Q, V) + 24 * V),
        "ms_fix_twins": ("twin repair: read the halfedge records", 16 * H),
        "ms_split": ("split: per iteration read the face attributes and halfedge records", s["split_iterations"] * (4 * F + 16 * H)),
        "ms_detect": ("NM-vertex detection: read the halfedge records, write vertex flags", 16 * H + Vf),
    }
    out = {}
    for k, (what, b) in model.items():
        ms = stats[k]
        gbs = b / (ms / 1e3) / 1e9 if ms > 0 else 0.0
        out[k[3:]] = {"ms": round(ms, 4), "bytes": int(b), "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4), "model": what}
    return out


def make_field(n):
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    from manifolddmc_amd import grids

    out = np.empty((n, n, n), np.float32)
    chunk = 16

    def work(z0):
        out[z0:min(n, z0 + chunk)] = grids.gyroid(n, z0=z0, z1=min(n, z0 + chunk))

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        list(ex.map(work, range(0, n, chunk)))
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_threads():
    """Threads for the all-cores run: the process's CPU share (OMP_NUM_THREADS when set, else the
    affinity mask), at most 16."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        n = min(n, int(os.environ["OMP_NUM_THREADS"]))
    return max(1, min(16, n))


def cpu_baseline(n, planes, flags, reps=3):
    """The CPU oracle (the same algorithm, single-threaded C++, -O3) on z-slabs of the same field:
    one slab on one thread, and one slab per thread on all of the process's host threads at
    once (independent extractions, the way a single-threaded extractor uses several cores).
    Best of `reps` for each."""
    from concurrent.futures import ThreadPoolExecutor

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as oa

    from manifolddmc_amd import grids

    threads = _host_threads()
    slabs = [grids.gyroid(n, z0=k * planes, z1=(k + 1) * planes) for k in range(threads)]
    cells = (n + 1) * (n + 1) * (planes + 1)

    def one(k):
        return len(oa.extract(slabs[k], (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0, flags=flags).face_he)

    t1 = []
    for _ in range(reps):
        t0 = time.perf_counter()
        faces = one(0)
        t1.append(time.perf_counter() - t0)
    tn = []
    with ThreadPoolExecutor(max_workers=threads) as ex:
        for _ in range(reps):
            t0 = time.perf_counter()
            list(ex.map(one, range(threads)))
            tn.append(time.perf_counter() - t0)
    single = cells / min(t1) / 1e6
    multi = threads * cells / min(tn) / 1e6
    return {"value": round(multi, 3), "unit": "Mcells/s", "cores": threads, "kind": "port",
            "single_thread_value": round(single, 3), "cpu_model": _cpu_model(),
            "sample": "z-slabs of %d planes of the same %d^3 gyroid (%dx%dx%d each, padded to %d cells), same flags; "
                      "1 thread: one slab, best of %d = %.2f s (%d faces); %d threads: one slab each at once, best of %d = %.2f s"
                      % (planes, n, n, n, planes, cells, reps, min(t1), faces, threads, reps, min(tn))}


def fp64_check(n, planes, flags):
    """FP32 positions of the GPU against the oracle evaluated in FP64 (checker use, in the CPU
    leg) on z-planes 0..planes-1 of the bench field: per stage the largest relative position
    error |p32 - p64|inf / max(|p64|inf, 1) and the largest error in cells (the cell size is 1)."""
    import numpy as np

    import manifolddmc_amd as mdmc
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as oa

    from manifolddmc_amd import grids

    slab = grids.gyroid(n, z0=0, z1=planes)
    out = {"workload": "z-planes 0..%d of the %d^3 gyroid" % (planes - 1, n)}
    d = mdmc.DualMarchingCubes(device=0)
    try:
        for label, fl in (("dual_vertices", mdmc.PAD1 | mdmc.SURF_ITER),
                          ("simplified", mdmc.PAD1 | mdmc.SURF_ITER | mdmc.P3X3Y_COLOR | mdmc.P3X3Y_OLD | mdmc.P3333), ("all_post", flags)):
            d.extract_raw(slab.ctypes.data, (n, n, planes), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0, fl, 0, False)
            g = d.mesh_numpy()
            o = oa.extract(slab, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0, flags=fl, fp64=True)
            if not np.array_equal(g.halfedges, o.he):
                out[label] = {"topology_equal": False}
                continue
            e = np.abs(g.positions.astype(np.float64) - o.pos64)
            rel = e.max(axis=1) / np.maximum(np.abs(o.pos64).max(axis=1), 1.0)
            out[label] = {"topology_equal": True, "vertices": int(len(g.positions)), "max_rel": float(rel.max()),
                          "max_cells": float(e.max()), "normal_max_abs": float(np.abs(g.normals.astype(np.float64) - o.nrm64).max())}
    finally:
        d.close()
    return out


def batch_fields(n_grids=256, isos_per_grid=4, seed=5):
    """256 grids, sides log-uniform in [16, 192] (independent per axis); even ones are sums
    of 4-32 Gaussian blobs (centres U[0.2, 0.8], s U[0.05, 0.15]), odd ones gyroids with 2-12
    periods and 0.05 N(0,1) noise; 4 iso-values per grid, uniform in the field's [10%, 90%]
    quantile range.  Generated on the GPU with torch, resident in HBM."""
    import numpy as np
    import torch

    rng = np.random.default_rng(seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    out = []
    for i in range(n_grids):
        nx, ny, nz = (int(round(float(np.exp(rng.uniform(np.log(16), np.log(192)))))) for _ in range(3))
        ax = [torch.arange(n, device="cuda", dtype=torch.float32) / (n - 1) for n in (nz, ny, nx)]
        z, y, x = torch.meshgrid(*ax, indexing="ij")
        if i % 2 == 0:
            f = torch.zeros((nz, ny, nx), device="cuda")
            for _ in range(int(rng.integers(4, 33))):
                c = rng.uniform(0.2, 0.8, size=3)
                sg = rng.uniform(0.05, 0.15)
                f += torch.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (2 * sg * sg))
        else:
            w = 2 * np.pi * int(rng.integers(2, 13))
            f = (torch.sin(w * x) * torch.cos(w * y) + torch.sin(w * y) * torch.cos(w * z) + torch.sin(w * z) * torch.cos(w * x)
                 + 0.05 * torch.randn((nz, ny, nx), device="cuda", generator=gen))
        f = f.contiguous()
        q = torch.quantile(f.reshape(-1)[:: max(1, f.numel() // 1000000)], torch.tensor([0.1, 0.9], device="cuda")).tolist()
        isos = sorted(rng.uniform(q[0], q[1], size=isos_per_grid).tolist())
        out.append((f, (0.0, 0.0, 0.0), (1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1)), isos))
    return out


def batch_bench(flags):
    """Batched form vs plain per-problem launches on the same 1024 problems."""
    import torch

    import manifolddmc_amd as mdmc

    fields = batch_fields()
    torch.cuda.synchronize()
    grids = [(f.data_ptr(), (f.shape[2], f.shape[1], f.shape[0]), o, d, True) for f, o, d, _ in fields]
    isos = [i for _, _, _, i in fields]
    n_problems = len(fields) * len(isos[0])
    cells = 0
    d = mdmc.DualMarchingCubes(device=0)
    # plain: one mdmc_extract per problem; warm-up over every grid once so that the tile-shape
    # tuning of each workload class happens outside the timed loop
    for (ptr, dims, o, dl, _), iso_list in zip(grids, isos):
        d.extract_raw(ptr, dims, o, dl, iso_list[0], flags, 0, True)
    t0 = time.perf_counter()
    for (ptr, dims, o, dl, _), iso_list in zip(grids, isos):
        for iso in iso_list:
            cells += d.extract_raw(ptr, dims, o, dl, iso, flags, 0, True).cells
    t_plain = time.perf_counter() - t0
    # batched: all problems in one mdmc_extract_batch, with the shape-determined front replayed
    # from a HIP graph (captured on the second call) and without
    reps = 3

    def batch_time(graphs):
        d.set_graphs(graphs)
        st = d.extract_batch_raw(grids, isos, flags)
        st = d.extract_batch_raw(grids, isos, flags)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            st = d.extract_batch_raw(grids, isos, flags)
        return (time.perf_counter() - t0) / reps, st

    t_eager, _ = batch_time(False)
    t_batch, st = batch_time(True)
    replays = d.graph_replays()
    d.close()
    return {"workload": "256 grids, sides log-U[16,192], half 4-32 Gaussian blobs / half gyroids (2-12 periods, 0.05 noise), "
                        "4 iso-values each in the [10%,90%] quantiles; pad-by-1, all post-processes",
            "problems": n_problems, "cells": int(st.cells), "vertices": int(st.n_vertices), "faces": int(st.n_faces),
            "plain_s": round(t_plain, 4), "batched_s": round(t_batch, 4),
            "plain_grids_per_s": round(n_problems / t_plain, 1), "batched_grids_per_s": round(n_problems / t_batch, 1),
            "batched_no_graph_grids_per_s": round(n_problems / t_eager, 1), "graph_replays": replays,
            "speedup": round(t_plain / t_batch, 2), "batched_stages_ms": {k: round(v, 3) for k, v in st.as_dict().items() if k.startswith("ms_")}}


def adjoint_bench(steps=100):
    """Differentiable form: 32-sphere SDF 256^3, iso 0, pad-by-1, dual vertices as edge means;
    loss = mean squared distance of the vertices to the sphere of radius 0.3 at the centre.
    Times the forward (mdmc_extract recording the topology + vertex copy) and forward +
    loss + backward (mdmc_backward), then runs `steps` descent steps with re-extraction."""
    import torch

    from manifolddmc_amd import grids
    from manifolddmc_amd.autograd import dual_marching_cubes

    field_np, origin, delta, iso = grids.workload("spheres256")
    field = torch.from_numpy(field_np).cuda()
    centre = torch.tensor([0.5, 0.5, 0.5], device="cuda")

    def loss_of(f):
        pos, _ = dual_marching_cubes(f, iso, origin, delta)
        r = torch.linalg.norm(pos - centre, dim=1)
        return ((r - 0.3) ** 2).mean()

    def fwd():
        with torch.no_grad():
            return loss_of(field)

    def fwd_bwd():
        f = field.detach().requires_grad_(True)
        l = loss_of(f)
        l.backward()
        return f.grad

    for _ in range(3):
        fwd()
        fwd_bwd()
    torch.cuda.synchronize()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        fwd()
    torch.cuda.synchronize()
    t_f = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        fwd_bwd()
    torch.cuda.synchronize()
    t_fb = (time.perf_counter() - t0) * 1e3 / reps
    f = field.detach().clone()
    losses = []
    for _ in range(steps):
        f.requires_grad_(True)
        l = loss_of(f)
        l.backward()
        losses.append(float(l.detach()))
        with torch.no_grad():  # move field values by at most a quarter cell
            f = f - (0.25 * delta[0]) * f.grad / f.grad.abs().max()
    return {"workload": "32-sphere SDF 256^3, iso 0, pad-by-1, no move-to-surface; loss = mean (|p - c| - 0.3)^2",
            "fwd_ms": round(t_f, 3), "fwd_bwd_ms": round(t_fb, 3), "ratio": round(t_fb / t_f, 3),
            "descent_steps": steps, "loss_first": losses[0], "loss_last": losses[-1]}


def main():
    args = parse()
    if args.copy_calibration:
        import torch
        a = torch.empty(args.copy_calibration // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 4, "copies": 5}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only; --gpus %d is not supported (run with --gpus 1)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import manifolddmc_amd as mdmc

    n = args.size
    print("generating %d^3 gyroid" % n, file=sys.stderr, flush=True)
    field = make_field(n)
    dev = torch.device("cuda:0")
    t_field = torch.from_numpy(field).to(dev)
    torch.cuda.synchronize()
    flags = mdmc.PAD1 | mdmc.SURF_ITER | mdmc.ALL_POST
    d = mdmc.DualMarchingCubes(device=0)
    if args.tile_shape:
        d.set_tile_shape(args.tile_shape)

    def step():
        return d.extract_raw(t_field.data_ptr(), (n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0, flags, 0, True)

    for i in range(args.warmup):
        st = step()
        print("warmup %d: %.1f ms" % (i, st.ms_total), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(step().as_dict())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    shape, tune_ms = d.tile_shape()
    ms = wall * 1e3 / args.steps
    s = stats[-1]
    cells = s["cells"]
    value = cells / (ms / 1e3) / 1e6
    # dominant kernel: the grid pass (classification, dual vertices, quads)
    P = (n + 2) ** 3
    alg_bytes = 4 * P + 24 * s["n_dmc_vertices"] + 16 * s["n_dmc_quads"]
    gen_ms = float(np.mean([x["ms_grid_pass_kernel"] for x in stats]))
    achieved = alg_bytes / (gen_ms / 1e3) / 1e9
    stage_keys = ["ms_grid", "ms_dmc", "ms_halfedge", "ms_simplify", "ms_fix_twins", "ms_split", "ms_detect", "ms_total"]
    out = {
        "metric": "Mcells/s DMC+repair @512^3 fp32",
        "value": round(value, 2),
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "gyroid %d^3, 8 periods, +0.1*N(0,1), iso 0, pad-by-1, all post-processes" % n,
                   "cells": cells, "vertices": s["n_vertices"], "faces": s["n_faces"], "dmc_vertices": s["n_dmc_vertices"],
                   "dmc_quads": s["n_dmc_quads"], "split_iterations": s["split_iterations"], "nm_edges4": s["nm_edges4"],
                   "nm_vertices": s["nm_vertices"]},
        "roofline": {"kernel": "grid pass: kClassify + tile-count scan + kEmit + kVertices", "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(n, shape),
                     "algorithmic_bytes": alg_bytes, "kernel_ms": round(gen_ms, 4),
                     "tile_shape": mdmc.TILE_SHAPES.get(shape, shape), "tune_ms": {mdmc.TILE_SHAPES[k]: round(v, 4) for k, v in tune_ms.items()}},
        "stages_ms": {k: round(float(np.mean([x[k] for x in stats])), 3) for k in stage_keys},
    }
    out["stages_roofline"] = stage_roofline(s, n, {k: float(np.mean([x[k] for x in stats])) for k in stage_keys})
    out["anomalies"] = s["anomalies"]
    out["split_sequential_merges"] = s["split_sequential_merges"]
    out["lookback_fallbacks"] = max(x["lookback_fallbacks"] for x in stats)
    d.close()
    if not args.no_batch:
        print("batched", file=sys.stderr, flush=True)
        out["batched"] = batch_bench(flags)
    if not args.no_adjoint:
        print("adjoint", file=sys.stderr, flush=True)
        out["adjoint"] = adjoint_bench()
    if not args.no_cpu:
        print("cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_planes, flags)
        print("fp64 check", file=sys.stderr, flush=True)
        out["fp64_error"] = fp64_check(n, 32, flags)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
