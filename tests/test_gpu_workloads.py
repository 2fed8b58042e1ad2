"""The two benchmark workloads the other GPU tests do not reach at their real size.

Batched form: 256 grids (sides log-uniform in [16, 192] per axis; half Gaussian-blob sums,
half noisy gyroids) x 4 iso-values each = 1024 problems in one mdmc_extract_batch, pad-by-1
and all post-processes -- the exact fields bench.py times.
  * every problem equals its own mdmc_extract, array for array (compared on the GPU);
  * halfedge invariants hold on all 1024 meshes, no edge keeps 4 faces, no anomalies;
  * 16 problems -- the four largest grids and twelve others -- equal the oracle bit for bit.

1024^3 gyroid (8 periods, +0.1 N(0,1), iso 0, pad-by-1, all post-processes), beyond what
the oracle can run here: halfedge invariants (on the GPU), statistics against the arrays,
no edge with 4 faces, no anomalies, and two extractions identical.
"""
import concurrent.futures

import numpy as np
import pytest

import meshcheck
import oracle_api as oa

pytestmark = pytest.mark.gpu

NAMES = ("positions", "normals", "halfedges", "face_halfedge", "face_attr", "vertex_flags", "vertex_halfedge")


def check_halfedges(t, V, F):
    """Halfedge structure of a (possibly concatenated) mesh held in torch tensors."""
    import torch
    he = t["halfedges"].long()
    H = he.shape[0]
    assert H == 4 * F
    if H == 0:
        return
    origin, face, nxt, twin = he[:, 0], he[:, 1], he[:, 2], he[:, 3]
    assert int(origin.min()) >= 0 and int(origin.max()) < V
    assert int(face.min()) >= 0 and int(face.max()) < F
    assert int(nxt.min()) >= 0 and int(nxt.max()) < H
    assert int(twin.min()) >= -1 and int(twin.max()) < H
    idx = torch.arange(H, device=he.device)
    assert torch.equal(nxt[nxt[nxt[nxt]]], idx), "next does not cycle in 4 steps"
    assert torch.equal(face[nxt], face), "next leaves the face"
    fh = t["face_halfedge"].long()
    assert torch.equal(face[fh], torch.arange(F, device=he.device)), "face -> halfedge -> face"
    has = twin >= 0
    tw = twin[has]
    assert torch.equal(twin[tw], idx[has]), "twin is not an involution"
    assert torch.equal(origin[tw], origin[nxt[has]]) and torch.equal(origin[nxt[tw]], origin[has]), "twin runs the same way"
    del he, origin, face, nxt, twin, idx, has, tw


# ------------------------------------------------------------------------------------ batched
@pytest.fixture(scope="module")
def batch():
    import torch

    import bench
    import manifolddmc_amd as mdmc
    print("batched workload: generating 256 grids", flush=True)
    fields = bench.batch_fields()
    torch.cuda.synchronize()
    flags = mdmc.PAD1 | mdmc.SURF_ITER | mdmc.ALL_POST
    grids = [(f.data_ptr(), (f.shape[2], f.shape[1], f.shape[0]), o, d, True) for f, o, d, _ in fields]
    isos = [i for _, _, _, i in fields]
    d = mdmc.DualMarchingCubes()
    print("batched workload: 1024 problems in one batch", flush=True)
    tot = d.extract_batch_raw(grids, isos, flags).as_dict()
    t = d.mesh_torch()
    vo, ho, fo = d.batch_offsets()
    per = [s.as_dict() for s in d.batch_stats()]
    d.close()
    return dict(fields=fields, grids=grids, isos=isos, flags=flags, tot=tot, t=t, vo=vo, ho=ho, fo=fo, per=per)


def problem_slice(b, t, vo, ho, fo):
    """Problem b of the concatenated batch with its own ids."""
    he = t["halfedges"][ho[b]:ho[b + 1]].clone()
    if len(he):
        he[:, 0] -= int(vo[b])
        he[:, 1] -= int(fo[b])
        he[:, 2] -= int(ho[b])
        tw = he[:, 3]
        he[:, 3] = tw.where(tw < 0, tw - int(ho[b]))
    return {"positions": t["positions"][vo[b]:vo[b + 1]], "normals": t["normals"][vo[b]:vo[b + 1]], "halfedges": he,
            "face_halfedge": t["face_halfedge"][fo[b]:fo[b + 1]] - int(ho[b]), "face_attr": t["face_attr"][fo[b]:fo[b + 1]],
            "vertex_flags": t["vertex_flags"][vo[b]:vo[b + 1]], "vertex_halfedge": t["vertex_halfedge"][vo[b]:vo[b + 1]]}


def test_batched_workload_equals_single_extractions(batch):
    import torch

    import manifolddmc_amd as mdmc
    B = len(batch["per"])
    assert B == 1024
    t, vo, ho, fo = batch["t"], batch["vo"], batch["ho"], batch["fo"]
    tot = batch["tot"]
    assert tot["n_vertices"] == vo[-1] and tot["n_halfedges"] == ho[-1] and tot["n_faces"] == fo[-1]
    assert tot["anomalies"] == 0 and tot["nm_edges4"] == 0 and tot["lookback_fallbacks"] == 0
    print("batched workload: halfedge invariants over %d halfedges" % int(ho[-1]), flush=True)
    check_halfedges(t, int(vo[-1]), int(fo[-1]))
    one = mdmc.DualMarchingCubes()
    try:
        b = 0
        for (ptr, dims, o, dl, _), iso_list in zip(batch["grids"], batch["isos"]):
            for iso in iso_list:
                st = one.extract_raw(ptr, dims, o, dl, iso, batch["flags"], 0, True).as_dict()
                s = one.mesh_torch()
                g = problem_slice(b, t, vo, ho, fo)
                bad = [k for k in NAMES if not torch.equal(g[k], s[k])]
                assert not bad, ("problem %d, grid %s, iso %r" % (b, dims, iso), bad)
                assert st["anomalies"] == 0 and st["nm_edges4"] == 0, (b, st)
                ps = batch["per"][b]
                for k in ("n_vertices", "n_halfedges", "n_faces", "n_dmc_vertices", "n_dmc_quads", "nm_edges4", "nm_edges3",
                          "n_edges", "quads_with_nm", "nm_vertices"):
                    assert ps[k] == st[k], (b, k, ps[k], st[k])
                b += 1
                if b % 256 == 0:
                    print("batched workload: %d problems equal their single extraction" % b, flush=True)
    finally:
        one.close()


def test_batched_workload_problems_match_oracle(batch):
    fields, isos = batch["fields"], batch["isos"]
    cells = np.array([(f.shape[0] + 1) * (f.shape[1] + 1) * (f.shape[2] + 1) for f, _, _, _ in fields])
    largest = [int(g) for g in np.argsort(-cells, kind="stable")[:4]]
    rng = np.random.default_rng(11)
    others = [int(g) for g in rng.choice([g for g in range(len(fields)) if g not in largest], size=12, replace=False)]
    picks = [(g, 0) for g in largest] + [(g, int(rng.integers(4))) for g in others]
    t, vo, ho, fo = batch["t"], batch["vo"], batch["ho"], batch["fo"]

    def run_oracle(gk):
        g, k = gk
        f, origin, delta, _ = fields[g]
        return oa.extract(f.cpu().numpy(), origin, delta, isos[g][k], flags=batch["flags"])

    print("batched workload: oracle on %d problems (largest grids %s)" % (len(picks), [tuple(fields[g][0].shape) for g in largest]),
          flush=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as ex:
        results = list(ex.map(run_oracle, picks))
    for (g, k), o in zip(picks, results):
        b = 4 * g + k
        s = {n: v.cpu().numpy() for n, v in problem_slice(b, t, vo, ho, fo).items()}
        errs = []
        for name, ref in (("positions", o.pos), ("normals", o.nrm), ("halfedges", o.he), ("face_halfedge", o.face_he),
                          ("face_attr", o.face_attr.view(np.int16)), ("vertex_flags", o.vflags), ("vertex_halfedge", o.vhe)):
            e = meshcheck.exact_diff(name, s[name], ref)
            if e:
                errs.append(e)
        assert not errs, ("problem %d (grid %d %s, iso %r)" % (b, g, tuple(fields[g][0].shape), isos[g][k]), errs)
        assert o.stats["anomalies"] == 0


# ------------------------------------------------------------------------------------ 1024^3
def test_gyroid_1024_structure_statistics_determinism():
    import torch

    import bench
    import manifolddmc_amd as mdmc
    n = 1024
    print("gyroid 1024^3: generating the field", flush=True)
    field = torch.from_numpy(bench.make_field(n)).cuda()
    torch.cuda.synchronize()
    flags = mdmc.PAD1 | mdmc.SURF_ITER | mdmc.ALL_POST
    d = mdmc.DualMarchingCubes()
    try:
        print("gyroid 1024^3: extraction 1", flush=True)
        st = d.extract_raw(field.data_ptr(), (n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0, flags, 0, True).as_dict()
        del field
        t = d.mesh_torch()
        V, H, F = d.info()
        assert st["cells"] == (n + 1) ** 3
        assert H == st["n_halfedges"] and V == st["n_vertices"] and F == st["n_faces"]
        assert st["anomalies"] == 0 and st["nm_edges4"] == 0 and st["nm_vertices"] == 0 and st["lookback_fallbacks"] == 0
        assert st["n_dmc_vertices"] > 100_000_000  # beyond the reference's 20 M capacities
        print("gyroid 1024^3: %d vertices, %d faces; halfedge invariants" % (V, F), flush=True)
        check_halfedges(t, V, F)
        print("gyroid 1024^3: extraction 2", flush=True)
        field = torch.from_numpy(bench.make_field(n)).cuda()
        st2 = d.extract_raw(field.data_ptr(), (n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0, flags, 0, True).as_dict()
        del field
        t2 = d.mesh_torch()
        bad = [k for k in NAMES if not torch.equal(t[k], t2[k])]
        assert not bad, bad
        assert {k: v for k, v in st.items() if not k.startswith("ms_")} == {k: v for k, v in st2.items() if not k.startswith("ms_")}
    finally:
        d.close()
