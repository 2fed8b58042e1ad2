"""mdmc_cli on the GPU: the .obj it writes for each reference problem matches the
reference's own .obj (Data-and-Results) and equals, byte for byte, what the Python
binding writes for the same extraction."""
import json
import os
import subprocess

import numpy as np
import pytest

import fixtures
import oracle_api as oa

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "manifolddmc_amd", "mdmc_cli")

def cli_args(setting, problem):
    kw = fixtures.settings_for(setting, problem)
    args = []
    if kw["flags"] & oa.PAD1:
        args.append("--pad")
    if not kw["flags"] & oa.SPLIT_NM:
        args.append("--no-split")
    args += ["--split-step", str(kw["split_step"])]
    return args


@pytest.mark.parametrize("setting,problem", [("0_No-Fixes", "2-Edge"), ("1_Grid-Extension-Fix", "Non-Manifold-Vertex"),
                                             ("2_One-Iteration", "4-Edge"), ("3_Fully-Fixed", "Iterative-Fix"),
                                             ("3_Fully-Fixed", "12-Merge")])
def test_cli_obj_matches_reference_and_python(tmp_path, setting, problem):
    import manifolddmc_amd as mdmc
    bname, iso = fixtures.PROBLEMS[problem]
    src = os.path.join(fixtures.GOLDEN, "Base-Problems", bname + ".bin")
    out = str(tmp_path / "mesh.obj")
    cmd = [CLI, src, "--iso", repr(iso), "--out", out] + cli_args(setting, problem)
    print("case", setting, problem, " ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    st = json.loads(r.stdout)
    # against the reference's .obj: positions to its 6 printed digits, faces as oriented cycles
    rv, rn, rf = oa.read_obj(out)
    res = fixtures.compare_with_obj(rv, rf, fixtures.obj_path(setting, problem), nrm=rn)
    assert res == {"n_vertices": 0, "n_faces": 0, "bad_position": 0, "bad_normal": 0, "face_mismatch": 0, "unreferenced_delta": 0}, res
    assert st["vertices"] == len(rv) and st["faces"] == len(rf)
    # against the Python binding on the same input
    vals, origin, deltas, _ = fixtures.load_problem(problem)
    kw = fixtures.settings_for(setting, problem)
    d = mdmc.DualMarchingCubes()
    try:
        d.splitting_step = kw["split_step"]
        v = np.ascontiguousarray(vals, np.float32)
        d.extract_raw(v.ctypes.data, (v.shape[2], v.shape[1], v.shape[0]), origin, deltas, iso, kw["flags"], 0, False)
        py = str(tmp_path / "py.obj")
        d.mesh_numpy().write_obj(py)
    finally:
        d.close()
    assert open(out).read() == open(py).read()
