"""The reference's golden data (Data-and-Results/) as test cases.

Settings (inferred from the meshes, confirmed by the oracle reproducing all of them):
  0_No-Fixes            no padding, no splitting
  1_Grid-Extension-Fix  pad by one, no splitting
  2_One-Iteration       pad by one, split up to MoveVerticesToSurface (one iteration)
  3_Fully-Fixed         pad by one, full iterative split
The file 1_Grid-Extension-Fix/4-Edge.obj holds the split mesh (it equals the
2_One-Iteration result face for face), so that case is compared with splitting on.
"""
import os

import numpy as np

import oracle_api as oa

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "data_and_results")

PROBLEMS = {  # obj name -> (bin name, iso)
    "1-Edge": ("1-Edge", 900.0),
    "2-Edge": ("2-Edge", 900.0),
    "3-Edge": ("3-Edge", 900.0),
    "4-Edge": ("4-Edge", 900.0),
    "Non-Manifold-Vertex": ("Non-Manifold-Vertex", 900.0),
    "Iterative-Fix": ("Iterative-Fix", 900.0),
    "12-Merge": ("12-Merge-iso-241.525", 241.525),
}
BASE = oa.SURF_ITER | oa.DETECT_NM_VERTS
SETTINGS = {
    "0_No-Fixes": dict(flags=BASE, split_step=oa.STEP_ALL),
    "1_Grid-Extension-Fix": dict(flags=BASE | oa.PAD1, split_step=oa.STEP_ALL),
    "2_One-Iteration": dict(flags=BASE | oa.PAD1 | oa.SPLIT_NM, split_step=oa.STEP_MOVE),
    "3_Fully-Fixed": dict(flags=BASE | oa.PAD1 | oa.SPLIT_NM, split_step=oa.STEP_ALL),
}


def cases():
    out = []
    for s in SETTINGS:
        for p in PROBLEMS:
            out.append((s, p))
    return out


def obj_path(setting, problem):
    p = os.path.join(GOLDEN, "Base-Results", setting, problem + ".obj")
    if not os.path.exists(p) and problem == "Iterative-Fix":
        p = os.path.join(GOLDEN, "Base-Results", setting, "Iteration-Fix.obj")
    return p


def settings_for(setting, problem):
    kw = dict(SETTINGS[setting])
    if setting == "1_Grid-Extension-Fix" and problem == "4-Edge":
        kw = dict(SETTINGS["2_One-Iteration"])
    return kw


def load_problem(problem):
    bname, iso = PROBLEMS[problem]
    vals, origin, deltas = oa.read_bin(os.path.join(GOLDEN, "Base-Problems", bname + ".bin"))
    return vals, origin, deltas, iso


def faces_from(he, face_he):
    h = np.asarray(face_he).astype(np.int64)
    out = np.empty((len(h), 4), np.int64)
    for i in range(4):
        out[:, i] = he[h, 0]
        h = he[h, 2]
    return out


# Normals in the fixtures are printed with 6 significant digits after FP32 arithmetic the
# reference's compiler may contract into FMAs; 5e-5 absolute covers both (the largest
# difference seen over all 28 fixtures is 2.3e-5).
NORMAL_TOL = 5e-5


def compare_with_obj(pos, faces, objfile, tol=2e-5, nrm=None):
    """Vertices matched by position (the .obj prints 6 significant digits), faces as
    oriented cycles; with nrm, every referenced vertex's normal against the file's vn line
    (NORMAL_TOL absolute).  Returns a dict of mismatch counts (all zero on success)."""
    rv, rn, rf = oa.read_obj(objfile)
    res = {"n_vertices": len(pos) - len(rv), "n_faces": len(faces) - len(rf), "bad_position": 0, "face_mismatch": 0}
    if nrm is not None:
        res["bad_normal"] = 0
        n64 = np.asarray(nrm, np.float64)
    used = np.unique(rf)
    mp = {}
    p64 = np.asarray(pos, np.float64)
    for r in used:
        d = np.abs(p64 - rv[r]).max(axis=1)
        j = int(np.argmin(d))
        if d[j] > tol * max(1.0, float(np.abs(rv[r]).max())):
            res["bad_position"] += 1
        if nrm is not None and np.abs(n64[j] - rn[r]).max() > NORMAL_TOL:
            res["bad_normal"] += 1
        mp[int(r)] = j

    def cyc(q):
        q = [int(x) for x in q]
        k = q.index(min(q))
        return tuple(q[k:] + q[:k])

    ref = {cyc([mp[int(x)] for x in q]) for q in rf}
    got = {cyc(q) for q in faces}
    res["face_mismatch"] = len(ref ^ got)
    res["unreferenced_delta"] = (len(pos) - len(np.unique(faces))) - (len(rv) - len(used))
    return res
