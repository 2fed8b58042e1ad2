"""Pins the oracle's vertex-valence simplifications with a second, independent
restatement written here in numpy from the reference kernels:

  simplify3X3Y (colour)   HalfedgeMesh.cu:296-339, :368-449, :500-538, :859-923
  simplify3X3YOld         HalfedgeMesh.cu:341-366, :455-498, :925-990
  simplify3333            HalfedgeMesh.cu:547-683, :992-1048
  quad attributes         DeviceVectors.h:334-471

Starting from the oracle's own DMC mesh (no post-process), each simplification is
applied here -- valence, pattern marks, merges (position = y + 0.5 (w - y) in float32,
normal renormalised as glm does), vertex / quad compaction with the reference's output
attributes, and the re-initialised halfedge mesh -- and the result must equal the
oracle's mesh for that flag set array for array.  Inputs are closed manifold surfaces
(Gaussian blobs inside the box, noisy gyroids cut off by a ball), where every halfedge
has a twin: the reference reads out of bounds on a missing twin in
getNeighboringFaces (HalfedgeVectors.h:104-122).
"""
import numpy as np
import pytest

import oracle_api as oa

NM, PATTERN, REMOVE, BOUNDARY, INVALID_COLOR = 0x20, 0x40, 0x80, 0x100, 0x1F


class Mesh:
    """Quads (each quad q owns halfedges 4q+i starting at quad[i]), attributes, vertices."""

    def __init__(self, quads, attr, pos, nrm):
        self.quads = np.asarray(quads, np.int64)
        self.attr = np.asarray(attr, np.int64)
        self.pos = np.asarray(pos, np.float32).copy()
        self.nrm = np.asarray(nrm, np.float32).copy()
        F = len(self.quads)
        self.origin = self.quads.reshape(-1)
        self.nxt = (np.arange(4 * F) // 4) * 4 + (np.arange(4 * F) % 4 + 1) % 4
        dest = self.origin[self.nxt]
        key = np.minimum(self.origin, dest) * (int(self.origin.max()) + 1 if F else 1) + np.maximum(self.origin, dest)
        order = np.argsort(key, kind="stable")
        ks = key[order]
        starts = np.r_[True, ks[1:] != ks[:-1]]
        run = np.diff(np.r_[np.nonzero(starts)[0], len(ks)])
        if F and (run != 2).any():
            raise ValueError("not a closed 2-manifold")
        twin = np.empty(4 * F, np.int64)
        twin[order[0::2]] = order[1::2]
        twin[order[1::2]] = order[0::2]
        self.twin = twin
        V = len(self.pos)
        self.valence = np.bincount(self.origin, minlength=V)  # every halfedge has a twin: no boundary term

    def neighbours(self, f):
        return [int(self.twin[4 * f + i]) // 4 for i in range(4)]


def normalize32(v):
    d = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]
    inv = np.float32(1.0) / np.sqrt(d)
    return v * inv


def merge(m, x, y, z, w, vtype, target):
    p = m.pos[y] + np.float32(0.5) * (m.pos[w] - m.pos[y])
    n = normalize32(m.nrm[y] + np.float32(0.5) * (m.nrm[w] - m.nrm[y]))
    m.pos[x] = p
    m.nrm[x] = n
    vtype[z] = 1
    target[z] = x


def compact(m, attr, vtype, target, p3333):
    keepV = vtype == 0
    address = np.cumsum(keepV) - 1
    quads, attrs = [], []
    for f in range(len(m.quads)):
        a = int(attr[f])
        if a & REMOVE:
            continue
        q = m.quads[f]
        if p3333:
            nq = [address[target[v]] if (a & PATTERN) else address[v] for v in q]
        else:
            nq = [address[target[v]] if vtype[v] else address[v] for v in q]
        quads.append(nq)
        attrs.append(INVALID_COLOR | (a & NM) | (a & BOUNDARY))
    return Mesh(np.array(quads, np.int64).reshape(-1, 4), attrs, m.pos[keepV], m.nrm[keepV])


def flag1(v):
    return v[0] == 3 and v[1] >= 5 and v[2] == 3 and v[3] >= 5


def flag2(v):
    return v[0] >= 5 and v[1] == 3 and v[2] >= 5 and v[3] == 3


def p3x3y_colour(m):
    attr = m.attr.copy()
    for f in range(len(m.quads)):  # markElementsP3X3Y
        if attr[f] & (NM | BOUNDARY):
            continue
        a = attr[f] & 0x3F
        if flag1(m.valence[m.quads[f]]) or flag2(m.valence[m.quads[f]]):
            a |= PATTERN
        attr[f] = a
    marked = attr.copy()
    vtype = np.zeros(len(m.pos), np.int64)
    target = np.full(len(m.pos), -1, np.int64)
    for f in range(len(m.quads)):  # mergeVerticesP3X3YColor
        if not marked[f] & PATTERN:
            continue
        nb = m.neighbours(f)
        if any(marked[n] & NM for n in nb) or marked[f] & BOUNDARY:
            continue
        colour = marked[f] & 0x1F
        if any((marked[n] & PATTERN) and (marked[n] & 0x1F) <= colour for n in nb):
            continue
        q = m.quads[f]
        v = m.valence[q]
        if flag1(v):
            merge(m, q[0], q[1], q[2], q[3], vtype, target)
            attr[f] |= REMOVE
        elif flag2(v):
            merge(m, q[1], q[0], q[3], q[2], vtype, target)
            attr[f] |= REMOVE
    return compact(m, attr, vtype, target, False)


def p3x3y_old(m):
    attr = m.attr.copy()
    count = np.zeros(len(m.pos), np.int64)
    for f in range(len(m.quads)):  # countVerticesP3X3Y
        a = attr[f] & 0x3F
        q = m.quads[f]
        v = m.valence[q]
        if flag1(v):
            count[q[0]] += 1
            count[q[2]] += 1
            a |= PATTERN
        elif flag2(v):
            count[q[1]] += 1
            count[q[3]] += 1
            a |= PATTERN
        attr[f] = a
    marked = attr.copy()
    vtype = np.zeros(len(m.pos), np.int64)
    target = np.full(len(m.pos), -1, np.int64)
    for f in range(len(m.quads)):  # mergeVerticesP3X3Y
        if marked[f] & (NM | BOUNDARY) or not marked[f] & PATTERN:
            continue
        nb = m.neighbours(f)
        if any(marked[n] & NM for n in nb) or marked[f] & BOUNDARY:
            continue
        q = m.quads[f]
        v = m.valence[q]
        if flag1(v) and count[q[0]] == 1 and count[q[2]] == 1:
            merge(m, q[0], q[1], q[2], q[3], vtype, target)
            attr[f] |= REMOVE
        elif flag2(v) and count[q[1]] == 1 and count[q[3]] == 1:
            merge(m, q[1], q[0], q[3], q[2], vtype, target)
            attr[f] |= REMOVE
    return compact(m, attr, vtype, target, False)


def p3333(m):
    attr = m.attr.copy()
    F = len(m.quads)
    for f in range(F):  # the clears of markElementsP3333 (all before any mark)
        if not attr[f] & (NM | BOUNDARY):
            attr[f] &= 0x3F
    before = attr.copy()
    vtype = np.zeros(len(m.pos), np.int64)
    target = np.full(len(m.pos), -1, np.int64)
    for f in range(F):  # markElementsP3333
        if before[f] & (NM | BOUNDARY):
            continue
        q = m.quads[f]
        if any(m.valence[v] != 3 for v in q):
            continue
        nbs, far, nval = [], [], []
        ok = True
        for i in range(4):
            t = int(m.twin[4 * f + i])
            n = t // 4
            if before[n] & (NM | BOUNDARY):
                ok = False
                break
            fv = int(m.origin[m.nxt[m.nxt[t]]])
            if m.valence[fv] == 3:
                ok = False
                break
            nbs.append(n)
            far.append(fv)
            nval.append(int(m.valence[fv]))
        if not ok:
            continue
        if any(nval[i] == 3 and nval[(i + 1) % 4] == 3 for i in range(4)):
            continue
        if all(x == 4 for x in nval):
            continue
        attr[f] |= PATTERN
        for i in range(4):
            attr[nbs[i]] |= REMOVE
            vtype[q[i]] = 1
            target[q[i]] = far[i]
    return compact(m, attr, vtype, target, True)


def oracle_mesh(field, iso, flags):
    o = oa.extract(field, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), iso, flags=flags)
    he = np.asarray(o.he)
    F = len(o.face_he)
    quads = np.empty((F, 4), np.int64)
    h = np.asarray(o.face_he).astype(np.int64)
    for i in range(4):
        quads[:, i] = he[h, 0]
        h = he[h, 2]
    return o, quads


def blobs(n, seed):
    from manifolddmc_amd import grids
    return np.ascontiguousarray(grids.gaussian_blobs(n, n_blobs=8, seed=seed), np.float32)


def masked_gyroid(n, periods, noise, seed):
    """A noisy gyroid kept inside a ball of radius 0.4 (outside it the field is positive),
    so that its iso-0 surface is closed."""
    from manifolddmc_amd import grids
    g = grids.gyroid(n, periods=periods, noise=noise, seed=seed)
    x = np.arange(n) / (n - 1) - 0.5
    r = np.sqrt(x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2)
    return np.ascontiguousarray(np.maximum(g, (r - 0.4) * 20), np.float32)


BASE = oa.SURF_ITER | oa.PAD1
FIELDS = {
    "blobs40-s0": (lambda: blobs(40, 0), 0.5),
    "blobs40-s1": (lambda: blobs(40, 1), 0.5),
    "gyroid40": (lambda: masked_gyroid(40, 2, 0.1, 0), 0.0),
    "gyroid32": (lambda: masked_gyroid(32, 2, 0.1, 4), 0.0),
}
CASES = [(f, s) for f in FIELDS for s in [("colour",), ("old",), ("3333",), ("colour", "old", "3333")]
         if f.startswith("gyroid") or s in (("colour",), ("old",))]


@pytest.mark.parametrize("field_name,steps", CASES)
def test_simplifications_match_independent_restatement(field_name, steps):
    make, iso = FIELDS[field_name]
    field = make()
    o0, q0 = oracle_mesh(field, iso, BASE)
    m = Mesh(q0, np.asarray(o0.face_attr), o0.pos, o0.nrm)
    assert not (m.attr & (NM | BOUNDARY)).any()
    flags = BASE
    for s in steps:
        m = {"colour": p3x3y_colour, "old": p3x3y_old, "3333": p3333}[s](m)
        flags |= {"colour": oa.P3X3Y_COLOR, "old": oa.P3X3Y_OLD, "3333": oa.P3333}[s]
    o1, q1 = oracle_mesh(field, iso, flags)
    assert len(o1.pos) < len(o0.pos)  # something was simplified
    assert np.array_equal(q1, m.quads)
    assert np.array_equal(np.asarray(o1.pos).view(np.uint32), m.pos.view(np.uint32))
    assert np.array_equal(np.asarray(o1.nrm).view(np.uint32), m.nrm.view(np.uint32))
    assert np.array_equal(np.asarray(o1.face_attr).astype(np.int64) & 0x1FF, m.attr & 0x1FF)


def test_restatement_detects_a_planted_mistake():
    """The comparison is sharp: the same restatement with the merged normal left
    unnormalised (a dropped term) no longer matches the oracle."""
    make, iso = FIELDS["gyroid40"]
    field = make()
    o0, q0 = oracle_mesh(field, iso, BASE)
    m = Mesh(q0, np.asarray(o0.face_attr), o0.pos, o0.nrm)
    global normalize32
    keep = normalize32
    normalize32 = lambda v: v  # noqa: E731
    try:
        r = p3x3y_colour(m)
    finally:
        normalize32 = keep
    o1, q1 = oracle_mesh(field, iso, BASE | oa.P3X3Y_COLOR)
    assert np.array_equal(q1, r.quads)
    assert np.array_equal(np.asarray(o1.pos).view(np.uint32), r.pos.view(np.uint32))
    assert not np.array_equal(np.asarray(o1.nrm).view(np.uint32), r.nrm.view(np.uint32))
