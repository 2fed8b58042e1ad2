"""Structural checks and comparisons of halfedge meshes (numpy)."""
import numpy as np


def invariants(he, face_he, n_vertices):
    """Errors of the halfedge structure (empty list = consistent)."""
    errs = []
    H = len(he)
    if H == 0:
        return errs
    o, f, n, t = he[:, 0], he[:, 1], he[:, 2], he[:, 3]
    if o.min() < 0 or o.max() >= n_vertices:
        errs.append("origin out of range")
    if n.min() < 0 or n.max() >= H:
        errs.append("next out of range")
        return errs
    # next cycles of length 4 with a common face
    h4 = n[n[n[n]]]
    if not np.array_equal(h4, np.arange(H)):
        errs.append("next cycle length != 4")
    if not (np.array_equal(f[n], f)):
        errs.append("face differs along next")
    fh = np.asarray(face_he)
    if not np.array_equal(f[fh], np.arange(len(fh))):
        errs.append("face_halfedge does not point into its face")
    tw = t >= 0
    if tw.any():
        tt = t[tw]
        if tt.max() >= H:
            errs.append("twin out of range")
            return errs
        if not np.array_equal(t[tt], np.nonzero(tw)[0]):
            errs.append("twin not an involution")
        # twin runs the other way: origin(twin) == origin(next)
        if not np.array_equal(o[tt], o[n[np.nonzero(tw)[0]]]):
            errs.append("twin orientation")
    return errs


def exact_diff(name, a, b):
    """Index and values of the first difference of two arrays compared bitwise, or None."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return "%s: shape %s vs %s" % (name, a.shape, b.shape)
    if a.dtype.kind == "f":
        av = a.view(np.uint32 if a.dtype == np.float32 else np.uint64)
        bv = b.view(av.dtype)
    else:
        av, bv = a, b
    bad = np.nonzero((av != bv).reshape(len(a), -1).any(axis=1))[0] if a.ndim > 0 and len(a) else []
    if len(bad) == 0:
        return None
    i = int(bad[0])
    return "%s: %d rows differ, first at %d: %s vs %s" % (name, len(bad), i, a[i], b[i])
