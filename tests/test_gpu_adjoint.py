"""The adjoint (mdmc_backward and the torch Function) against FP64 references.

With the topology frozen, dL/dfield for a linear functional L = sum_v g_v . p_v of the
dual vertex positions is checked
  * against central finite differences (h = 1e-3) of the FP64 oracle forward, at
    voxels whose perturbation flips no sign and no decider,
  * against an FP64 reference gradient assembled here by scattering the dual-vertex
    formula over every vertex (cells and contours from the oracle's polygons), with
    padding / refinement applied as dense per-axis weight matrices.
Tolerance: 1e-4 relative to the gradient's scale.

A plain central difference at h = 1e-3 has a truncation error of about (h / dv)^2, where
dv is the field step across a cell edge (dv ~ 4e-3 on the 256^3 sphere field: ~6%), so
the differences are Richardson-extrapolated from h = 1e-3, 5e-4 and 2.5e-4 (error
O(h^6)); the perturbation at h = 1e-3 must keep the topology.
"""
import numpy as np
import pytest

import oracle_api as oa

pytestmark = pytest.mark.gpu

# corner offsets (bit 0 = x, bit 1 = y, bit 2 = z) and edge endpoints / axes (CellIntersection.inl:383)
EDGE_AB = [(0, 1), (1, 3), (2, 3), (0, 2), (4, 5), (5, 7), (6, 7), (4, 6), (0, 4), (1, 5), (3, 7), (2, 6)]
EDGE_AXIS = [0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2]


def corner_index(cx, cy, cz, k, nx, ny):
    return ((cz + ((k >> 2) & 1)) * ny + (cy + ((k >> 1) & 1))) * nx + (cx + (k & 1))


def enumerate_vertices(proc, iso):
    """Dual vertices of a processed float32 grid in (cell, contour) order: list of
    (cell (x, y, z), edges of the contour)."""
    nz, ny, nx = proc.shape
    flat = proc.reshape(-1)
    out = []
    ge = proc >= iso
    cell_any = ge[:-1, :-1, :-1].copy()
    cell_all = ge[:-1, :-1, :-1].copy()
    for k in range(1, 8):
        sl = ge[(k >> 2) & 1:nz - 1 + ((k >> 2) & 1), (k >> 1) & 1:ny - 1 + ((k >> 1) & 1), k & 1:nx - 1 + (k & 1)]
        cell_any |= sl
        cell_all &= sl
    active = np.argwhere(cell_any & ~cell_all)  # (z, y, x), already in linear cell order
    for cz, cy, cx in active:
        s = np.array([flat[corner_index(cx, cy, cz, k, nx, ny)] for k in range(8)], np.float32)
        for contour in oa.cell_polygon(s, iso, False):
            out.append(((int(cx), int(cy), int(cz)), contour))
    return out


def positions64(proc64, verts, iso, origin, delta, subset=None):
    nz, ny, nx = proc64.shape
    flat = proc64.reshape(-1)
    ids = range(len(verts)) if subset is None else subset
    res = {}
    for v in ids:
        (cx, cy, cz), edges = verts[v]
        acc = np.zeros(3)
        for e in edges:
            a, b = EDGE_AB[e]
            sa, sb = flat[corner_index(cx, cy, cz, a, nx, ny)], flat[corner_index(cx, cy, cz, b, nx, ny)]
            t = (iso - sa) / (sb - sa)
            off = np.array([a & 1, (a >> 1) & 1, (a >> 2) & 1], np.float64)
            off[EDGE_AXIS[e]] += t
            acc += off
        res[v] = np.asarray(origin, np.float64) + (np.array([cx, cy, cz], np.float64) + acc / len(edges)) * np.asarray(delta, np.float64)
    return res


def processed_gradient64(proc64, verts, iso, delta, gpos):
    """dL/d(processed grid) for L = sum g_v . p_v, in FP64 (scatter form)."""
    nz, ny, nx = proc64.shape
    flat = proc64.reshape(-1)
    out = np.zeros(flat.size)
    d = np.asarray(delta, np.float64)
    for v, ((cx, cy, cz), edges) in enumerate(verts):
        n = len(edges)
        for e in edges:
            a, b = EDGE_AB[e]
            ia, ib = corner_index(cx, cy, cz, a, nx, ny), corner_index(cx, cy, cz, b, nx, ny)
            den = flat[ib] - flat[ia]
            t = (iso - flat[ia]) / den
            gw = gpos[v, EDGE_AXIS[e]] * d[EDGE_AXIS[e]] / n
            out[ia] += gw * (t - 1.0) / den
            out[ib] += gw * (-t) / den
    return out.reshape(proc64.shape)


def axis_matrix(n_in, refinement, pad):
    """W[p, j]: weight of input coordinate j in processed coordinate p (refineGrid on one
    axis, clamp-duplicated border when padded)."""
    rr = refinement + 1
    un = n_in + (n_in - 1) * refinement
    on = un + 2 * pad
    W = np.zeros((on, n_in))
    for p in range(on):
        u = min(max(p, 1), on - 2) - 1 if pad else p
        f = np.float32(u) / np.float32(rr)
        m = int(np.floor(f))
        dd = float(np.float32(f - np.float32(m)))
        W[p, min(max(m, 0), n_in - 1)] += 1.0 - dd
        W[p, min(max(m + 1, 0), n_in - 1)] += dd
    return W


def input_gradient64(dproc, shape_in, refinement, pad):
    nz, ny, nx = shape_in
    Wz, Wy, Wx = axis_matrix(nz, refinement, pad), axis_matrix(ny, refinement, pad), axis_matrix(nx, refinement, pad)
    g = np.tensordot(Wz.T, dproc, axes=(1, 0))
    g = np.tensordot(Wy.T, g, axes=(1, 1)).transpose(1, 0, 2)
    g = np.tensordot(g, Wx, axes=(2, 0))
    return g


def gpu_forward_backward(field, origin, delta, iso, gpos_fn, pad=True, refinement=0):
    import torch

    import manifolddmc_amd as mdmc
    dmc = mdmc.DualMarchingCubes()
    try:
        t = torch.from_numpy(np.ascontiguousarray(field, np.float32)).cuda()
        dmc.set_stream(torch.cuda.current_stream().cuda_stream)
        nz, ny, nx = field.shape
        flags = mdmc.RECORD_ADJOINT | (mdmc.PAD1 if pad else 0)
        print("adjoint case shape", field.shape, "pad", pad, "refinement", refinement, flush=True)
        dmc.extract_raw(t.data_ptr(), (nx, ny, nz), origin, delta, iso, flags, refinement, True)
        mesh = dmc.mesh_numpy()
        gpos = np.ascontiguousarray(gpos_fn(mesh.positions), np.float32)
        gd = torch.from_numpy(gpos).cuda()
        dfield = torch.empty_like(t)
        dmc.backward(gd.data_ptr() if gpos.size else 0, dfield.data_ptr())
        torch.cuda.synchronize()
        return mesh, gpos, dfield.cpu().numpy().astype(np.float64)
    finally:
        dmc.close()


def richardson(diff, h=1e-3):
    """Central difference diff(h) extrapolated from h, h/2, h/4."""
    d1, d2, d3 = diff(h), diff(h / 2), diff(h / 4)
    r1, r2 = (4 * d2 - d1) / 3, (4 * d3 - d2) / 3
    return (16 * r2 - r1) / 15


def small_field(kind, n, seed):
    from manifolddmc_amd import grids
    if kind == "blobs":
        return grids.gaussian_blobs(n, n_blobs=4, seed=seed), 0.5
    if kind == "spheres":
        return grids.sphere_sdf(n, n_spheres=4, seed=seed, radius_range=(0.15, 0.3)), 0.0
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, size=(n, n, n)).astype(np.float32), 0.0


@pytest.mark.parametrize("kind,n,pad,refinement", [("blobs", 20, True, 0), ("spheres", 18, False, 0), ("blobs", 12, True, 1),
                                                   ("random", 10, True, 0)])
def test_backward_matches_fp64_reference(kind, n, pad, refinement):
    field, iso = small_field(kind, n, 3)
    origin, delta = (0.1, -0.2, 0.3), (0.5, 0.25, 0.75)
    rng = np.random.default_rng(11)
    mesh, gpos, g_gpu = gpu_forward_backward(field, origin, delta, iso, lambda p: rng.normal(size=p.shape), pad, refinement)
    o = oa.extract(field, origin, delta, iso, flags=oa.PAD1 if pad else 0, refinement=refinement, want_grid=True)
    proc = o.grid
    verts = enumerate_vertices(proc, iso)
    assert len(verts) == len(mesh.positions) > 0
    # the enumeration gives the same vertices as the library (ids and positions)
    po = np.asarray(o.grid_origin, np.float64)
    pd = np.asarray(o.grid_delta, np.float64)
    p64 = positions64(proc.astype(np.float64), verts, iso, po, pd)
    P = np.array([p64[v] for v in range(len(verts))])
    assert np.max(np.abs(P - mesh.positions)) <= 1e-5 * max(1.0, np.max(np.abs(P)))
    dproc = processed_gradient64(proc.astype(np.float64), verts, iso, pd, gpos.astype(np.float64))
    g_ref = input_gradient64(dproc, field.shape, refinement, 1 if pad else 0)
    err = np.linalg.norm(g_gpu - g_ref) / np.linalg.norm(g_ref)
    assert err <= 1e-4, err


@pytest.mark.parametrize("kind,n,pad,refinement", [("blobs", 16, True, 0), ("spheres", 14, True, 1), ("random", 9, False, 0)])
def test_backward_matches_oracle_central_differences(kind, n, pad, refinement):
    """Central FD (h = 1e-3) of the FP64 oracle forward at every voxel with a nonzero
    gradient whose perturbation keeps the mesh topology."""
    field, iso = small_field(kind, n, 5)
    origin, delta = (0.0, 0.0, 0.0), (1.0 / (n - 1),) * 3
    rng = np.random.default_rng(12)
    flags = oa.PAD1 if pad else 0
    mesh, gpos, g_gpu = gpu_forward_backward(field, origin, delta, iso, lambda p: rng.normal(size=p.shape), pad, refinement)
    g64 = gpos.astype(np.float64)
    scale = np.max(np.abs(g_gpu))
    candidates = np.argwhere(np.abs(g_gpu) > 1e-3 * scale)
    rng.shuffle(candidates)
    checked = 0
    for z, y, x in candidates[:200]:
        v0 = field[z, y, x]
        if abs(float(v0) - iso) < 1e-2:
            continue
        frozen = [True]

        def diff(h):
            vals, runs = [], []
            for sgn in (1.0, -1.0):
                f = field.copy()
                f[z, y, x] = np.float32(float(v0) + sgn * h)
                vals.append(float(f[z, y, x]))
                runs.append(oa.extract(f, origin, delta, iso, flags=flags, refinement=refinement, fp64=True))
            if runs[0].pos64.shape != gpos.shape or runs[0].he.shape != runs[1].he.shape or not np.array_equal(runs[0].he, runs[1].he):
                frozen[0] = False  # a decider flipped: topology not frozen
                return 0.0
            return (np.sum(g64 * runs[0].pos64) - np.sum(g64 * runs[1].pos64)) / (vals[0] - vals[1])

        fd = richardson(diff)
        if not frozen[0]:
            continue
        assert abs(g_gpu[z, y, x] - fd) <= 1e-4 * scale + 1e-4 * abs(fd), (z, y, x, g_gpu[z, y, x], fd)
        checked += 1
    assert checked >= min(50, len(candidates) // 2), checked


def test_sphere_sdf_256_loss_gradient():
    """The 256^3 union of 32 spheres at iso 0, loss = mean squared distance of the dual
    vertices to the sphere of radius 0.3 at the centre: the GPU adjoint against the FP64
    reference gradient, and against FP64 central differences at 1000 voxels."""
    from manifolddmc_amd import grids
    field, origin, delta, iso = grids.workload("spheres256")
    centre = np.array([0.5, 0.5, 0.5])

    def loss_grad(p):
        p = p.astype(np.float64)
        r = np.linalg.norm(p - centre, axis=1)
        return (2.0 / len(p)) * ((r - 0.3) / r)[:, None] * (p - centre)

    mesh, gpos, g_gpu = gpu_forward_backward(field, origin, delta, iso, loss_grad, True, 0)
    proc = np.pad(field, 1, mode="edge")
    verts = enumerate_vertices(proc, iso)
    assert len(verts) == len(mesh.positions)
    pd = np.asarray(delta, np.float64)
    po = np.asarray(origin, np.float64) - pd
    proc64 = proc.astype(np.float64)
    g64 = gpos.astype(np.float64)
    dproc = processed_gradient64(proc64, verts, iso, pd, g64)
    g_ref = input_gradient64(dproc, field.shape, 0, 1)
    err = np.linalg.norm(g_gpu - g_ref) / np.linalg.norm(g_ref)
    assert err <= 1e-4, err
    # central differences at 1000 voxels away from the iso-value, recomputing only the
    # vertices of the cells around the voxel
    by_cell = {}
    for v, (c, _) in enumerate(verts):
        by_cell.setdefault(c, []).append(v)
    nz, ny, nx = proc.shape
    rng = np.random.default_rng(21)
    cand = np.argwhere(g_gpu != 0)
    rng.shuffle(cand)
    scale = np.max(np.abs(g_gpu))
    checked = 0
    for z, y, x in cand:
        if checked == 1000:
            break
        if min(z, y, x) == 0 or z == field.shape[0] - 1 or y == field.shape[1] - 1 or x == field.shape[2] - 1:
            continue  # border voxels also feed the duplicated padding layer
        pz, py, px = z + 1, y + 1, x + 1
        if abs(proc64[pz, py, px] - iso) < 2e-3:
            continue
        cells = [(px - dx, py - dy, pz - dz) for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)]
        ids = [v for c in cells for v in by_cell.get(c, [])]
        same = True
        for sgn in (1.0, -1.0):
            p32 = proc.copy()
            p32[pz, py, px] = np.float32(proc[pz, py, px] + sgn * 1e-3)
            for c in cells:
                if c not in by_cell:
                    continue
                cx, cy, cz = c
                s = np.array([p32.reshape(-1)[corner_index(cx, cy, cz, k, nx, ny)] for k in range(8)], np.float32)
                if [verts[v][1] for v in by_cell[c]] != oa.cell_polygon(s, iso, False):
                    same = False
        if not same:
            continue
        def diff(h):
            lv = []
            for sgn in (1.0, -1.0):
                q = proc64.copy()
                q[pz, py, px] += sgn * h
                p = positions64(q, verts, iso, po, pd, ids)
                lv.append(sum(float(np.dot(g64[v], p[v])) for v in ids))
            return (lv[0] - lv[1]) / (2 * h)

        fd = richardson(diff)
        assert abs(g_gpu[z, y, x] - fd) <= 1e-4 * scale + 1e-4 * abs(fd), (z, y, x, g_gpu[z, y, x], fd)
        checked += 1
    assert checked == 1000, checked


def test_torch_function_gradient_and_descent():
    import torch

    from manifolddmc_amd import grids
    from manifolddmc_amd.autograd import dual_marching_cubes
    field_np, origin, delta, iso = grids.workload("spheres64")
    field = torch.from_numpy(field_np).cuda().requires_grad_(True)
    centre = torch.tensor([0.5, 0.5, 0.5], device="cuda")

    def loss_of(f):
        pos, quads = dual_marching_cubes(f, iso, origin, delta)
        r = torch.linalg.norm(pos - centre, dim=1)
        return ((r - 0.3) ** 2).mean(), pos, quads

    loss, pos, quads = loss_of(field)
    assert quads.shape[1] == 4 and int(quads.max()) < len(pos)
    loss.backward()
    g_torch = field.grad.detach().cpu().numpy().astype(np.float64)
    # the same gradient through the C ABI directly
    p = pos.detach().cpu().numpy().astype(np.float64)
    r = np.linalg.norm(p - 0.5, axis=1)
    gp = (2.0 / len(p)) * ((r - 0.3) / r)[:, None] * (p - 0.5)
    _, _, g_abi = gpu_forward_backward(field_np, origin, delta, iso, lambda q: gp, True, 0)
    assert np.linalg.norm(g_torch - g_abi) <= 1e-5 * np.linalg.norm(g_abi)
    # a few descent steps with re-extraction lower the loss
    f = field.detach().clone()
    losses = []
    for _ in range(5):
        f.requires_grad_(True)
        l, _, _ = loss_of(f)
        l.backward()
        losses.append(float(l.detach()))
        with torch.no_grad():  # move field values by at most a quarter cell
            f = f - (0.25 * delta[0]) * f.grad / f.grad.abs().max()
    assert losses[-1] < losses[0], losses
