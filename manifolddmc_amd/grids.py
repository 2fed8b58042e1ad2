"""Grid I/O and the seeded synthetic fields of the benchmark workloads.

Grids are float32 arrays laid out [z][y][x] (x fastest), the layout of
UniformGrid (UniformGrid.h:50-52).  Every generator is a pure function of its
seed and of the point index, so any z-slab of a field can be produced on its
own and is bit-identical to the same slab of the full field.
"""
import struct

import numpy as np


# ---------------------------------------------------------------------------- reference file format
def read_bin(path):
    """UniformGridHost::readFromFile (UniformGrid.inl:512-571): int32 dims[3], float32 deltas[3],
    then uint16 values x fastest.  Returns (values, origin, deltas) with the reference's centred
    origin -0.5 * (dims - 1) * deltas, computed in float32."""
    with open(path, "rb") as f:
        b = f.read()
    dims = struct.unpack("<3i", b[:12])
    deltas = np.array(struct.unpack("<3f", b[12:24]), dtype=np.float32)
    n = dims[0] * dims[1] * dims[2]
    vals = np.frombuffer(b[24:24 + 2 * n], dtype="<u2").astype(np.float32)
    vals = np.where(np.isnan(vals), np.float32(0), vals)
    origin = (np.float32(-0.5) * (np.array(dims, dtype=np.int32) - 1).astype(np.float32)) * deltas
    return vals.reshape(dims[2], dims[1], dims[0]), origin.astype(np.float32), deltas


def write_bin(path, values, deltas):
    """The same format (UniformGrid.inl:639-695 writes uint16 samples)."""
    v = np.asarray(values)
    nz, ny, nx = v.shape
    with open(path, "wb") as f:
        f.write(struct.pack("<3i", nx, ny, nz))
        f.write(struct.pack("<3f", *[float(d) for d in deltas]))
        f.write(np.clip(np.rint(v), 0, 65535).astype("<u2").tobytes())


# ---------------------------------------------------------------------------- counter-based RNG
_M = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M
    return z ^ (z >> np.uint64(31))


def _uniform01(seed, idx, stream):
    """(0,1] doubles from a SplitMix64 hash of (seed, stream, index)."""
    with np.errstate(over="ignore"):
        key = _splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream))
        h = _splitmix64(idx.astype(np.uint64) ^ key)
    return ((h >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)


def normal(seed, idx):
    """N(0,1) per index by Box-Muller on two counter-based uniforms."""
    u1 = _uniform01(seed, idx, 1)
    u2 = _uniform01(seed, idx, 2)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


# ---------------------------------------------------------------------------- workload fields
def gaussian_blobs(n, n_blobs=8, seed=0, z0=0, z1=None, centre_range=(0.2, 0.8), sigma_range=(0.05, 0.15)):
    """Sum of n_blobs Gaussians exp(-|x-c|^2 / (2 s^2)) on x_i = i/(n-1) in [0,1]^3
    (centres uniform in centre_range^3, s uniform in sigma_range, amplitude 1).  Rows z0..z1."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(centre_range[0], centre_range[1], size=(n_blobs, 3))
    sig = rng.uniform(sigma_range[0], sigma_range[1], size=n_blobs)
    z1 = n if z1 is None else z1
    t = np.arange(n, dtype=np.float64) / (n - 1)
    tz = t[z0:z1]
    f = np.zeros((z1 - z0, n, n), np.float64)
    for c, s in zip(centres, sig):
        ex = np.exp(-((t - c[0]) ** 2) / (2 * s * s))
        ey = np.exp(-((t - c[1]) ** 2) / (2 * s * s))
        ez = np.exp(-((tz - c[2]) ** 2) / (2 * s * s))
        f += ez[:, None, None] * ey[None, :, None] * ex[None, None, :]
    return f.astype(np.float32)


def random_lattice_upsampled(n_lattice=32, factor=4, seed=1):
    """An i.i.d. uniform[-1,1] lattice of n_lattice^3 trilinearly upsampled to (n_lattice*factor)^3
    samples placed at x = i*(n_lattice-1)/(n-1) lattice units."""
    rng = np.random.default_rng(seed)
    lat = rng.uniform(-1.0, 1.0, size=(n_lattice,) * 3).astype(np.float32).astype(np.float64)
    n = n_lattice * factor
    x = np.arange(n, dtype=np.float64) * (n_lattice - 1) / (n - 1)
    i0 = np.minimum(np.floor(x).astype(np.int64), n_lattice - 2)
    fr = x - i0
    out = np.zeros((n, n, n), np.float64)
    for dz in (0, 1):
        wz = fr if dz else 1.0 - fr
        for dy in (0, 1):
            wy = fr if dy else 1.0 - fr
            for dx in (0, 1):
                wx = fr if dx else 1.0 - fr
                out += (wz[:, None, None] * wy[None, :, None] * wx[None, None, :]) * lat[np.ix_(i0 + dz, i0 + dy, i0 + dx)]
    return out.astype(np.float32)


def gyroid(n, periods=8, noise=0.1, seed=2, z0=0, z1=None):
    """sin(x)cos(y) + sin(y)cos(z) + sin(z)cos(x) with x = 2*pi*periods*i/n, plus noise * N(0,1)
    per sample from a counter-based generator keyed by the linear index.  Rows z0..z1."""
    z1 = n if z1 is None else z1
    t = 2.0 * np.pi * periods * np.arange(n, dtype=np.float64) / n
    sx, cx = np.sin(t), np.cos(t)
    sz, cz = sx[z0:z1], cx[z0:z1]
    out = np.empty((z1 - z0, n, n), np.float32)
    for k in range(z1 - z0):
        z = z0 + k
        f = sx[None, :] * cx[:, None] + sx[:, None] * cz[k] + sz[k] * cx[None, :]
        if noise:
            idx = np.arange(z * n * n, (z + 1) * n * n, dtype=np.int64)
            f = f + noise * normal(seed, idx).reshape(n, n)
        out[k] = f.astype(np.float32)
    return out


def sphere_sdf(n, n_spheres=32, seed=4, centre_range=(0.15, 0.85), radius_range=(0.03, 0.12)):
    """min_i(|x - c_i| - r_i) on x_i = i/(n-1) in [0,1]^3."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(centre_range[0], centre_range[1], size=(n_spheres, 3))
    radii = rng.uniform(radius_range[0], radius_range[1], size=n_spheres)
    t = np.arange(n, dtype=np.float64) / (n - 1)
    out = np.full((n, n, n), np.inf)
    for c, r in zip(centres, radii):
        d = np.sqrt((t[:, None, None] - c[2]) ** 2 + (t[None, :, None] - c[1]) ** 2 + (t[None, None, :] - c[0]) ** 2) - r
        np.minimum(out, d, out=out)
    return out.astype(np.float32)


# workload definitions: (values, origin, delta, iso)
def workload(name):
    if name == "blobs64":
        n = 64
        return gaussian_blobs(n), (0.0, 0.0, 0.0), (1.0 / (n - 1),) * 3, 0.5
    if name == "random128":
        v = random_lattice_upsampled()
        n = v.shape[0]
        return v, (0.0, 0.0, 0.0), (1.0 / (n - 1),) * 3, 0.0
    if name.startswith("gyroid"):
        n = int(name[len("gyroid"):])
        return gyroid(n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0
    if name.startswith("spheres"):
        n = int(name[len("spheres"):])
        return sphere_sdf(n), (0.0, 0.0, 0.0), (1.0 / (n - 1),) * 3, 0.0
    raise KeyError(name)
