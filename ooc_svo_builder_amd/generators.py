"""Seeded synthetic meshes for the partition -> voxelize -> SVO-build workloads.

The reference ships no sample meshes (its inputs come from tri_convert on user
models, tri_convert.cpp:85-150), so every workload here is generated. Records
follow tri_util.h:29-54: binary = v0, v1, v2 (9 float32); payload = v0, v1, v2,
one per-triangle normal, c0, c1, c2 (21 float32). The per-triangle normal is the
normalized mean of the three vertex normals, the rule of
tri_convert_util.h:43-52 (getShadingFaceNormal).

All meshes live in the cubic bbox [0,1]^3 that the .tri header states, so the
grid origin is 0 as partitioner.cpp:45 and voxelizer.cpp:189-196 assume.
"""
from __future__ import annotations

import numpy as np

# ----------------------------------------------------------------------------
# icosphere: 20 * 4**s triangles, 10 * 4**s + 2 shared vertices
# ----------------------------------------------------------------------------


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v, f


def _subdivide(v, f):
    """Split every face into 4; shared edge midpoints are created once."""
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    e.sort(axis=1)
    uniq, inv = np.unique(e, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    mid = v[uniq[:, 0]] + v[uniq[:, 1]]
    mid /= np.linalg.norm(mid, axis=1, keepdims=True)
    nv = len(v)
    v2 = np.concatenate([v, mid])
    nf = len(f)
    a = nv + inv[:nf]
    b = nv + inv[nf:2 * nf]
    c = nv + inv[2 * nf:]
    f2 = np.concatenate([np.stack([f[:, 0], a, c], 1), np.stack([f[:, 1], b, a], 1),
                         np.stack([f[:, 2], c, b], 1), np.stack([a, b, c], 1)])
    return v2, f2


def icosphere(subdivisions: int = 5, radius: float = 0.4, center=(0.5, 0.5, 0.5),
              noise: float = 0.02, seed: int = 1, payload: bool = False) -> np.ndarray:
    """Watertight icosphere; each shared vertex is displaced radially once by
    U(-noise, noise). Returns (N, 9) float32 binary records, or (N, 21) payload records
    (vertex normal = radial direction, per-vertex RGB U[0,1])."""
    v, f = _icosahedron()
    for _ in range(subdivisions):
        v, f = _subdivide(v, f)
    rng = np.random.default_rng(seed)
    r = radius + rng.uniform(-noise, noise, size=len(v))
    p = (v * r[:, None] + np.asarray(center, dtype=np.float64)).astype(np.float32)
    geo = p[f].reshape(len(f), 9)
    if not payload:
        return geo
    rec = np.empty((len(f), 21), dtype=np.float32)
    rec[:, 0:9] = geo
    m = v[f].sum(axis=1) / 3.0
    m /= np.linalg.norm(m, axis=1, keepdims=True)
    rec[:, 9:12] = m.astype(np.float32)
    C = rng.uniform(0.0, 1.0, size=(len(v), 3)).astype(np.float32)
    rec[:, 12:21] = C[f].reshape(len(f), 9)
    return rec


# ----------------------------------------------------------------------------
# heightfield terrain: (n-1)^2 * 2 triangles over [0,1]^2
# ----------------------------------------------------------------------------


def _hash_u32(ix, iy, seed):
    h = (ix.astype(np.uint64) * np.uint64(0x8DA6B343) ^ iy.astype(np.uint64) * np.uint64(0xD8163841)
         ^ np.uint64(seed & 0xFFFFFFFF) * np.uint64(0xCB1AB31F)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(15)
    h = (h * np.uint64(0x2C1B3C6D)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(12)
    h = (h * np.uint64(0x297A2D39)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(15)
    return h


def _value_noise(x, y, seed):
    ix = np.floor(x).astype(np.int64)
    iy = np.floor(y).astype(np.int64)
    fx = x - ix
    fy = y - iy
    sx = fx * fx * (3.0 - 2.0 * fx)
    sy = fy * fy * (3.0 - 2.0 * fy)
    ixu = ix & 0xFFFFFFFF
    iyu = iy & 0xFFFFFFFF

    def val(a, b):
        return _hash_u32(a, b, seed).astype(np.float64) / 4294967296.0

    v00 = val(ixu, iyu)
    v10 = val((ixu + 1) & 0xFFFFFFFF, iyu)
    v01 = val(ixu, (iyu + 1) & 0xFFFFFFFF)
    v11 = val((ixu + 1) & 0xFFFFFFFF, (iyu + 1) & 0xFFFFFFFF)
    a = v00 + (v10 - v00) * sx
    b = v01 + (v11 - v01) * sx
    return a + (b - a) * sy


def fbm(x, y, seed, octaves=6, lacunarity=2.0, gain=0.5, base_freq=4.0):
    """fBm of value noise normalized to [0, 1]."""
    total = np.zeros_like(x)
    amp, freq, norm = 1.0, base_freq, 0.0
    for o in range(octaves):
        total += amp * _value_noise(x * freq, y * freq, seed * 1000003 + o)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


def terrain(n: int = 1025, seed: int = 2, payload: bool = True) -> np.ndarray:
    """n x n vertex heightfield over [0,1]^2, z = 0.2 + 0.3 * fBm. Returns (N, 21)
    payload records (normals by central differences, per-vertex RGB U[0,1]) or (N, 9)
    binary records of the same geometry."""
    g = np.linspace(0.0, 1.0, n, dtype=np.float64)
    X, Y = np.meshgrid(g, g, indexing="ij")  # X[i, j] = g[i]
    Z = 0.2 + 0.3 * fbm(X, Y, seed)
    P = np.stack([X, Y, Z], axis=-1).astype(np.float32)
    i0 = np.arange(n - 1)
    I, J = np.meshgrid(i0, i0, indexing="ij")
    I = I.reshape(-1)
    J = J.reshape(-1)
    a = (I, J)
    b = (I + 1, J)
    c = (I + 1, J + 1)
    d = (I, J + 1)
    # two triangles per cell, interleaved: (a, b, c), (a, c, d)
    tri_idx = [(a, b, c), (a, c, d)]
    ntri = 2 * len(I)
    rec = np.empty((ntri, 21 if payload else 9), dtype=np.float32)
    for k, (p0, p1, p2) in enumerate(tri_idx):
        rec[k::2, 0:3] = P[p0]
        rec[k::2, 3:6] = P[p1]
        rec[k::2, 6:9] = P[p2]
    if not payload:
        return rec
    h = 1.0 / (n - 1)
    dzdx = np.gradient(Z, h, axis=0)
    dzdy = np.gradient(Z, h, axis=1)
    N = np.stack([-dzdx, -dzdy, np.ones_like(Z)], axis=-1)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    N = N.astype(np.float32)
    rng = np.random.default_rng(seed)
    C = rng.uniform(0.0, 1.0, size=(n, n, 3)).astype(np.float32)
    for k, (p0, p1, p2) in enumerate(tri_idx):
        m = (N[p0].astype(np.float64) + N[p1] + N[p2]) / 3.0
        m /= np.linalg.norm(m, axis=1, keepdims=True)
        rec[k::2, 9:12] = m.astype(np.float32)
        rec[k::2, 12:15] = C[p0]
        rec[k::2, 15:18] = C[p1]
        rec[k::2, 18:21] = C[p2]
    return rec


# ----------------------------------------------------------------------------
# sphere-surface triangle soup
# ----------------------------------------------------------------------------


def _soup_chunk(args):
    """Chunk k of a sphere soup: its own generator, seeded by (seed, k)."""
    seed, k, m, centers, radii, cdf, median_edge, sigma = args
    rng = np.random.default_rng([seed, k])
    n_spheres = len(radii)
    sid = np.minimum(np.searchsorted(cdf, rng.uniform(size=m)), n_spheres - 1)
    d = rng.normal(size=(m, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ctr = centers[sid] + d * radii[sid, None]
    # tangent basis
    helper = np.where(np.abs(d[:, :1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
    t1 = np.cross(d, helper)
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    t2 = np.cross(d, t1)
    edge = median_edge * np.exp(sigma * rng.normal(size=m))
    theta = rng.uniform(0.0, 2.0 * np.pi, size=m)
    rad = edge / np.sqrt(3.0)
    out = np.empty((m, 9), dtype=np.float32)
    for k3 in range(3):
        ang = theta + 2.0 * np.pi * k3 / 3.0
        v = ctr + rad[:, None] * (np.cos(ang)[:, None] * t1 + np.sin(ang)[:, None] * t2)
        out[:, 3 * k3:3 * k3 + 3] = v.astype(np.float32)
    return out


def sphere_soup(n_tris: int, n_spheres: int = 200, seed: int = 3,
                median_edge: float = 4.0 / 8192.0, sigma: float = 1.0,
                chunk: int = 1 << 20, workers: int = 1) -> np.ndarray:
    """Equilateral triangles centred on the surfaces of seeded spheres (centres
    U[0.1,0.9]^3, radii log-uniform in [0.01, 0.2]); sphere picked by area, edge
    length log-normal (median median_edge, sigma), random in-plane rotation. Returns
    (N, 9) float32. Chunks of 'chunk' triangles draw from their own seeded generators,
    so 'workers' processes give the same soup as one."""
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0.1, 0.9, size=(n_spheres, 3))
    radii = np.exp(rng.uniform(np.log(0.01), np.log(0.2), size=n_spheres))
    w = radii ** 2
    cdf = np.cumsum(w / w.sum())
    jobs = [(seed, k, min(chunk, n_tris - s), centers, radii, cdf, median_edge, sigma)
            for k, s in enumerate(range(0, n_tris, chunk))]
    out = np.empty((n_tris, 9), dtype=np.float32)
    if workers > 1 and len(jobs) > 1:
        from concurrent.futures import ProcessPoolExecutor
        with ProcessPoolExecutor(min(workers, len(jobs))) as ex:
            parts = ex.map(_soup_chunk, jobs)
            for k, part in enumerate(parts):
                out[k * chunk:k * chunk + len(part)] = part
    else:
        for k, j in enumerate(jobs):
            out[k * chunk:k * chunk + j[2]] = _soup_chunk(j)
    return out


# ----------------------------------------------------------------------------
# job mixes for the batched and served forms
# ----------------------------------------------------------------------------

MIX_GRIDS = (64, 128, 256, 512, 1024)


def splitmix64(seed: int, n: int) -> list:
    """n successive SplitMix64 outputs from 'seed' (the per-mesh seeds of a mix)."""
    out = []
    x = seed & 0xFFFFFFFFFFFFFFFF
    for _ in range(n):
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        out.append(z ^ (z >> 31))
    return out


def mix_spec(n_jobs: int = 512, seed: int = 4, n_min: int = 1000, n_max: int = 2_000_000):
    """Job shapes of a mix: per mesh (kind, size parameter, payload, gridsize, seed).
    50/50 icosphere/terrain, triangle count log-uniform in [n_min, n_max] (the icosphere
    subdivision or terrain resolution nearest to it), gridsize uniform over MIX_GRIDS,
    30% payload."""
    spec = []
    for s in splitmix64(seed, n_jobs):
        rng = np.random.default_rng(s)
        kind = "icosphere" if rng.uniform() < 0.5 else "terrain"
        n_tri = float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        if kind == "icosphere":  # 20 * 4**k triangles
            size = int(np.clip(np.round(np.log(n_tri / 20.0) / np.log(4.0)), 0, 8))
        else:  # 2 * (n - 1)**2 triangles
            size = int(max(2, np.round(np.sqrt(n_tri / 2.0)) + 1))
        payload = bool(rng.uniform() < 0.3)
        G = int(MIX_GRIDS[rng.integers(0, len(MIX_GRIDS))])
        spec.append((kind, size, payload, G, int(s & 0x7FFFFFFF)))
    return spec


def make_job_mesh(spec_item) -> np.ndarray:
    kind, size, payload, G, s = spec_item
    if kind == "icosphere":
        return icosphere(subdivisions=size, seed=s, payload=payload)
    return terrain(n=size, seed=s, payload=payload)


def job_mix(n_jobs: int = 512, seed: int = 4, workers: int = 1, **kw) -> list:
    """The mix as job dicts (records, payload, gridsize), meshes built with 'workers'
    processes."""
    spec = mix_spec(n_jobs, seed, **kw)
    if workers > 1:
        from concurrent.futures import ProcessPoolExecutor
        with ProcessPoolExecutor(workers) as ex:
            meshes = list(ex.map(make_job_mesh, spec, chunksize=4))
    else:
        meshes = [make_job_mesh(x) for x in spec]
    return [dict(records=m, payload=x[2], gridsize=x[3]) for m, x in zip(meshes, spec)]
