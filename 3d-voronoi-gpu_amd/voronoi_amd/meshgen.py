"""Seeded synthetic tet meshes and site sets.

Every random number comes from splitmix64 turned into a float with a 24-bit
mantissa, so the same seed gives the same bytes from Python and from C++.

Meshes:
  * jittered_cube(n)  -- unit cube split into n^3 cubes, each cut into the 6
    Kuhn tets along its main diagonal; interior vertex coordinates jittered by
    U(-0.1, 0.1) * h, boundary coordinates kept so the domain stays [0,1]^3.
  * torus_voxels(n)   -- voxels of an n^3 grid whose centre lies inside the
    torus (R=0.35, r=0.15) centred in the unit cube, 6 Kuhn tets each.

Every tet is positively oriented in the reference's convention: the first
vertex lies on the positive side of the plane of face (c, b, d)
(voronoi.cu:179, 119-123).

`normalize` restates the reference loader's mapping of the mesh into
[0,1000]^3 (test_voronoi.cu:14-38, 103-118) in FP32 with the same operation
order; `map_sites` applies the same affine map to sites, which the reference
expects to be already in that frame (test_voronoi.cu:135-147).
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Outputs offset..offset+n-1 of the splitmix64 stream started at `seed`."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform24(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n floats in [0,1) with 24-bit mantissa (exact in float32)."""
    z = splitmix64(seed, n, offset) >> np.uint64(40)
    return (z.astype(np.float64) * (1.0 / (1 << 24))).astype(np.float32)


# Kuhn subdivision: the 6 tets of a cube share the main diagonal 000-111; each
# follows one ordering of the axes.  Corner index = x + 2y + 4z.
def _kuhn_tets() -> np.ndarray:
    import itertools

    tets = []
    for perm in itertools.permutations(range(3)):
        c = [0, 0, 0]
        path = [0]
        for a in perm:
            c[a] = 1
            path.append(c[0] + 2 * c[1] + 4 * c[2])
        tets.append(path)
    return np.array(tets, dtype=np.int64)


_KUHN = _kuhn_tets()


def orientation(verts: np.ndarray, tets: np.ndarray) -> np.ndarray:
    """(b-c) x (d-c) . (a-c) per tet, in float64: > 0 is the reference's positive orientation."""
    v = verts.astype(np.float64)
    a, b, c, d = (v[tets[:, j]] for j in range(4))
    n = np.cross(b - c, d - c)
    return np.einsum("ij,ij->i", n, a - c)


def _orient(verts: np.ndarray, tets: np.ndarray) -> np.ndarray:
    o = orientation(verts, tets)
    flip = o < 0
    tets = tets.copy()
    tets[flip, 2], tets[flip, 3] = tets[flip, 3].copy(), tets[flip, 2].copy()
    return tets


def _cube_tets(cells: np.ndarray, n1: int) -> np.ndarray:
    """Tets of the cubes with lower corners `cells` (int array (m,3)) on a grid
    with n1 vertices per axis; vertex id = i + n1*(j + n1*k)."""
    corner = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)], dtype=np.int64)
    # corner index x + 2y + 4z matches the order above
    ids = cells[:, None, :] + corner[None, :, :]  # (m, 8, 3)
    vid = ids[..., 0] + n1 * (ids[..., 1] + n1 * ids[..., 2])  # (m, 8)
    return vid[:, _KUHN].reshape(-1, 4)


def jittered_cube(n: int, seed: int = 1, jitter: float = 0.1):
    """(verts float32 (V,3) in [0,1]^3, tets int32 (6n^3, 4))."""
    n1 = n + 1
    h = 1.0 / n
    k, j, i = np.meshgrid(np.arange(n1), np.arange(n1), np.arange(n1), indexing="ij")
    idx = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)  # vertex id = i + n1*(j + n1*k)
    u = uniform24(seed, idx.size).reshape(-1, 3).astype(np.float64)
    off = (2.0 * u - 1.0) * jitter * h
    interior = (idx > 0) & (idx < n)
    pos = idx * h + np.where(interior, off, 0.0)
    verts = pos.astype(np.float32)
    c = np.arange(n)
    kk, jj, ii = np.meshgrid(c, c, c, indexing="ij")
    cells = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1)
    tets = _orient(verts, _cube_tets(cells, n1))
    return verts, tets.astype(np.int32)


def torus_inside(p: np.ndarray, R: float = 0.35, r: float = 0.15) -> np.ndarray:
    x, y, z = p[:, 0] - 0.5, p[:, 1] - 0.5, p[:, 2] - 0.5
    return (np.sqrt(x * x + y * y) - R) ** 2 + z * z < r * r


def torus_voxels(n: int = 96):
    """(verts float32, tets int32, kept-voxel mask (n,n,n) indexed [k,j,i])."""
    h = 1.0 / n
    c = np.arange(n)
    kk, jj, ii = np.meshgrid(c, c, c, indexing="ij")
    cells = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1)
    keep = torus_inside((cells + 0.5) * h)
    cells = cells[keep]
    n1 = n + 1
    tets = _cube_tets(cells, n1)
    used, inv = np.unique(tets.ravel(), return_inverse=True)
    i = used % n1
    j = (used // n1) % n1
    k = used // (n1 * n1)
    verts = (np.stack([i, j, k], axis=1) * h).astype(np.float32)
    tets = _orient(verts, inv.reshape(-1, 4))
    return verts, tets.astype(np.int32), keep.reshape(n, n, n)


def uniform_sites(m: int, seed: int = 2) -> np.ndarray:
    return uniform24(seed, 3 * m).reshape(m, 3)


def gmm_torus_sites(m: int, n_vox: int = 96, comps: int = 16, seed: int = 3):
    """m sites from an isotropic 16-component Gaussian mixture (sigma in
    [0.02, 0.08]) with centres inside the torus, rejected to the voxelised torus."""
    rng = np.random.Generator(np.random.PCG64(seed))
    centres = []
    while len(centres) < comps:
        p = rng.random((64, 3))
        centres.extend(p[torus_inside(p)].tolist())
    centres = np.array(centres[:comps])
    sig = 0.02 + 0.06 * rng.random(comps)
    out = []
    got = 0
    _, _, keep = torus_voxels(n_vox) if n_vox else (None, None, None)
    while got < m:
        c = rng.integers(0, comps, size=4 * m)
        p = centres[c] + rng.standard_normal((4 * m, 3)) * sig[c, None]
        ok = np.all((p >= 0) & (p < 1), axis=1)
        p = p[ok]
        vi = np.floor(p * n_vox).astype(np.int64)
        p = p[keep[vi[:, 2], vi[:, 1], vi[:, 0]]]
        out.append(p)
        got += len(p)
    return np.concatenate(out)[:m].astype(np.float32)


def bbox_grown(verts: np.ndarray):
    """get_bbox (test_voronoi.cu:14-38) in FP32: bbox grown by 0.1% of its largest side."""
    v = verts.astype(np.float32)
    lo = v.min(axis=0)
    hi = v.max(axis=0)
    d = np.float32(max(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]))
    d = np.float32(np.float32(0.001) * d)
    return (lo - d).astype(np.float32), (hi + d).astype(np.float32)


def frame(verts: np.ndarray):
    """(lo, maxside) of the reference's [0,1000]^3 normalisation."""
    lo, hi = bbox_grown(verts)
    ext = (hi - lo).astype(np.float32)
    return lo, np.float32(max(ext[0], ext[1], ext[2]))


def to_frame(p: np.ndarray, lo: np.ndarray, maxside: np.float32) -> np.ndarray:
    """1000.f * (p - lo) / maxside in FP32 (test_voronoi.cu:112-114)."""
    p = p.astype(np.float32)
    return ((np.float32(1000.0) * (p - lo)) / maxside).astype(np.float32)


def normalize(verts: np.ndarray):
    lo, ms = frame(verts)
    return to_frame(verts, lo, ms)


def batch_sizes(n_inst: int = 256, lo: int = 12, hi: int = 45, seed: int = 11) -> np.ndarray:
    """Cube subdivisions n = round(exp(U[log lo, log hi])) of a batch of
    instances: 6 n^3 tets each (10,368 .. 546,750 for 12 .. 45)."""
    u = uniform24(seed, n_inst).astype(np.float64)
    return np.rint(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)


def make_batch(n_inst: int = 256, sites_per_tet: float = 0.1, seed: int = 11):
    """A batch of independent jittered-cube instances with sizes from
    batch_sizes and round(sites_per_tet * tets) uniform sites each."""
    out = []
    for i, n in enumerate(batch_sizes(n_inst, seed=seed)):
        n = int(n)
        out.append(make_problem("cube", n, max(2, int(round(sites_per_tet * 6 * n ** 3))), seed=seed * 1000 + i))
    return out


def serving_jobs(n_jobs: int, rate: float, seed: int = 21, torus_frac: float = 0.05):
    """A job stream: Poisson arrivals at `rate` jobs/s; each job a jittered
    cube of 6n^3 tets, n = round(logU[12,45]) with tets/10 sites, or (with
    probability torus_frac) a voxelised torus of 32..64 voxels per side with
    tets/10 Gaussian-mixture sites; 5..50 Lloyd iterations.  Returns a list of
    (arrival_s, (verts, tets, sites), n_iter, kind) with kind "cube" or "torus"."""
    u = uniform24(seed, 4 * n_jobs).astype(np.float64).reshape(n_jobs, 4)
    sizes = batch_sizes(n_jobs, seed=seed + 1)
    t = np.cumsum(-np.log1p(-u[:, 0]) / rate)
    out = []
    for i in range(n_jobs):
        if u[i, 1] < torus_frac:
            nv = 32 + int(u[i, 2] * 33)
            verts, tets, _ = torus_voxels(nv)
            prob = make_problem("torus", nv, max(100, len(tets) // 10), seed=seed * 7919 + i)
            kind = "torus"
        else:
            n = int(sizes[i])
            prob = make_problem("cube", n, max(100, (6 * n ** 3) // 10), seed=seed * 7919 + i)
            kind = "cube"
        out.append((float(t[i]), prob, 5 + int(u[i, 3] * 46), kind))
    return out


def make_problem(kind: str, n: int, n_sites: int, seed: int = 7):
    """Normalised (verts, tets, sites) for a named construction."""
    if kind == "cube":
        verts, tets = jittered_cube(n, seed=seed)
        sites = uniform_sites(n_sites, seed=seed + 1)
    elif kind == "torus":
        verts, tets, _ = torus_voxels(n)
        sites = gmm_torus_sites(n_sites, n_vox=n, seed=seed + 1)
    else:
        raise ValueError(kind)
    lo, ms = frame(verts)
    return to_frame(verts, lo, ms), tets, to_frame(sites, lo, ms)
