"""Synthetic inputs for the predicate and voxelization path.

Meshes are closed, consistently wound triangle soups (float64 vertices, int32
triangles). Grids "N^3 over [0,1]^3" are expressed through the unchanged
voxelize API the way the reference allows it: two unreferenced anchor points
(0,0,0) and (1,1,1) are appended to the vertex list and the grid is requested
with ``resolution=N, pad=0``; make_grid takes the bounding box over all points
(src/voxelize/voxelize.cpp:52-74) while triangles come only from facets, which
gives origin 1/(2N) and spacing 1/N.

Workloads:
  icosphere_c1      642 V / 1,280 T icosphere (radius 0.45), seeded rotation, 64^3
  tuples_orient3d   batched orient3d tuples, 90 % uniform, 10 % near-degenerate
  tuples_insphere   batched insphere tuples, same mix
  torus_c3          65,536-triangle jittered torus, 256^3
  harmonic_c4       icosphere L5 with spherical-harmonic displacement, 128^3
  batch_c5          64 uneven meshes with per-mesh resolutions
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

ANCHORS = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]])


@dataclass
class MeshCase:
    """A closed mesh plus the grid request that voxelizes it."""
    name: str
    points: np.ndarray       # (V, 3) float64, anchors included when anchored
    tris: np.ndarray         # (T, 3) int32
    resolution: int
    pad: int = 0
    meta: dict = field(default_factory=dict)


# ---- meshes -------------------------------------------------------------------------

def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]],
                 dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere(level: int):
    """Unit icosphere by recursive midpoint subdivision with projection:
    10*4^L + 2 vertices, 20*4^L triangles, outward winding."""
    v, f = _icosahedron()
    for _ in range(level):
        edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.sort(edges, axis=1)
        uniq, inv = np.unique(key, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid])
        nf = f.shape[0]
        m01 = base + inv[:nf]
        m12 = base + inv[nf:2 * nf]
        m20 = base + inv[2 * nf:]
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], 1),
            np.stack([f[:, 1], m12, m01], 1),
            np.stack([f[:, 2], m20, m12], 1),
            np.stack([m01, m12, m20], 1)])
    return v, f.astype(np.int32)


def geodesic_sphere(freq: int):
    """Unit geodesic sphere: every icosahedron face split into freq^2 triangles
    (20*freq^2 triangles), vertices projected to the sphere. Covers triangle counts
    between the icosphere levels."""
    v0, f0 = _icosahedron()
    n = max(1, int(freq))
    # lattice points (i, j) with i + j <= n on one face: weights (i, j, n-i-j) on (A, B, C)
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    keep = ii + jj <= n
    ii, jj = ii[keep], jj[keep]
    lid = -np.ones((n + 1, n + 1), dtype=np.int64)
    lid[ii, jj] = np.arange(ii.size)
    # triangles of one face in local lattice ids
    tri = []
    for i in range(n):
        for j in range(n - i):
            tri.append((lid[i, j], lid[i + 1, j], lid[i, j + 1]))
            if i + j + 1 < n:
                tri.append((lid[i + 1, j], lid[i + 1, j + 1], lid[i, j + 1]))
    tri = np.array(tri, dtype=np.int64)
    npf = ii.size
    # canonical integer key per point: (vertex, weight) pairs sorted by vertex id, with
    # zero weights dropped, so points shared by faces get one id
    verts = f0[:, None, :].repeat(npf, 1)                     # (20, npf, 3)
    w = np.stack([ii, jj, n - ii - jj], -1)[None].repeat(20, 0)  # (20, npf, 3)
    vid = np.where(w > 0, verts, 1 << 20)
    order = np.argsort(vid, axis=-1)
    vid = np.take_along_axis(vid, order, -1)
    w = np.take_along_axis(w, order, -1)
    key = np.concatenate([vid, w], -1).reshape(-1, 6)
    uniq, inv = np.unique(key, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    pos = np.zeros((uniq.shape[0], 3))
    for c in range(3):
        vv = uniq[:, c]
        ok = vv < (1 << 20)
        pos[ok] += v0[vv[ok]] * uniq[ok, 3 + c][:, None]
    pos /= np.linalg.norm(pos, axis=1, keepdims=True)
    faces = (inv.reshape(20, npf)[:, tri]).reshape(-1, 3).astype(np.int32)
    c = pos[faces].mean(axis=1)
    nrm = np.cross(pos[faces[:, 1]] - pos[faces[:, 0]], pos[faces[:, 2]] - pos[faces[:, 0]])
    flip = (nrm * c).sum(1) < 0
    faces[flip] = faces[flip][:, [0, 2, 1]]
    return pos, faces


def torus(R: float, r: float, nu: int, nv: int):
    """Torus around the z axis: nu x nv parametric grid, 2*nu*nv triangles, with the
    outward unit normal per vertex."""
    u = 2 * np.pi * np.arange(nu) / nu
    w = 2 * np.pi * np.arange(nv) / nv
    U, W = np.meshgrid(u, w, indexing="ij")
    x = (R + r * np.cos(W)) * np.cos(U)
    y = (R + r * np.cos(W)) * np.sin(U)
    z = r * np.sin(W)
    v = np.stack([x, y, z], -1).reshape(-1, 3)
    nrm = np.stack([np.cos(W) * np.cos(U), np.cos(W) * np.sin(U), np.sin(W)],
                   -1).reshape(-1, 3)
    i = np.arange(nu)[:, None]
    j = np.arange(nv)[None, :]
    a = (i * nv + j)
    b = (((i + 1) % nu) * nv + j)
    c = (((i + 1) % nu) * nv + (j + 1) % nv)
    d = (i * nv + (j + 1) % nv)
    f = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3),
                        np.stack([a, c, d], -1).reshape(-1, 3)]).astype(np.int32)
    return v, f, nrm


def real_sph_harm_sum(dirs: np.ndarray, lmax: int, rng: np.random.Generator):
    """A seeded random sum of real spherical harmonics up to degree lmax evaluated at
    unit directions, scaled so that its maximum magnitude over `dirs` is 1."""
    from scipy.special import lpmv
    x, y, z = dirs[:, 0], dirs[:, 1], dirs[:, 2]
    theta = np.arccos(np.clip(z, -1, 1))
    phi = np.arctan2(y, x)
    ct = np.cos(theta)
    s = np.zeros(dirs.shape[0])
    for l in range(1, lmax + 1):
        for m in range(-l, l + 1):
            am = abs(m)
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) *
                             math.factorial(l - am) / math.factorial(l + am))
            P = lpmv(am, l, ct)
            if m > 0:
                y_lm = math.sqrt(2) * norm * P * np.cos(m * phi)
            elif m < 0:
                y_lm = math.sqrt(2) * norm * P * np.sin(am * phi)
            else:
                y_lm = norm * P
            s += rng.standard_normal() * y_lm
    return s / np.max(np.abs(s))


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Haar-random rotation from a uniformly random unit quaternion."""
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def anchored(points: np.ndarray) -> np.ndarray:
    """Append the (0,0,0) and (1,1,1) anchors so that resolution=N, pad=0 yields the
    N^3 grid over [0,1]^3."""
    return np.concatenate([np.asarray(points, dtype=np.float64), ANCHORS])


def _drop_degenerate(v, f):
    nrm = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    keep = np.abs(nrm).sum(1) > 0
    return f[keep]


# ---- configurations -----------------------------------------------------------------

def icosphere_c1(seed: int = 1, snap: float | None = None) -> MeshCase:
    """Icosphere L3 (642 V, 1,280 T), radius 0.45, seeded Haar rotation, centred at
    (0.5,0.5,0.5); 64^3 grid over [0,1]^3. snap=2**-6 snaps every coordinate to that
    lattice (rays then hit edges exactly); snap=2**-7 also puts vertices on the column
    lines (vertex hits and the nudge path). Zero-area triangles after snapping are
    dropped."""
    rng = np.random.default_rng(seed)
    v, f = icosphere(3)
    v = v @ random_rotation(rng).T * 0.45 + 0.5
    name = "icosphere_l3_rot"
    if snap:
        v = np.round(v / snap) * snap
        f = _drop_degenerate(v, f)
        name += f"_snap{int(round(-math.log2(snap)))}"
    return MeshCase(name, anchored(v), f.astype(np.int32), 64, 0,
                    {"seed": seed, "snap": snap})


def torus_c3(seed: int = 3, res: int = 256, nu: int = 256, nv: int = 128) -> MeshCase:
    """Torus R=0.3, r=0.12 centred at (0.5,0.5,0.5), nu x nv parametric grid
    (32,768 V / 65,536 T at 256 x 128), each vertex moved along its normal by
    U(-1e-3, 1e-3); res^3 grid over [0,1]^3."""
    rng = np.random.default_rng(seed)
    v, f, nrm = torus(0.3, 0.12, nu, nv)
    v = v + nrm * rng.uniform(-1e-3, 1e-3, size=(v.shape[0], 1)) + 0.5
    return MeshCase(f"torus_{nu}x{nv}", anchored(v), f, res, 0,
                    {"seed": seed, "jitter": "U(-1e-3,1e-3) along normals"})


def harmonic_sphere(level_or_freq: int, amplitude: float, r0: float, rng,
                    geodesic=False):
    v, f = geodesic_sphere(level_or_freq) if geodesic else icosphere(level_or_freq)
    s = real_sph_harm_sum(v, 6, rng)
    return v * (r0 * (1.0 + amplitude * s))[:, None], f


def harmonic_c4(seed: int = 4, res: int = 128, level: int = 5) -> MeshCase:
    """Icosphere L5 (10,242 V / 20,480 T), radius 0.35*(1 + 0.05*S) with S a seeded
    random real spherical-harmonic sum up to degree 6 scaled to max |S| = 1, centred at
    (0.5,0.5,0.5); res^3 grid over [0,1]^3."""
    rng = np.random.default_rng(seed)
    v, f = harmonic_sphere(level, 0.05, 0.35, rng)
    return MeshCase(f"harmonic_l{level}", anchored(v + 0.5), f, res, 0, {"seed": seed})


def upstream_grad(n: int, seed: int = 40) -> np.ndarray:
    """Upstream dL/dSDF, i.i.d. N(0,1) float32."""
    return np.random.default_rng(seed).standard_normal(n).astype(np.float32)


def batch_c5(seed: int = 5, count: int = 64, tri_range=(500, 200_000),
             res_range=(32, 256)) -> list[MeshCase]:
    """`count` closed meshes drawn from {geodesic sphere, torus, harmonic sphere}, each
    with a Haar rotation and a scale U[0.3, 0.9] (the mesh's bounding radius becomes
    scale/2, centred in [0,1]^3); triangle counts and per-mesh grid resolutions are
    log-uniform in the given ranges."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(count):
        kind = int(rng.integers(3))
        target = math.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1])))
        res = int(round(math.exp(rng.uniform(math.log(res_range[0]),
                                             math.log(res_range[1])))))
        scale = rng.uniform(0.3, 0.9)
        rot = random_rotation(rng)
        if kind == 1:
            nv = max(3, int(round(math.sqrt(target / 4))))
            nu = 2 * nv
            v, f, nrm = torus(0.3, 0.12, nu, nv)
            name = f"torus_{nu}x{nv}"
        else:
            freq = max(1, int(round(math.sqrt(target / 20))))
            if kind == 0:
                v, f = geodesic_sphere(freq)
                name = f"sphere_f{freq}"
            else:
                v, f = harmonic_sphere(freq, 0.05, 1.0, rng, geodesic=True)
                name = f"harmonic_f{freq}"
        v = v @ rot.T
        v = v / np.max(np.linalg.norm(v, axis=1)) * (scale / 2) + 0.5
        out.append(MeshCase(f"c5_{m:02d}_{name}", anchored(v), f.astype(np.int32), res,
                            0, {"kind": kind, "scale": scale}))
    return out


# ---- predicate tuples ---------------------------------------------------------------

_NORM81 = None


def _norm_vectors(n2: int = 81) -> np.ndarray:
    global _NORM81
    if _NORM81 is None:
        r = int(math.isqrt(n2))
        g = np.arange(-r, r + 1)
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        m = X * X + Y * Y + Z * Z == n2
        _NORM81 = np.stack([X[m], Y[m], Z[m]], -1).astype(np.float64)
    return _NORM81


def _grid24(rng, shape, lo=-1.0, hi=1.0):
    """Uniform values on the 2^-24 grid in [lo, hi]."""
    return np.round(rng.uniform(lo, hi, size=shape) * 2.0 ** 24) / 2.0 ** 24


def _ulp_perturb(rng, soa, cols):
    """Move one random coordinate of each selected tuple by +-1 ulp (nextafter)."""
    n = cols.size
    if n == 0:
        return
    plane = rng.integers(soa.shape[0], size=n)
    up = rng.integers(2, size=n).astype(bool)
    vals = soa[plane, cols]
    soa[plane, cols] = np.where(up, np.nextafter(vals, np.inf), np.nextafter(vals, -np.inf))


def tuples_orient3d(n: int, seed: int = 2, degenerate: float = 0.10):
    """(12, n) float64 SoA orient3d tuples (plane = slot*3 + axis). 90 % uniform in
    [-1,1]^3. 10 % near-degenerate, half exactly coplanar (a, b, c on the 2^-24 grid and
    d = a + s(b-a) + t(c-a) with s, t in {k/16}, all exact in FP64) and half the same
    with one coordinate moved by +-1 ulp. Returns (soa, kind) with kind 0 = uniform,
    1 = coplanar, 2 = coplanar +-1 ulp."""
    rng = np.random.default_rng(seed)
    soa = rng.uniform(-1.0, 1.0, size=(12, n))
    kind = np.zeros(n, dtype=np.int8)
    nd = int(round(n * degenerate))
    idx = rng.choice(n, size=nd, replace=False) if nd else np.zeros(0, np.int64)
    if nd:
        a = _grid24(rng, (3, nd))
        b = _grid24(rng, (3, nd))
        c = _grid24(rng, (3, nd))
        s = rng.integers(-16, 17, size=nd) / 16.0
        t = rng.integers(-16, 17, size=nd) / 16.0
        d = a + s * (b - a) + t * (c - a)
        # the four points go to random slots so the coplanar point is not always d
        pts = np.stack([a, b, c, d])            # (4, 3, nd)
        perm = np.argsort(rng.random((4, nd)), axis=0)
        pts = np.take_along_axis(pts, perm[:, None, :], axis=0)
        soa[:, idx] = pts.reshape(12, nd)
        kind[idx] = 1
        half = idx[: nd // 2]
        _ulp_perturb(rng, soa, half)
        kind[half] = 2
    return np.ascontiguousarray(soa), kind


def tuples_insphere(n: int, seed: int = 22, degenerate: float = 0.10):
    """(15, n) float64 SoA insphere tuples. 90 % uniform in [-1,1]^3. 10 % near-
    degenerate, half exactly cospherical (centre on the 2^-24 grid plus distinct integer
    vectors of norm 9 times m*2^-24, all exact in FP64) and half the same with one
    coordinate moved by +-1 ulp. kind: 0 uniform, 1 cospherical, 2 cospherical +-1 ulp."""
    rng = np.random.default_rng(seed)
    soa = rng.uniform(-1.0, 1.0, size=(15, n))
    kind = np.zeros(n, dtype=np.int8)
    nd = int(round(n * degenerate))
    idx = rng.choice(n, size=nd, replace=False) if nd else np.zeros(0, np.int64)
    if nd:
        vecs = _norm_vectors(81)
        centre = _grid24(rng, (3, nd), -0.5, 0.5)
        m = rng.integers(1, 2 ** 24 // 20, size=nd).astype(np.float64) * 2.0 ** -24
        # five distinct vectors per tuple
        choice = np.argsort(rng.random((vecs.shape[0], nd)), axis=0)[:5]   # (5, nd)
        pts = centre[None, :, :] + vecs[choice].transpose(0, 2, 1) * m[None, None, :]
        soa[:, idx] = pts.reshape(15, nd)
        kind[idx] = 1
        half = idx[: nd // 2]
        _ulp_perturb(rng, soa, half)
        kind[half] = 2
    return np.ascontiguousarray(soa), kind


def tuples_wide_spread(n: int, npts: int, seed: int = 23):
    """(3*npts, n) float64 SoA tuples whose coordinates carry random 53-bit significands
    and independent random exponents in [-300, 300) (orient3d) or [-150, 150) (insphere),
    the ranges in which expansion arithmetic is exact: coordinate differences span far
    more than 126 bits, so the fixed-point stages decline and the expansion passes decide.
    Every tuple repeats a point (orient3d: b = a; insphere: e = a), so its determinant is
    exactly 0; in the second half one coordinate of the repeated copy is then moved by
    +-1 ulp (tiny nonzero determinants)."""
    lim = 300 if npts == 4 else 150
    rng = np.random.default_rng(seed)
    q = 3 * npts

    def spread(shape):
        m = rng.integers(2 ** 52, 2 ** 53, size=shape).astype(np.float64)
        sg = np.where(rng.random(shape) < 0.5, -1.0, 1.0)
        return sg * np.ldexp(m, rng.integers(-lim, lim, size=shape) - 53)

    t = spread((n, q))
    dst = 3 if npts == 4 else 12          # the repeated copy: b (orient3d) or e (insphere)
    t[:, dst:dst + 3] = t[:, 0:3]
    rows = np.arange(n // 2, n)
    col = dst + rng.integers(0, 3, size=rows.size)
    up = rng.random(rows.size) < 0.5
    t[rows, col] = np.nextafter(t[rows, col], np.where(up, np.inf, -np.inf))
    return np.ascontiguousarray(t.T)


def binary_stl_bytes(nu: int = 2048, nv: int = 1024, seed: int = 6) -> bytes:
    """A binary STL file (80-byte header, uint32 count, 50-byte records) of a closed torus
    (R = 0.3, r = 0.12, centred at 0.5) tessellated nu x nv (2 * nu * nv triangles, every
    vertex shared by six of them), float32 coordinates with seeded jitter: the input of
    the STL decode and vertex-merge workload."""
    rng = np.random.default_rng(seed)
    pts, tris, _ = torus(0.3, 0.12, nu, nv)
    pts = pts + 0.5 + rng.normal(scale=1e-4, size=pts.shape)
    soup = pts.astype("<f4")[tris]
    rec = np.zeros(len(tris), dtype=[("normal", "<f4", 3), ("v", "<f4", (3, 3)),
                                     ("attr", "<u2")])
    rec["v"] = soup
    head = b"torus workload".ljust(80, b"\0") + np.uint32(len(tris)).astype("<u4").tobytes()
    return head + rec.tobytes()


def lattice_tets(m: int, jitter: float = 0.0, seed: int = 7):
    """A tetrahedral mesh of the unit cube: an (m+1)^3 lattice of points, every cell split
    into the 6 tets of its Kuhn (path) subdivision (6 m^3 tets, all congruent when
    jitter = 0: dihedral angles of 45, 60 and 90 degrees), points moved by a seeded
    uniform jitter of up to `jitter` cells on interior points."""
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(*[np.arange(m + 1)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pts = g.astype(np.float64) / m
    if jitter:
        inner = ((g > 0) & (g < m)).all(1)
        pts[inner] += rng.uniform(-jitter, jitter, (int(inner.sum()), 3)) / m
    idx = lambda i, j, k: (i * (m + 1) + j) * (m + 1) + k  # noqa: E731
    c = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)
    tets = []
    for perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        cur = c.copy()
        verts = [idx(*cur.T)]
        for ax in perm:
            cur = cur.copy()
            cur[:, ax] += 1
            verts.append(idx(*cur.T))
        tets.append(np.stack(verts, 1))
    return pts, np.concatenate(tets).astype(np.int32)


def positive_tets(pts: np.ndarray, tets: np.ndarray) -> np.ndarray:
    """tets reordered (v2 <-> v3 where needed) so that det[v1-v0, v2-v0, v3-v0] > 0, the
    orientation the Delaunay insertion keeps internally (incremental.cpp:290-293)."""
    p = pts[tets]
    det = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0],
                                                             p[:, 3] - p[:, 0]))
    # orient3d(a,b,c,d) = det[a-d; b-d; c-d] has the opposite sign of det[b-a; c-a; d-a]
    out = tets.copy()
    flip = det > 0
    out[flip, 2], out[flip, 3] = tets[flip, 3], tets[flip, 2]
    return out


def tet_neighbors(tets: np.ndarray) -> np.ndarray:
    """nbrs[t, k] = the tet sharing the face opposite vertex k of tet t, or -1."""
    n = len(tets)
    faces = []
    for k in range(4):
        f = np.sort(np.delete(tets, k, axis=1), axis=1)
        faces.append(f)
    F = np.concatenate(faces)                      # (4n, 3)
    owner = np.tile(np.arange(n), 4)
    slot = np.repeat(np.arange(4), n)
    order = np.lexsort((F[:, 2], F[:, 1], F[:, 0]))
    Fs = F[order]
    same = (Fs[1:] == Fs[:-1]).all(1)
    nbrs = np.full((n, 4), -1, dtype=np.int32)
    a, b = order[:-1][same], order[1:][same]
    nbrs[owner[a], slot[a]] = owner[b]
    nbrs[owner[b], slot[b]] = owner[a]
    return nbrs



def delaunay_tets(n: int, seed: int = 8, lattice: int = 0):
    """A Delaunay tetrahedralization (scipy's Qhull) of n seeded points uniform in the
    unit cube, or, with lattice > 0, drawn from a lattice of `lattice` cells per axis
    (many exactly cospherical tuples). Returns (points, tets positively oriented, nbrs)."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    if lattice:
        g = rng.integers(0, lattice + 1, size=(4 * n, 3))
        pts = np.unique(g, axis=0)[:n].astype(np.float64) / lattice
        pts = pts[rng.permutation(len(pts))]
    else:
        pts = rng.uniform(0.0, 1.0, (n, 3))
    tets = Delaunay(pts).simplices.astype(np.int32)
    tets = positive_tets(pts, tets)
    return pts, tets, tet_neighbors(tets)
