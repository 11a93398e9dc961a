"""CPU oracle for the predicate and voxelization path (ctypes over liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against. Only tests/,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this module;
the shipped package never does.

Every function restates the CyberMeshGenerator behaviour cited in oracle.h.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

EXACT = 0
COMPAT = 1


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def _load() -> ctypes.CDLL:
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    P = ctypes.c_void_p
    sz = ctypes.c_size_t
    i32 = ctypes.c_int
    d = ctypes.c_double
    lib.orc_orient3d_naive.restype = d
    lib.orc_orient3d_naive.argtypes = [P] * 4
    lib.orc_insphere_naive.restype = d
    lib.orc_insphere_naive.argtypes = [P] * 5
    lib.orc_orient3d_exact.restype = i32
    lib.orc_orient3d_exact.argtypes = [P] * 4
    lib.orc_insphere_exact.restype = i32
    lib.orc_insphere_exact.argtypes = [P] * 5
    lib.orc_orient3d_sign.restype = i32
    lib.orc_orient3d_sign.argtypes = [P] * 4 + [i32]
    lib.orc_insphere_sign.restype = i32
    lib.orc_insphere_sign.argtypes = [P] * 5 + [i32]
    lib.orc_orient3d_batch.argtypes = [P, sz, sz, i32, P, i32]
    lib.orc_insphere_batch.argtypes = [P, sz, sz, i32, P, i32]
    lib.orc_orient3d_indexed.argtypes = [P, P, sz, i32, P, i32]
    lib.orc_fan_triangulate.restype = sz
    lib.orc_fan_triangulate.argtypes = [P, sz, P, P]
    lib.orc_make_grid.restype = i32
    lib.orc_make_grid.argtypes = [P, sz, i32, i32, P]
    lib.orc_occupancy.argtypes = [P, P, sz, P, P, P, i32]
    lib.orc_point_tri_dist2.restype = d
    lib.orc_point_tri_dist2.argtypes = [P] * 5
    lib.orc_sdf_shell.argtypes = [P, P, sz, P, P, P, sz, sz, sz]
    lib.orc_sdf_brute.argtypes = [P, P, sz, P, P, P, i32]
    lib.orc_sdf_bvh.argtypes = [P, P, sz, P, P, P, P, P, sz, sz, i32]
    lib.orc_sdf_grad.argtypes = [P, sz, P, sz, P, P, P, P, P, i32]
    lib.orc_sdf_dist2_to.argtypes = [P, P, P, P, P, i32]
    lib.orc_sdf_fd.restype = d
    lib.orc_sdf_fd.argtypes = [P, sz, P, sz, P, P, P, P, P, i32, i32, d, P, i32]
    lib.orc_x86_trunc.restype = i32
    lib.orc_x86_trunc.argtypes = [d]
    lib.orc_ray_hits.restype = i32
    lib.orc_ray_hits.argtypes = [P] * 5
    lib.orc_carve_resolution.restype = i32
    lib.orc_carve_resolution.argtypes = [sz]
    lib.orc_carve_create.restype = P
    lib.orc_carve_create.argtypes = [P, sz, P, sz]
    lib.orc_carve_destroy.argtypes = [P]
    lib.orc_carve_cell_entries.restype = i32
    lib.orc_carve_cell_entries.argtypes = [P, i32]
    lib.orc_carve_inside.restype = i32
    lib.orc_carve_inside.argtypes = [P, P]
    lib.orc_carve_inside_brute.restype = i32
    lib.orc_carve_inside_brute.argtypes = [P, P, sz, P]
    lib.orc_tet_centroid.argtypes = [P, P, P]
    lib.orc_carve_points.argtypes = [P, P, sz, P, i32]
    lib.orc_carve_tets.argtypes = [P, P, sz, P, i32]
    lib.orc_carve_parity_points.argtypes = [P, P, sz, P, P, sz, P, i32]
    i64 = ctypes.c_int64
    PP = ctypes.POINTER(P)
    lib.orc_weld.restype = i64
    lib.orc_weld.argtypes = [P, i64, P, P]
    lib.orc_read_stl.restype = i64
    lib.orc_read_stl.argtypes = [ctypes.c_char_p, PP, PP, ctypes.POINTER(i64),
                                 ctypes.POINTER(i32), ctypes.c_char_p, sz]
    lib.orc_stl_soup.restype = i64
    lib.orc_stl_soup.argtypes = [ctypes.c_char_p, PP, ctypes.POINTER(i32), ctypes.c_char_p, sz]
    lib.orc_read_obj.restype = i32
    lib.orc_read_obj.argtypes = [ctypes.c_char_p, PP, ctypes.POINTER(i64), PP, PP,
                                 ctypes.POINTER(i64), ctypes.c_char_p, sz]
    lib.orc_free.argtypes = [P]
    lib.orc_tet_volume.restype = d
    lib.orc_tet_volume.argtypes = [P] * 4
    lib.orc_tet_dihedrals.argtypes = [P] * 5
    lib.orc_tet_quality.argtypes = [P, P, P, P, P, P]
    lib.orc_quality_report.argtypes = [P, P, sz, i32, P, P]
    lib.orc_dihedrals_batch.argtypes = [P, P, sz, P]
    lib.orc_locate.restype = i32
    lib.orc_locate.argtypes = [P, P, P, sz, P, i32, P]
    lib.orc_locate_batch.argtypes = [P, P, P, sz, P, sz, P, i32, P, P, i32]
    lib.orc_cavity_batch.argtypes = [P, P, P, P, sz, P, i32, P, P, i32]
    return lib


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


class _Grid(ctypes.Structure):
    _fields_ = [("corner", ctypes.c_double * 3), ("spacing", ctypes.c_double),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int)]


@dataclass
class Grid:
    corner: tuple
    spacing: float
    dims: tuple  # (nx, ny, nz)

    @property
    def origin(self):
        """Centre of cell (0,0,0) (voxelize.cpp:389-391)."""
        return tuple(c + 0.5 * self.spacing for c in self.corner)

    @property
    def ncell(self) -> int:
        nx, ny, nz = self.dims
        return nx * ny * nz

    def _c(self) -> _Grid:
        g = _Grid()
        for a in range(3):
            g.corner[a] = self.corner[a]
        g.spacing = self.spacing
        g.nx, g.ny, g.nz = self.dims
        return g


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---- predicates ---------------------------------------------------------------------

def orient3d_naive(a, b, c, d) -> float:
    a, b, c, d = (_f64(x) for x in (a, b, c, d))
    return lib().orc_orient3d_naive(_p(a), _p(b), _p(c), _p(d))


def insphere_naive(a, b, c, d, e) -> float:
    a, b, c, d, e = (_f64(x) for x in (a, b, c, d, e))
    return lib().orc_insphere_naive(_p(a), _p(b), _p(c), _p(d), _p(e))


def orient3d_sign(a, b, c, d, mode=EXACT) -> int:
    a, b, c, d = (_f64(x) for x in (a, b, c, d))
    return lib().orc_orient3d_sign(_p(a), _p(b), _p(c), _p(d), mode)


def insphere_sign(a, b, c, d, e, mode=EXACT) -> int:
    a, b, c, d, e = (_f64(x) for x in (a, b, c, d, e))
    return lib().orc_insphere_sign(_p(a), _p(b), _p(c), _p(d), _p(e), mode)


def orient3d_batch(soa: np.ndarray, mode=EXACT, nthreads=0) -> np.ndarray:
    """soa: (12, n) float64 planes (slot*3 + axis)."""
    soa = _f64(soa)
    assert soa.shape[0] == 12
    n = soa.shape[1]
    out = np.empty(n, dtype=np.int8)
    lib().orc_orient3d_batch(_p(soa), n, n, mode, _p(out), nthreads)
    return out


def insphere_batch(soa: np.ndarray, mode=EXACT, nthreads=0) -> np.ndarray:
    """soa: (15, n) float64 planes (slot*3 + axis)."""
    soa = _f64(soa)
    assert soa.shape[0] == 15
    n = soa.shape[1]
    out = np.empty(n, dtype=np.int8)
    lib().orc_insphere_batch(_p(soa), n, n, mode, _p(out), nthreads)
    return out


def orient3d_indexed(xyz, quads, mode=COMPAT, nthreads=0) -> np.ndarray:
    xyz = _f64(xyz).reshape(-1, 3)
    quads = _i32(quads).reshape(-1, 4)
    out = np.empty(quads.shape[0], dtype=np.int8)
    lib().orc_orient3d_indexed(_p(xyz), _p(quads), quads.shape[0], mode, _p(out), nthreads)
    return out


# ---- voxelization -------------------------------------------------------------------

def fan_triangulate(polygons) -> np.ndarray:
    off = np.zeros(len(polygons) + 1, dtype=np.int32)
    for i, p in enumerate(polygons):
        off[i + 1] = off[i] + len(p)
    idx = _i32(np.concatenate([np.asarray(p, dtype=np.int32) for p in polygons])
               if polygons else np.zeros(0, np.int32))
    cap = int(sum(max(len(p) - 2, 0) for p in polygons))
    tris = np.zeros((max(cap, 1), 3), dtype=np.int32)
    m = lib().orc_fan_triangulate(_p(off), len(polygons), _p(idx), _p(tris))
    return tris[:m].copy()


def make_grid(pts, resolution: int, pad: int) -> Grid:
    pts = _f64(pts).reshape(-1, 3)
    g = _Grid()
    rc = lib().orc_make_grid(_p(pts), pts.shape[0], resolution, pad, ctypes.byref(g))
    if rc != 0:
        raise ValueError("zero-extent point set")
    return Grid(tuple(g.corner), g.spacing, (g.nx, g.ny, g.nz))


def occupancy(pts, tris, grid: Grid, nthreads=0, with_stats=False):
    """uint8 occupancy, flat idx = (k*ny + j)*nx + i."""
    pts = _f64(pts).reshape(-1, 3)
    tris = _i32(tris).reshape(-1, 3)
    occ = np.zeros(grid.ncell, dtype=np.uint8)
    stats = np.zeros(2, dtype=np.int64)
    g = grid._c()
    lib().orc_occupancy(_p(pts), _p(tris), tris.shape[0], ctypes.byref(g), _p(occ),
                        _p(stats), nthreads)
    return (occ, stats) if with_stats else occ


def point_tri_dist2(p, a, b, c):
    p, a, b, c = (_f64(x) for x in (p, a, b, c))
    bary = np.zeros(3)
    d2 = lib().orc_point_tri_dist2(_p(p), _p(a), _p(b), _p(c), _p(bary))
    return d2, bary


def sdf_shell(pts, tris, grid: Grid, occ, first=0, count=0, stride=1) -> np.ndarray:
    pts = _f64(pts).reshape(-1, 3)
    tris = _i32(tris).reshape(-1, 3)
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    out = np.zeros(grid.ncell, dtype=np.float32)
    g = grid._c()
    lib().orc_sdf_shell(_p(pts), _p(tris), tris.shape[0], ctypes.byref(g), _p(occ),
                        _p(out), first, count, stride)
    return out


def sdf_brute(pts, tris, grid: Grid, occ, nthreads=0) -> np.ndarray:
    pts = _f64(pts).reshape(-1, 3)
    tris = _i32(tris).reshape(-1, 3)
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    out = np.zeros(grid.ncell, dtype=np.float32)
    g = grid._c()
    lib().orc_sdf_brute(_p(pts), _p(tris), tris.shape[0], ctypes.byref(g), _p(occ),
                        _p(out), nthreads)
    return out


def sdf_bvh(pts, tris, grid: Grid, occ, first=0, count=0, nthreads=0):
    """Returns (dist float32, best FP64 d^2, argmin triangle) over cells
    [first, first+count) (count=0: all); entries outside the range are zero."""
    pts = _f64(pts).reshape(-1, 3)
    tris = _i32(tris).reshape(-1, 3)
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    dist = np.zeros(grid.ncell, dtype=np.float32)
    best = np.zeros(grid.ncell, dtype=np.float64)
    arg = np.full(grid.ncell, -1, dtype=np.int32)
    g = grid._c()
    lib().orc_sdf_bvh(_p(pts), _p(tris), tris.shape[0], ctypes.byref(g), _p(occ),
                      _p(dist), _p(best), _p(arg), first, count, nthreads)
    return dist, best, arg


def sdf_grad(pts, tris, grid: Grid, occ, gout, nthreads=0, arg=None) -> np.ndarray:
    """FP64 adjoint dL/dpts; arg (per cell, optional) names the triangle whose closest
    point each cell uses, instead of the FP64 nearest one."""
    pts = _f64(pts).reshape(-1, 3)
    tris = _i32(tris).reshape(-1, 3)
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    gout = np.ascontiguousarray(gout, dtype=np.float32)
    grad = np.zeros_like(pts)
    g = grid._c()
    pa = None
    if arg is not None:
        arg = np.ascontiguousarray(arg, dtype=np.int32)
        if arg.size != grid.ncell:
            raise ValueError("arg must have one entry per cell")
        pa = _p(arg)
    lib().orc_sdf_grad(_p(pts), pts.shape[0], _p(tris), tris.shape[0], ctypes.byref(g),
                       _p(occ), _p(gout), pa, _p(grad), nthreads)
    return grad


def sdf_dist2_to(pts, tris, grid: Grid, arg, nthreads=0) -> np.ndarray:
    """FP64 squared distance from each cell centre to triangle arg[cell] (-1 if < 0)."""
    pts = _f64(pts).reshape(-1, 3)
    tris = _i32(tris).reshape(-1, 3)
    arg = np.ascontiguousarray(arg, dtype=np.int32)
    if arg.size != grid.ncell:
        raise ValueError("arg must have one entry per cell")
    out = np.zeros(grid.ncell, dtype=np.float64)
    g = grid._c()
    lib().orc_sdf_dist2_to(_p(pts), _p(tris), ctypes.byref(g), _p(arg), _p(out), nthreads)
    return out


def sdf_fd(pts, tris, grid: Grid, occ, gout, vert: int, axis: int, h=1e-5,
           nthreads=0, base=None, with_ties=False):
    """Central difference of L = sum gout * sdf w.r.t. pts[vert, axis]; base =
    (best, arg) from sdf_bvh lets many calls share the unperturbed search. with_ties:
    returns (fd, tie_bound), tie_bound = sum over the cells whose nearest triangle
    differs between the samples -h, 0, +h of |g| times the spread of those triangles'
    derivatives."""
    pts = _f64(pts).reshape(-1, 3)
    tris = _i32(tris).reshape(-1, 3)
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    gout = np.ascontiguousarray(gout, dtype=np.float32)
    g = grid._c()
    bb = ba = None
    if base is not None:
        best = np.ascontiguousarray(base[0], dtype=np.float64)
        arg = np.ascontiguousarray(base[1], dtype=np.int32)
        bb, ba = _p(best), _p(arg)
    ties = ctypes.c_double(0.0)
    fd = lib().orc_sdf_fd(_p(pts), pts.shape[0], _p(tris), tris.shape[0],
                          ctypes.byref(g), _p(occ), _p(gout), bb, ba, vert, axis, h,
                          ctypes.byref(ties), nthreads)
    return (fd, ties.value) if with_ties else fd


def voxelize(pts, polygons_or_tris, resolution=64, pad=1, mode="occupancy", nthreads=0):
    """cmg::voxelize::voxelize restated (voxelize.cpp:362-398): returns
    (grid, occupancy-or-distance flat array). polygons_or_tris: (M,3) int array of
    triangles or a list of polygons (fan-triangulated)."""
    pts = _f64(pts).reshape(-1, 3)
    if resolution < 1:
        raise ValueError("voxelize: resolution must be >= 1")
    if pts.shape[0] == 0:
        raise ValueError("voxelize: PLC has no points")
    if isinstance(polygons_or_tris, np.ndarray) and polygons_or_tris.ndim == 2 \
            and polygons_or_tris.shape[1] == 3:
        tris = _i32(polygons_or_tris)
    else:
        tris = fan_triangulate(list(polygons_or_tris))
    if tris.shape[0] == 0:
        raise ValueError("voxelize: PLC has no boundary triangles")
    grid = make_grid(pts, resolution, pad)
    occ = occupancy(pts, tris, grid, nthreads)
    if mode in ("sdf", 1):
        return grid, sdf_bvh(pts, tris, grid, occ, nthreads=nthreads)[0]
    return grid, occ


# ---- PLC carve ----------------------------------------------------------------------

def x86_trunc(v: float) -> int:
    return lib().orc_x86_trunc(float(v))


def ray_hits(o, d, a, b, c) -> bool:
    o, d, a, b, c = (_f64(x) for x in (o, d, a, b, c))
    return bool(lib().orc_ray_hits(_p(o), _p(d), _p(a), _p(b), _p(c)))


def carve_resolution(ntris: int) -> int:
    return lib().orc_carve_resolution(ntris)


class Carve:
    """CarveGrid restated (tetrahedralize_plc.cpp:60-168) over pts and tris."""

    def __init__(self, pts, tris):
        self.pts = _f64(pts).reshape(-1, 3)
        self.tris = _i32(tris).reshape(-1, 3)
        self._h = lib().orc_carve_create(_p(self.pts), self.pts.shape[0], _p(self.tris),
                                         self.tris.shape[0])

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            lib().orc_carve_destroy(h)

    def entries(self, k: int) -> int:
        return lib().orc_carve_cell_entries(self._h, k)

    def inside(self, q) -> bool:
        q = _f64(q)
        return bool(lib().orc_carve_inside(self._h, _p(q)))

    def points(self, q, nthreads=0) -> np.ndarray:
        q = _f64(q).reshape(-1, 3)
        out = np.zeros(q.shape[0], dtype=np.uint8)
        lib().orc_carve_points(self._h, _p(q), q.shape[0], _p(out), nthreads)
        return out

    def tets(self, tets, nthreads=0) -> np.ndarray:
        tets = _i32(tets).reshape(-1, 4)
        out = np.zeros(tets.shape[0], dtype=np.uint8)
        lib().orc_carve_tets(self._h, _p(tets), tets.shape[0], _p(out), nthreads)
        return out


def carve_inside_brute(pts, tris, q) -> bool:
    pts = _f64(pts).reshape(-1, 3)
    tris = _i32(tris).reshape(-1, 3)
    q = _f64(q)
    return bool(lib().orc_carve_inside_brute(_p(pts), _p(tris), tris.shape[0], _p(q)))


def tet_centroid(pts, tet) -> np.ndarray:
    pts = _f64(pts).reshape(-1, 3)
    tet = _i32(tet)
    c = np.zeros(3)
    lib().orc_tet_centroid(_p(pts), _p(tet), _p(c))
    return c


def carve_parity_points(pts, tris, grid: Grid, q, nthreads=0) -> np.ndarray:
    pts = _f64(pts).reshape(-1, 3)
    tris = _i32(tris).reshape(-1, 3)
    q = _f64(q).reshape(-1, 3)
    out = np.zeros(q.shape[0], dtype=np.uint8)
    g = grid._c()
    lib().orc_carve_parity_points(_p(pts), _p(tris), tris.shape[0], ctypes.byref(g),
                                  _p(q), q.shape[0], _p(out), nthreads)
    return out


# ---- surface file readers -------------------------------------------------------------

def weld(soup):
    """(points, idx): first-occurrence merge of a vertex soup (N, 3) (stl.cpp:24-61)."""
    soup = _f64(soup).reshape(-1, 3)
    n = soup.shape[0]
    pts = np.zeros((max(n, 1), 3))
    idx = np.zeros(max(n, 1), dtype=np.int32)
    m = lib().orc_weld(_p(soup), n, _p(pts), _p(idx))
    return pts[:m].copy(), idx[:n].copy()


def _take(ptr, count, dtype):
    if not ptr or count == 0:
        if ptr:
            lib().orc_free(ptr)
        return np.zeros(0, dtype=dtype)
    ct = ctypes.c_double if dtype == np.float64 else ctypes.c_int32
    arr = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)), shape=(count,)).copy()
    lib().orc_free(ptr)
    return arr


class ReadError(ValueError):
    pass


def read_stl(path: str):
    """(points (V,3), tris (T,3), binary) as src/io/stl.cpp reads them; ReadError with
    the reference's message on failure."""
    P_, T_ = ctypes.c_void_p(), ctypes.c_void_p()
    nt = ctypes.c_int64()
    binary = ctypes.c_int()
    err = ctypes.create_string_buffer(512)
    n = lib().orc_read_stl(path.encode(), ctypes.byref(P_), ctypes.byref(T_),
                           ctypes.byref(nt), ctypes.byref(binary), err, 512)
    if n < 0:
        raise ReadError(err.value.decode())
    pts = _take(P_.value, 3 * n, np.float64).reshape(-1, 3)
    tris = _take(T_.value, 3 * nt.value, np.int32).reshape(-1, 3)
    return pts, tris, bool(binary.value)


def stl_soup(path: str):
    """The vertex soup (3 per triangle, (N,3) float64) before merging, and binary."""
    S = ctypes.c_void_p()
    binary = ctypes.c_int()
    err = ctypes.create_string_buffer(512)
    n = lib().orc_stl_soup(path.encode(), ctypes.byref(S), ctypes.byref(binary), err, 512)
    if n < 0:
        raise ReadError(err.value.decode())
    return _take(S.value, 3 * n, np.float64).reshape(-1, 3), bool(binary.value)


def read_obj(path: str):
    """(points (V,3), polygons: list of int32 arrays) as src/io/obj.cpp reads them."""
    P_, O_, I_ = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    npts, nf = ctypes.c_int64(), ctypes.c_int64()
    err = ctypes.create_string_buffer(512)
    rc = lib().orc_read_obj(path.encode(), ctypes.byref(P_), ctypes.byref(npts),
                            ctypes.byref(O_), ctypes.byref(I_), ctypes.byref(nf), err, 512)
    if rc != 0:
        raise ReadError(err.value.decode())
    pts = _take(P_.value, 3 * npts.value, np.float64).reshape(-1, 3)
    off = _take(O_.value, nf.value + 1, np.int32)
    idx = _take(I_.value, int(off[-1]) if off.size else 0, np.int32)
    return pts, [idx[off[f]:off[f + 1]] for f in range(nf.value)]


# ---- mesh quality report ----------------------------------------------------------------

class _Quality(ctypes.Structure):
    _fields_ = [("num_tets", ctypes.c_int64),
                ("min_radius_edge", ctypes.c_double), ("max_radius_edge", ctypes.c_double),
                ("mean_radius_edge", ctypes.c_double),
                ("min_dihedral", ctypes.c_double), ("max_dihedral", ctypes.c_double),
                ("mean_dihedral", ctypes.c_double),
                ("min_volume", ctypes.c_double), ("max_volume", ctypes.c_double),
                ("total_volume", ctypes.c_double),
                ("dihedral_histogram", ctypes.c_int32 * 18), ("num_worst", ctypes.c_int64)]


def tet_dihedrals(a, b, c, d) -> np.ndarray:
    a, b, c, d = (_f64(x) for x in (a, b, c, d))
    out = np.zeros(6)
    lib().orc_tet_dihedrals(_p(a), _p(b), _p(c), _p(d), _p(out))
    return out


def tet_quality(pts, tet):
    """(ratio, valid, volume, dihedrals[6]) of one tet."""
    pts = _f64(pts).reshape(-1, 3)
    tet = _i32(tet)
    r, v, vol = ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
    dih = np.zeros(6)
    lib().orc_tet_quality(_p(pts), _p(tet), ctypes.byref(r), ctypes.byref(v),
                          ctypes.byref(vol), _p(dih))
    return r.value, bool(v.value), vol.value, dih


def quality_report(pts, tets, worst_count=10) -> dict:
    """quality::report (metrics.cpp:81-169) as a dict; 'worst_tets' descending ratio,
    ties by ascending index."""
    pts = _f64(pts).reshape(-1, 3)
    tets = _i32(tets).reshape(-1, 4)
    q = _Quality()
    n = tets.shape[0]
    worst = np.zeros(max(n, 1), dtype=np.int32)
    lib().orc_quality_report(_p(pts), _p(tets), n, worst_count, ctypes.byref(q), _p(worst))
    out = {f: getattr(q, f) for f, _ in _Quality._fields_ if f != "dihedral_histogram"}
    out["dihedral_histogram"] = list(q.dihedral_histogram)
    out["worst_tets"] = worst[: q.num_worst].tolist()
    return out


def dihedrals(pts, tets) -> np.ndarray:
    """(T, 6) dihedral angles in degrees (metrics.cpp:52-77)."""
    pts = _f64(pts).reshape(-1, 3)
    tets = _i32(tets).reshape(-1, 4)
    out = np.zeros((tets.shape[0], 6))
    lib().orc_dihedrals_batch(_p(pts), _p(tets), tets.shape[0], _p(out))
    return out


# ---- point location ------------------------------------------------------------------

def locate(pts, tets, nbrs, q, start=0, nthreads=0):
    """(tet, steps) per query: incremental.cpp:321-358's walk from start (an int or one
    start tet per query); steps -1 where brute force answered."""
    pts = _f64(pts).reshape(-1, 3)
    tets = _i32(tets).reshape(-1, 4)
    nbrs = _i32(nbrs).reshape(-1, 4)
    q = _f64(q).reshape(-1, 3)
    out = np.zeros(q.shape[0], dtype=np.int32)
    steps = np.zeros(q.shape[0], dtype=np.int32)
    if np.isscalar(start):
        lib().orc_locate_batch(_p(pts), _p(tets), _p(nbrs), tets.shape[0], _p(q), q.shape[0],
                               None, int(start), _p(out), _p(steps), nthreads)
    else:
        st = _i32(start)
        lib().orc_locate_batch(_p(pts), _p(tets), _p(nbrs), tets.shape[0], _p(q), q.shape[0],
                               _p(st), 0, _p(out), _p(steps), nthreads)
    return out, steps


def cavity(pts, tets, nbrs, q, located, cap=256, nthreads=0):
    """(count (nq,), cavities (nq, cap)) as Triangulation::insert floods them
    (incremental.cpp:148-161): count -1 past cap, 0 where located is -1; row i holds the
    cavity of q[i] in discovery order."""
    pts = _f64(pts).reshape(-1, 3)
    tets = _i32(tets).reshape(-1, 4)
    nbrs = _i32(nbrs).reshape(-1, 4)
    q = _f64(q).reshape(-1, 3)
    located = _i32(located).ravel()
    count = np.zeros(q.shape[0], dtype=np.int32)
    out = np.full((q.shape[0], cap), -1, dtype=np.int32)
    lib().orc_cavity_batch(_p(pts), _p(tets), _p(nbrs), _p(q), q.shape[0], _p(located), cap,
                           _p(count), _p(out), nthreads)
    return count, out
