"""cybermeshgenerator_amd — CyberMeshGenerator's predicate and voxelization path on
AMD Instinct MI355X.

A ctypes binding over libcmg_hip.so with the surface of the reference's ``cybermesh``
package for this path (bindings/python/src/cybermesh/__init__.py:416-494): ``PLC``,
``voxelize`` and ``VoxelGrid``, plus the batched predicates and the SDF adjoint. Every
compute call runs the hand-written HIP kernels; there is no CPU path, and importing
fails loudly if the library has not been built
(``python -m cybermeshgenerator_amd.build``).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcmg_hip.so")

EXACT = 0
COMPAT = 1
_MODES = {"exact": EXACT, "compat": COMPAT, EXACT: EXACT, COMPAT: COMPAT}

_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_void_pp = ctypes.POINTER(ctypes.c_void_p)
_P = ctypes.c_void_p
_sz = ctypes.c_size_t
_i = ctypes.c_int
_errbuf_t = ctypes.c_char_p


class PredStats(ctypes.Structure):
    """Per-stage lane counts of a batched-sign call (cmg_pred_stats)."""
    _fields_ = [("lanes", ctypes.c_uint64), ("fp32", ctypes.c_uint64),
                ("fp64", ctypes.c_uint64), ("exact", ctypes.c_uint64),
                ("wide", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("exact_inline", ctypes.c_uint64)]

    def as_dict(self) -> dict:
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


def _init_torch_runtime_first() -> None:
    """torch ships its own HIP and HSA runtime libraries; this library links the system
    ROCm ones, so a process that uses both holds two runtimes. torch's finds no device
    when it starts after the other, so it is started first whenever torch is present."""
    try:
        import torch
    except ImportError:
        return
    torch.cuda.is_available()


def _load() -> ctypes.CDLL:
    _init_torch_runtime_first()
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m cybermeshgenerator_amd.build`"
            " (this package has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    lib.cmg_plc_create.restype = _P
    lib.cmg_plc_destroy.argtypes = [_P]
    lib.cmg_plc_set_points.restype = _i
    lib.cmg_plc_set_points.argtypes = [_P, _c_double_p, _sz]
    lib.cmg_plc_add_facet.restype = _i
    lib.cmg_plc_add_facet.argtypes = [_P, _c_int_p, _sz, _i]
    lib.cmg_plc_voxelize.restype = _i
    lib.cmg_plc_voxelize.argtypes = [_P, _i, _i, _i, _c_void_pp, _errbuf_t, _sz]
    lib.cmg_voxels_destroy.argtypes = [_P]
    lib.cmg_voxels_dims.argtypes = [_P, _c_int_p, _c_int_p, _c_int_p]
    lib.cmg_voxels_origin.argtypes = [_P, _c_double_p, _c_double_p, _c_double_p]
    lib.cmg_voxels_spacing.restype = ctypes.c_double
    lib.cmg_voxels_spacing.argtypes = [_P]
    lib.cmg_voxels_occupancy.restype = ctypes.POINTER(ctypes.c_ubyte)
    lib.cmg_voxels_occupancy.argtypes = [_P]
    lib.cmg_voxels_distance.restype = ctypes.POINTER(ctypes.c_float)
    lib.cmg_voxels_distance.argtypes = [_P]
    lib.cmg_version.restype = ctypes.c_char_p
    lib.cmg_orient3d_signs.restype = _i
    lib.cmg_orient3d_signs.argtypes = [_P, _sz, _P, _sz, _i, _P, _errbuf_t, _sz]
    lib.cmg_insphere_signs.restype = _i
    lib.cmg_insphere_signs.argtypes = [_P, _sz, _P, _sz, _i, _P, _errbuf_t, _sz]
    lib.cmg_orient3d_values.restype = _i
    lib.cmg_orient3d_values.argtypes = [_P, _sz, _P, _sz, _P, _errbuf_t, _sz]
    lib.cmg_insphere_values.restype = _i
    lib.cmg_insphere_values.argtypes = [_P, _sz, _P, _sz, _P, _errbuf_t, _sz]
    lib.cmg_voxels_sdf_backward.restype = _i
    lib.cmg_voxels_sdf_backward.argtypes = [_P, _P, _P, _P, _errbuf_t, _sz]
    lib.cmg_voxels_sdf_nearest.restype = _i
    lib.cmg_voxels_sdf_nearest.argtypes = [_P, _P, _P, _errbuf_t, _sz]
    lib.cmg_backend_last.restype = ctypes.c_char_p
    lib.cmg_hip_available.restype = _i
    lib.cmg_hip_orient3d_signs.restype = _i
    lib.cmg_hip_orient3d_signs.argtypes = [_P, _i, _P, _i, _P]
    lib.cmg_hip_pred_workspace_size.restype = _sz
    lib.cmg_hip_pred_workspace_size.argtypes = [_sz, _i]
    for f in (lib.cmg_hip_orient3d_signs_soa, lib.cmg_hip_insphere_signs_soa):
        f.restype = _i
        f.argtypes = [_P, _sz, _sz, _i, _P, _P, _sz, _P, _errbuf_t, _sz]
    lib.cmg_hip_pred_stats.restype = _i
    lib.cmg_hip_pred_stats.argtypes = [_P, _sz, _i, _P, ctypes.POINTER(PredStats),
                                       _errbuf_t, _sz]
    lib.cmg_hip_kernel_timer.argtypes = [ctypes.c_char_p, _i]
    lib.cmg_hip_kernel_time.restype = _i
    lib.cmg_hip_kernel_time.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_uint64), _errbuf_t, _sz]
    return lib


class _LazyLib:
    """Loads libcmg_hip on first use, so that the package (and its build module) can be
    imported before the library exists; any call then fails loudly if it is missing."""

    _cdll = None

    def __getattr__(self, name):
        if _LazyLib._cdll is None:
            _LazyLib._cdll = _load()
        return getattr(_LazyLib._cdll, name)


_lib = _LazyLib()


def lib() -> ctypes.CDLL:
    """The loaded libcmg_hip (for callers binding further symbols)."""
    if _LazyLib._cdll is None:
        _LazyLib._cdll = _load()
    return _LazyLib._cdll


def _check(status: int, err) -> None:
    if status != 0:
        msg = err.value.decode(errors="replace") if err is not None else ""
        raise RuntimeError(msg or f"cmg error {status}")


def _err():
    return ctypes.create_string_buffer(512)


def hip_available() -> bool:
    return bool(_lib.cmg_hip_available())


def last_backend() -> str:
    """Backend that served this thread's last call (cmg::backend::last_backend)."""
    return _lib.cmg_backend_last().decode()


def kernel_timer(name: str, enable: bool = True) -> None:
    """Start/stop HIP-event timing of the named kernel's launches."""
    _lib.cmg_hip_kernel_timer(name.encode(), int(enable))


def kernel_time(name: str) -> tuple:
    """(total_ms, launches) of the named kernel since the last read."""
    ms, cnt = ctypes.c_double(), ctypes.c_uint64()
    err = _err()
    _check(_lib.cmg_hip_kernel_time(name.encode(), ctypes.byref(ms), ctypes.byref(cnt), err,
                                    len(err)), err)
    return ms.value, int(cnt.value)


def version() -> str:
    return _lib.cmg_version().decode()


def _mode(mode) -> int:
    try:
        return _MODES[mode]
    except KeyError:
        raise ValueError("mode must be 'exact'/0 or 'compat'/1") from None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---- batched predicates (host arrays) -----------------------------------------------

def orient3d_signs(points, quads, mode="compat") -> np.ndarray:
    """int8 orient3d signs for index quadruples into (N, 3) points — the batch of
    cmg::robust::orient3d_signs (src/predicates/robust.cpp:68-105). mode 'compat' gives
    what cmg::robust::orient3d returns; 'exact' the exact sign."""
    pts = _f64(points).reshape(-1, 3)
    q = _i32(quads).reshape(-1, 4)
    out = np.empty(q.shape[0], dtype=np.int8)
    err = _err()
    _check(_lib.cmg_orient3d_signs(_ptr(pts), pts.shape[0], _ptr(q), q.shape[0],
                                   _mode(mode), _ptr(out), err, len(err)), err)
    return out


def insphere_signs(points, quints, mode="compat") -> np.ndarray:
    """int8 insphere signs for index quintuples (a, b, c, d, e) into (N, 3) points."""
    pts = _f64(points).reshape(-1, 3)
    q = _i32(quints).reshape(-1, 5)
    out = np.empty(q.shape[0], dtype=np.int8)
    err = _err()
    _check(_lib.cmg_insphere_signs(_ptr(pts), pts.shape[0], _ptr(q), q.shape[0],
                                   _mode(mode), _ptr(out), err, len(err)), err)
    return out


def orient3d_batch(points, quads) -> np.ndarray:
    """float64 orient3d values (cmg::robust::orient3d_batch, robust.cpp:52-66)."""
    pts = _f64(points).reshape(-1, 3)
    q = _i32(quads).reshape(-1, 4)
    out = np.empty(q.shape[0], dtype=np.float64)
    err = _err()
    _check(_lib.cmg_orient3d_values(_ptr(pts), pts.shape[0], _ptr(q), q.shape[0],
                                    _ptr(out), err, len(err)), err)
    return out


def insphere_values(points, quints) -> np.ndarray:
    """float64 insphere values for index quintuples (a, b, c, d, e): what
    cmg::robust::insphere returns (robust.cpp:44-50) — the naive FP64 determinant when it
    is nonzero, else the exact determinant's leading component."""
    pts = _f64(points).reshape(-1, 3)
    q = _i32(quints).reshape(-1, 5)
    out = np.empty(q.shape[0], dtype=np.float64)
    err = _err()
    _check(_lib.cmg_insphere_values(_ptr(pts), pts.shape[0], _ptr(q), q.shape[0],
                                    _ptr(out), err, len(err)), err)
    return out


# ---- batched predicates (device tensors) --------------------------------------------

def pred_workspace_size(n: int, npts: int = 4) -> int:
    return int(_lib.cmg_hip_pred_workspace_size(n, npts))


def _stream_handle(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def signs_soa(soa, mode="exact", out=None, workspace=None, stream=None):
    """Batched signs over a device tensor of SoA tuples: soa is (12, n) float64 for
    orient3d or (15, n) for insphere (row q = slot*3 + axis). Returns an int8 tensor.
    With a caller workspace (uint8 tensor of pred_workspace_size bytes) the call is
    stream-ordered; without one it synchronises."""
    import torch
    if soa.dtype != torch.float64 or soa.dim() != 2 or soa.shape[0] not in (12, 15):
        raise ValueError("soa must be a (12, n) or (15, n) float64 tensor")
    if not soa.is_cuda:
        raise ValueError("soa must be a device tensor")
    if soa.stride(1) != 1:
        raise ValueError("soa rows must be contiguous")
    n = soa.shape[1]
    ld = soa.stride(0)
    if out is None:
        out = torch.empty(n, dtype=torch.int8, device=soa.device)
    fn = _lib.cmg_hip_orient3d_signs_soa if soa.shape[0] == 12 else \
        _lib.cmg_hip_insphere_signs_soa
    ws_ptr, ws_bytes = None, 0
    if workspace is not None:
        ws_ptr, ws_bytes = workspace.data_ptr(), workspace.numel() * workspace.element_size()
    err = _err()
    _check(fn(soa.data_ptr(), n, ld, _mode(mode), out.data_ptr(), ws_ptr, ws_bytes,
              _stream_handle(stream), err, len(err)), err)
    return out


def pred_stats(n: int, npts: int = 4, workspace=None, stream=None) -> dict:
    """Per-stage lane counts of the last signs_soa call over n tuples on `workspace`."""
    st = PredStats()
    err = _err()
    _check(_lib.cmg_hip_pred_stats(workspace.data_ptr() if workspace is not None else None,
                                   n, npts, _stream_handle(stream), ctypes.byref(st), err,
                                   len(err)), err)
    return st.as_dict()


# ---- PLC / voxelize (the cybermesh surface) -----------------------------------------

class PLC:
    """A piecewise linear complex: points and polygonal facets (cmg_plc handle)."""

    def __init__(self):
        self._handle = _lib.cmg_plc_create()
        self._destroy = _lib.cmg_plc_destroy  # usable at interpreter exit
        self._points = np.zeros((0, 3))
        self._facets: list[np.ndarray] = []

    def __del__(self):
        h = getattr(self, "_handle", None)
        destroy = getattr(self, "_destroy", None)
        if h and destroy is not None:
            destroy(h)
            self._handle = None

    def add_points(self, points) -> "PLC":
        """Set the vertex cloud from an (N, 3) array. Replaces any prior set."""
        pts = _f64(points)
        if pts.ndim != 2 or pts.shape[1] != 3:
            raise ValueError("points must be an (N, 3) array")
        _check(_lib.cmg_plc_set_points(self._handle, pts.ctypes.data_as(_c_double_p),
                                       pts.shape[0]), None)
        self._points = pts.copy()
        return self

    def add_facet(self, vertex_indices: Sequence[int], marker: int = 0) -> "PLC":
        """Append a single-polygon facet (a triangle when three indices)."""
        idx = _i32(vertex_indices).ravel()
        if idx.size < 3:
            raise ValueError("a facet needs at least 3 vertex indices")
        _check(_lib.cmg_plc_add_facet(self._handle, idx.ctypes.data_as(_c_int_p), idx.size,
                                      int(marker)), None)
        self._facets.append(idx.copy())
        return self

    def add_triangles(self, tris) -> "PLC":
        """Append every row of an (M, 3) index array as a triangle facet."""
        for t in _i32(tris).reshape(-1, 3):
            self.add_facet(t)
        return self

    @property
    def points(self) -> np.ndarray:
        return self._points.copy()

    @property
    def triangles(self) -> np.ndarray:
        """Facets fan-triangulated to an (M, 3) int32 array (voxelize.cpp:29-38)."""
        out = []
        for v in self._facets:
            for i in range(2, v.size):
                out.append((v[0], v[i - 1], v[i]))
        return np.array(out, dtype=np.int32).reshape(-1, 3)


@dataclass
class VoxelGrid:
    """A regular voxel grid returned by :func:`voxelize`: ``grid`` is (nx, ny, nz), uint8
    occupancy or float32 signed distance (negative inside); ``origin`` is the centre of
    cell (0, 0, 0) (include/cmg/voxelize/voxelize.hpp:40)."""

    grid: np.ndarray
    origin: np.ndarray
    spacing: float
    mode: str

    @property
    def dims(self) -> tuple:
        return tuple(int(d) for d in self.grid.shape)


_VOXEL_MODES = {"occupancy": 0, "sdf": 1}


def _voxelize_handle(plc: PLC, resolution: int, mode_code: int, pad: int):
    out = ctypes.c_void_p()
    err = _err()
    st = _lib.cmg_plc_voxelize(plc._handle, int(resolution), mode_code, int(pad),
                               ctypes.byref(out), err, len(err))
    _check(st, err)
    return out.value


def _grid_info(handle):
    nx, ny, nz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _lib.cmg_voxels_dims(handle, ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nz))
    ox, oy, oz = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    _lib.cmg_voxels_origin(handle, ctypes.byref(ox), ctypes.byref(oy), ctypes.byref(oz))
    spacing = float(_lib.cmg_voxels_spacing(handle))
    return (nx.value, ny.value, nz.value), np.array([ox.value, oy.value, oz.value]), spacing


def voxelize(plc: PLC, resolution: int = 64, mode="occupancy", pad: int = 1,
             flat: bool = False) -> VoxelGrid:
    """Voxelize a PLC (cybermesh.voxelize, __init__.py:439-494): ``mode`` "occupancy"/0
    gives a uint8 inside/outside grid, "sdf"/1 a float32 signed distance (negative
    inside). The grid is (nx, ny, nz); with flat=True it is the C ABI's flat array,
    idx = (k*ny + j)*nx + i."""
    if not isinstance(plc, PLC):
        raise TypeError("plc must be a PLC")
    if isinstance(mode, str):
        if mode not in _VOXEL_MODES:
            raise ValueError(f"mode must be one of {sorted(_VOXEL_MODES)} or 0/1")
        mode_code, mode_name = _VOXEL_MODES[mode], mode
    else:
        mode_code = int(mode)
        if mode_code not in (0, 1):
            raise ValueError("mode must be 'occupancy'/0 or 'sdf'/1")
        mode_name = "occupancy" if mode_code == 0 else "sdf"
    handle = _voxelize_handle(plc, resolution, mode_code, pad)
    try:
        (nx, ny, nz), origin, spacing = _grid_info(handle)
        n = nx * ny * nz
        if mode_code == 0:
            grid = np.ctypeslib.as_array(_lib.cmg_voxels_occupancy(handle), (n,)).copy()
        else:
            grid = np.ctypeslib.as_array(_lib.cmg_voxels_distance(handle), (n,)).copy()
        if not flat:
            grid = grid.reshape((nz, ny, nx)).transpose(2, 1, 0).copy()
    finally:
        _lib.cmg_voxels_destroy(handle)
    return VoxelGrid(grid=grid, origin=origin, spacing=spacing, mode=mode_name)


def sdf_with_grad(plc: PLC, resolution: int, pad: int, grad_sdf) -> tuple:
    """Signed distance (flat float32) and dL/dpoints (N, 3) float32 for an upstream
    gradient grad_sdf (flat, C ABI order), through cmg_voxels_sdf_backward."""
    handle = _voxelize_handle(plc, resolution, 1, pad)
    try:
        (nx, ny, nz), origin, spacing = _grid_info(handle)
        n = nx * ny * nz
        sdf = np.ctypeslib.as_array(_lib.cmg_voxels_distance(handle), (n,)).copy()
        g = np.ascontiguousarray(grad_sdf, dtype=np.float32).ravel()
        if g.size != n:
            raise ValueError("grad_sdf must have one value per cell")
        grad = np.zeros((plc._points.shape[0], 3), dtype=np.float32)
        err = _err()
        _check(_lib.cmg_voxels_sdf_backward(plc._handle, handle, _ptr(g), _ptr(grad), err,
                                            len(err)), err)
    finally:
        _lib.cmg_voxels_destroy(handle)
    return sdf, grad


def sdf_nearest(plc: PLC, resolution: int, pad: int) -> np.ndarray:
    """Per cell (flat, C ABI order), the index of the triangle whose closest point the
    SDF adjoint uses (the FP32 nearest triangle; -1 where there is none), through
    cmg_voxels_sdf_nearest."""
    handle = _voxelize_handle(plc, resolution, 1, pad)
    try:
        (nx, ny, nz), _, _ = _grid_info(handle)
        out = np.full(nx * ny * nz, -1, dtype=np.int32)
        err = _err()
        _check(_lib.cmg_voxels_sdf_nearest(plc._handle, handle, _ptr(out), err, len(err)),
               err)
    finally:
        _lib.cmg_voxels_destroy(handle)
    return out


def plc_from_mesh(points, tris) -> PLC:
    """A PLC whose facets are the rows of an (M, 3) triangle array."""
    p = PLC()
    p.add_points(points)
    p.add_triangles(tris)
    return p
