"""Point-in-domain classification of the PLC carve (cmg_carve_*, cmg_plc_carve).

The classification step of src/constrained/tetrahedralize_plc.cpp:256-268: which
Delaunay tetrahedra (by centroid) or which points lie inside a closed boundary surface.
Two modes: "rays" is the reference's own classifier (CarveGrid, :60-168: majority of
three Moller-Trumbore ray parities) and is bit-identical to it; "parity" is the
voxelizer's exact column parity (src/voxelize/voxelize.cpp:78-195) at the query point.

Device tensors for CarveGrid: points float64 (V, 3), triangles int32 (T, 3), queries
float64 (N, 3) or tets int32 (N, 4); results uint8.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import PLC, _check, _err, lib

_P = ctypes.c_void_p
_i = ctypes.c_int
_sz = ctypes.c_size_t

RAYS, PARITY = 0, 1
_MODES = {"rays": RAYS, "parity": PARITY, RAYS: RAYS, PARITY: PARITY}

_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L = lib()
    L.cmg_carve_create.restype = _i
    L.cmg_carve_create.argtypes = [_P, _sz, _P, _sz, _i, _i, _P, ctypes.POINTER(_P),
                                   ctypes.c_char_p, _sz]
    L.cmg_carve_destroy.argtypes = [_P]
    for f in (L.cmg_carve_points, L.cmg_carve_tets):
        f.restype = _i
        f.argtypes = [_P, _P, _sz, _P, _P, ctypes.c_char_p, _sz]
    L.cmg_carve_status.restype = _i
    L.cmg_carve_status.argtypes = [_P, _P, ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p,
                                   _sz]
    L.cmg_plc_carve.restype = _i
    L.cmg_plc_carve.argtypes = [_P, _P, _sz, _i, _P, ctypes.c_char_p, _sz]
    _bound = True


def _mode(mode) -> int:
    if mode not in _MODES:
        raise ValueError(f"carve mode must be 'rays' or 'parity', not {mode!r}")
    return _MODES[mode]


def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class CarveGrid:
    """The classifier over a closed surface held on the GPU.

    points: float64 (V, 3) and tris: int32 (T, 3), CUDA tensors kept alive by the object.
    mode: "rays" (the reference's CarveGrid) or "parity" (exact column parity over
    make_grid(points, resolution, 1); resolution <= 0 picks ceil(sqrt(T)) capped at 256).
    """

    def __init__(self, points, tris, mode="rays", resolution: int = 0, stream=None):
        import torch
        _bind()
        self.points = points.detach().to(torch.float64).contiguous()
        self.tris = tris.detach().to(torch.int32).contiguous()
        if not self.points.is_cuda or not self.tris.is_cuda:
            raise ValueError("CarveGrid needs CUDA tensors")
        if self.points.dim() != 2 or self.points.shape[1] != 3:
            raise ValueError("points must be (V, 3)")
        if self.tris.dim() != 2 or self.tris.shape[1] != 3:
            raise ValueError("tris must be (T, 3)")
        self.mode = _mode(mode)
        h = _P()
        err = _err()
        st = lib().cmg_carve_create(_P(self.points.data_ptr()), self.points.shape[0],
                                    _P(self.tris.data_ptr()), self.tris.shape[0], self.mode,
                                    int(resolution), _stream(stream), ctypes.byref(h), err,
                                    len(err))
        _check(st, err)
        self._h = h
        self._destroy = lib().cmg_carve_destroy  # usable at interpreter exit

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._destroy(h)

    def points_inside(self, q, out=None, stream=None):
        """uint8 (N,): 1 where q[i] (float64 (N, 3) CUDA tensor) is inside."""
        import torch
        q = q.detach().to(torch.float64).contiguous()
        if q.dim() != 2 or q.shape[1] != 3:
            raise ValueError("q must be (N, 3)")
        if out is None:
            out = torch.empty(q.shape[0], dtype=torch.uint8, device=q.device)
        err = _err()
        _check(lib().cmg_carve_points(self._h, _P(q.data_ptr()), q.shape[0],
                                      _P(out.data_ptr()), _stream(stream), err, len(err)),
               err)
        return out

    def tets_inside(self, tets, out=None, stream=None):
        """uint8 (N,): 1 where the centroid of tets[i] (int32 (N, 4) CUDA tensor of indices
        into the points) is inside; tets with out-of-range indices get 0 (bad_tets())."""
        import torch
        tets = tets.detach().to(torch.int32).contiguous()
        if tets.dim() != 2 or tets.shape[1] != 4:
            raise ValueError("tets must be (N, 4)")
        if out is None:
            out = torch.empty(tets.shape[0], dtype=torch.uint8, device=tets.device)
        err = _err()
        _check(lib().cmg_carve_tets(self._h, _P(tets.data_ptr()), tets.shape[0],
                                    _P(out.data_ptr()), _stream(stream), err, len(err)), err)
        return out

    def bad_tets(self, stream=None) -> int:
        """Tets rejected for out-of-range indices since the last call (synchronises)."""
        n = ctypes.c_uint64(0)
        err = _err()
        _check(lib().cmg_carve_status(self._h, _stream(stream), ctypes.byref(n), err,
                                      len(err)), err)
        return int(n.value)


def plc_carve(plc: PLC, tets, mode="rays") -> np.ndarray:
    """keep flags (uint8, one per tet) of tetrahedralize_plc.cpp:256-268 for host tets
    (int (N, 4) indices into the PLC's points) against the PLC's fan-triangulated
    facets."""
    _bind()
    tets = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
    keep = np.zeros(tets.shape[0], dtype=np.uint8)
    err = _err()
    st = lib().cmg_plc_carve(plc._handle, tets.ctypes.data_as(_P), tets.shape[0],
                             _mode(mode), keep.ctypes.data_as(_P), err, len(err))
    _check(st, err)
    return keep
