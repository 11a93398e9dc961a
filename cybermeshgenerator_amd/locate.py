"""Batched point location in a tetrahedral mesh (cmg_hip_locate): one visibility walk per
query with the rules of the Delaunay insertion's locate
(src/delaunay/incremental.cpp:321-358); and the Delaunay cavity of each located query
(cmg_hip_cavity, the insertion's flood, :148-161).

points float64 (V, 3), tets int32 (T, 4) positively oriented, nbrs int32 (T, 4) (the tet
across the face opposite each vertex, -1 on the hull), queries float64 (N, 3): numpy
arrays (copied to the GPU) or CUDA tensors. Returns (tets found, walk lengths, number of
walks that fell back to brute force); -1 walk length marks those.
"""
from __future__ import annotations

import ctypes

from . import _check, _err, lib

_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L = lib()
    P, sz = ctypes.c_void_p, ctypes.c_size_t
    L.cmg_hip_locate.restype = ctypes.c_int
    L.cmg_hip_locate.argtypes = [P, sz, P, P, sz, P, sz, P, ctypes.c_int, P, P,
                                 ctypes.POINTER(ctypes.c_uint64), P, ctypes.c_char_p, sz]
    L.cmg_hip_cavity.restype = ctypes.c_int
    L.cmg_hip_cavity.argtypes = [P, sz, P, P, sz, P, sz, P, ctypes.c_int, P, P,
                                 ctypes.POINTER(ctypes.c_uint64), P, ctypes.c_char_p, sz]
    _bound = True


def _dev(x, dtype, cols):
    import torch
    t = torch.as_tensor(x, dtype=dtype).reshape(-1, cols) if cols else \
        torch.as_tensor(x, dtype=dtype).reshape(-1)
    return (t if t.is_cuda else t.cuda()).contiguous()


def locate(points, tets, nbrs, queries, start=0, stream=None):
    import torch
    _bind()
    pts = _dev(points, torch.float64, 3)
    tt = _dev(tets, torch.int32, 4)
    nb = _dev(nbrs, torch.int32, 4)
    q = _dev(queries, torch.float64, 3)
    st_arr = None
    start_tet = 0
    if isinstance(start, int):
        start_tet = start
    else:
        st_arr = _dev(start, torch.int32, 0)
    n = q.shape[0]
    out = torch.empty(n, dtype=torch.int32, device=q.device)
    steps = torch.empty(n, dtype=torch.int32, device=q.device)
    brute = ctypes.c_uint64(0)
    s = stream if stream is not None else torch.cuda.current_stream()
    err = _err()
    _check(lib().cmg_hip_locate(
        ctypes.c_void_p(pts.data_ptr()), pts.shape[0], ctypes.c_void_p(tt.data_ptr()),
        ctypes.c_void_p(nb.data_ptr()), tt.shape[0], ctypes.c_void_p(q.data_ptr()), n,
        ctypes.c_void_p(st_arr.data_ptr()) if st_arr is not None else None, start_tet,
        ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(steps.data_ptr()),
        ctypes.byref(brute), ctypes.c_void_p(s.cuda_stream), err, len(err)), err)
    return out, steps, int(brute.value)


def cavity(points, tets, nbrs, queries, located, cap=64, stream=None, count=None, out=None):
    """Delaunay cavity of every query from its located tet (incremental.cpp:148-161):
    returns (count (N,) int32 CUDA tensor: the cavity size, -1 past cap or on a bad index,
    0 where located is -1; cavities (N, cap) int32 CUDA tensor, row i in discovery order
    and -1 after it unless `out` (N, cap) is given, whose entries past each cavity are left
    as they were; number of queries that met an index out of range)."""
    import torch
    _bind()
    pts = _dev(points, torch.float64, 3)
    tt = _dev(tets, torch.int32, 4)
    nb = _dev(nbrs, torch.int32, 4)
    q = _dev(queries, torch.float64, 3)
    loc = _dev(located, torch.int32, 0)
    n = q.shape[0]
    if loc.shape[0] != n:
        raise ValueError("one located tet per query expected")
    if count is None:
        count = torch.empty(n, dtype=torch.int32, device=q.device)
    cav = out if out is not None else torch.full((n, cap), -1, dtype=torch.int32,
                                                 device=q.device)
    if count.shape != (n,) or cav.shape != (n, cap) or not (count.is_contiguous() and
                                                            cav.is_contiguous()):
        raise ValueError("count must be (N,) and out (N, cap), contiguous")
    bad = ctypes.c_uint64(0)
    s = stream if stream is not None else torch.cuda.current_stream()
    err = _err()
    _check(lib().cmg_hip_cavity(
        ctypes.c_void_p(pts.data_ptr()), pts.shape[0], ctypes.c_void_p(tt.data_ptr()),
        ctypes.c_void_p(nb.data_ptr()), tt.shape[0], ctypes.c_void_p(q.data_ptr()), n,
        ctypes.c_void_p(loc.data_ptr()), int(cap), ctypes.c_void_p(count.data_ptr()),
        ctypes.c_void_p(cav.data_ptr()), ctypes.byref(bad), ctypes.c_void_p(s.cuda_stream),
        err, len(err)), err)
    return count, cav, int(bad.value)
