"""The mesh quality report (cmg_hip_quality_report): cmg::quality::report
(src/quality/metrics.cpp:81-169) over a tetrahedral mesh on the GPU.

Points float64 (V, 3) and tets int32 (T, 4), numpy arrays (copied to the GPU) or CUDA
tensors (used in place). Returns the QualityReport fields as a dict, with worst_tets the
indices of the largest radius-edge ratios, largest first.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _check, _err, lib

FIELDS = ["min_radius_edge", "max_radius_edge", "mean_radius_edge", "min_dihedral",
          "max_dihedral", "mean_dihedral", "min_volume", "max_volume", "total_volume"]


class _Report(ctypes.Structure):
    _fields_ = ([("num_tets", ctypes.c_uint64)] + [(f, ctypes.c_double) for f in FIELDS] +
                [("dihedral_histogram", ctypes.c_int * 18), ("num_worst", ctypes.c_uint64)])


_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L = lib()
    P, sz = ctypes.c_void_p, ctypes.c_size_t
    L.cmg_hip_quality_report.restype = ctypes.c_int
    L.cmg_hip_quality_report.argtypes = [P, sz, P, sz, ctypes.c_int,
                                         ctypes.POINTER(_Report), P, P, ctypes.c_char_p, sz]
    _bound = True


def report(points, tets, worst_count: int = 10, stream=None) -> dict:
    import torch
    _bind()
    pts = torch.as_tensor(points, dtype=torch.float64).reshape(-1, 3)
    tt = torch.as_tensor(tets, dtype=torch.int32).reshape(-1, 4)
    pts = (pts if pts.is_cuda else pts.cuda()).contiguous()
    tt = (tt if tt.is_cuda else tt.cuda()).contiguous()
    n = tt.shape[0]
    cap = n if worst_count < 0 else min(worst_count, n)
    worst = np.zeros(max(cap, 1), dtype=np.int32)
    r = _Report()
    s = stream if stream is not None else torch.cuda.current_stream()
    err = _err()
    _check(lib().cmg_hip_quality_report(
        ctypes.c_void_p(pts.data_ptr()), pts.shape[0], ctypes.c_void_p(tt.data_ptr()), n,
        int(worst_count), ctypes.byref(r), worst.ctypes.data_as(ctypes.c_void_p),
        ctypes.c_void_p(s.cuda_stream), err, len(err)), err)
    out = {"num_tets": int(r.num_tets)}
    out.update({f: getattr(r, f) for f in FIELDS})
    out["dihedral_histogram"] = list(r.dihedral_histogram)
    out["worst_tets"] = worst[: r.num_worst].tolist()
    return out
