"""Surface files to PLCs and to device arrays (cmg_read_plc, cmg_plc_points /
cmg_plc_triangles, cmg_hip_stl_decode, cmg_hip_weld).

read_plc follows cybermesh.read_plc / cmg::io::read_plc for .stl and .obj
(src/io/stl.cpp, src/io/obj.cpp): same grammar, same messages, coincident STL vertices
merged in first-occurrence order (on the GPU). read_stl_device keeps a binary STL on the
GPU end to end: the file's bytes are copied to device memory once and decoded and merged
there into the float64 points and int32 triangles the voxelizer and the carve take.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import PLC, _check, _err, lib

_P = ctypes.c_void_p
_i = ctypes.c_int
_sz = ctypes.c_size_t

_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L = lib()
    L.cmg_read_plc.restype = _i
    L.cmg_read_plc.argtypes = [ctypes.c_char_p, ctypes.POINTER(_P), ctypes.c_char_p, _sz]
    L.cmg_plc_num_points.restype = _sz
    L.cmg_plc_num_points.argtypes = [_P]
    L.cmg_plc_points.restype = ctypes.POINTER(ctypes.c_double)
    L.cmg_plc_points.argtypes = [_P]
    L.cmg_plc_num_triangles.restype = _sz
    L.cmg_plc_num_triangles.argtypes = [_P]
    L.cmg_plc_triangles.restype = ctypes.POINTER(ctypes.c_int)
    L.cmg_plc_triangles.argtypes = [_P]
    L.cmg_hip_stl_decode.restype = _i
    L.cmg_hip_stl_decode.argtypes = [_P, _sz, _P, _P, ctypes.POINTER(ctypes.c_int64), _P,
                                     ctypes.c_char_p, _sz]
    L.cmg_hip_weld.restype = _i
    L.cmg_hip_weld.argtypes = [_P, _sz, _P, _P, ctypes.POINTER(ctypes.c_int64), _P,
                               ctypes.c_char_p, _sz]
    _bound = True


def plc_geometry(handle) -> tuple:
    """(points (V,3) float64, triangles (T,3) int32) of a cmg_plc handle."""
    _bind()
    L = lib()
    n = L.cmg_plc_num_points(handle)
    pts = (np.ctypeslib.as_array(L.cmg_plc_points(handle), shape=(3 * n,)).reshape(-1, 3)
           .copy() if n else np.zeros((0, 3)))
    m = L.cmg_plc_num_triangles(handle)
    tris = (np.ctypeslib.as_array(L.cmg_plc_triangles(handle), shape=(3 * m,))
            .reshape(-1, 3).astype(np.int32) if m else np.zeros((0, 3), np.int32))
    return pts, tris


def read_plc(path: str) -> PLC:
    """A PLC read from a .stl or .obj file (cmg_read_plc); RuntimeError with the
    reference's message on failure."""
    _bind()
    h = _P()
    err = _err()
    _check(lib().cmg_read_plc(os.fsencode(path), ctypes.byref(h), err, len(err)), err)
    plc = PLC.__new__(PLC)
    plc._handle = h.value
    plc._destroy = lib().cmg_plc_destroy
    pts, tris = plc_geometry(h)
    plc._points = pts
    plc._facets = [t.copy() for t in tris]
    return plc


def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def stl_decode(data, stream=None):
    """A binary STL held as a uint8 CUDA tensor (the whole file) -> (points float64 (V,3),
    triangles int32 (T,3)) CUDA tensors, merged as the reference merges them."""
    import torch
    _bind()
    if data.dtype != torch.uint8 or not data.is_cuda or data.dim() != 1:
        raise ValueError("data must be a 1-D uint8 CUDA tensor holding the whole file")
    data = data.contiguous()
    if data.numel() < 84:
        raise ValueError("not a binary STL (shorter than its header)")
    n = int(np.frombuffer(bytes(data[80:84].cpu().numpy()), dtype="<u4")[0])
    xyz = torch.empty((max(3 * n, 1), 3), dtype=torch.float64, device=data.device)
    tris = torch.empty((max(n, 1), 3), dtype=torch.int32, device=data.device)
    npts = ctypes.c_int64(0)
    err = _err()
    _check(lib().cmg_hip_stl_decode(_P(data.data_ptr()), data.numel(), _P(xyz.data_ptr()),
                                    _P(tris.data_ptr()), ctypes.byref(npts),
                                    _stream(stream), err, len(err)), err)
    return xyz[: npts.value], tris[:n]


def read_stl_device(path: str, stream=None):
    """A binary STL file straight to device arrays: (points, triangles) CUDA tensors."""
    import torch
    raw = np.fromfile(path, dtype=np.uint8)
    host = torch.from_numpy(raw).pin_memory()
    dev = torch.empty(raw.size, dtype=torch.uint8, device="cuda")
    dev.copy_(host, non_blocking=True)
    return stl_decode(dev, stream)


def weld(soup, stream=None):
    """Merge coincident vertices of a float64 (N,3) CUDA tensor -> (points (V,3), idx (N,))
    in first-occurrence order (src/io/stl.cpp:24-61)."""
    import torch
    _bind()
    soup = soup.detach().to(torch.float64).contiguous()
    if soup.dim() != 2 or soup.shape[1] != 3 or not soup.is_cuda:
        raise ValueError("soup must be a float64 (N, 3) CUDA tensor")
    n = soup.shape[0]
    xyz = torch.empty_like(soup)
    idx = torch.empty(n, dtype=torch.int32, device=soup.device)
    npts = ctypes.c_int64(0)
    err = _err()
    _check(lib().cmg_hip_weld(_P(soup.data_ptr()), n, _P(xyz.data_ptr()), _P(idx.data_ptr()),
                              ctypes.byref(npts), _stream(stream), err, len(err)), err)
    return xyz[: npts.value], idx
