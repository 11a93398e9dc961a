"""Batched signed-distance voxelization of many meshes (cmg_voxelize_batch_*), and the
differentiable single-mesh SDF behind a torch.autograd.Function.

Device tensors throughout: points float64 (V, 3), triangles int32 (T, 3), the SDF of
all grids concatenated as one float32 tensor, gradients float32 (V, 3).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _check, _err, lib

_P = ctypes.c_void_p
_i = ctypes.c_int
_sz = ctypes.c_size_t

FORWARD, BACKWARD, GRAPH = 1, 2, 4

_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L = lib()
    L.cmg_voxelize_batch_create.restype = _i
    L.cmg_voxelize_batch_create.argtypes = [_i, _P, _P, _P, _P, _P, _P, _P,
                                            ctypes.POINTER(_P), ctypes.c_char_p, _sz]
    L.cmg_voxelize_batch_destroy.argtypes = [_P]
    L.cmg_voxelize_batch_cells.restype = ctypes.c_int64
    L.cmg_voxelize_batch_cells.argtypes = [_P]
    L.cmg_voxelize_batch_grid.restype = _i
    L.cmg_voxelize_batch_grid.argtypes = [_P, _i, _P, _P, _P, _P]
    L.cmg_voxelize_batch_run.restype = _i
    L.cmg_voxelize_batch_run.argtypes = [_P, _i, _P, _P, _P, _P, _P, ctypes.c_char_p, _sz]
    L.cmg_voxelize_batch_status.restype = _i
    L.cmg_voxelize_batch_status.argtypes = [_P, _P, ctypes.POINTER(ctypes.c_uint),
                                            ctypes.POINTER(ctypes.c_uint64),
                                            ctypes.c_char_p, _sz]
    L.cmg_voxelize_batch_autotune.restype = _i
    L.cmg_voxelize_batch_autotune.argtypes = [_P, _P, _P, _P, _P, ctypes.POINTER(_i),
                                              ctypes.POINTER(_i), _P, ctypes.c_char_p, _sz]
    L.cmg_voxelize_batch_set_variant.restype = _i
    L.cmg_voxelize_batch_set_variant.argtypes = [_P, _i, _i]
    L.cmg_voxelize_batch_occupancy.restype = _P
    L.cmg_voxelize_batch_occupancy.argtypes = [_P]
    L.cmg_voxelize_batch_nearest.restype = _i
    L.cmg_voxelize_batch_nearest.argtypes = [_P, _P, _P, ctypes.c_char_p, _sz]
    L.cmg_voxelize_batch_counters.restype = _i
    L.cmg_voxelize_batch_counters.argtypes = [_P, _P, _P, ctypes.c_char_p, _sz]
    _bound = True


def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class Grid:
    def __init__(self, dims, origin, spacing, offset):
        self.dims = tuple(int(d) for d in dims)
        self.origin = np.array(origin, dtype=np.float64)
        self.spacing = float(spacing)
        self.offset = int(offset)

    @property
    def ncell(self):
        nx, ny, nz = self.dims
        return nx * ny * nz


class SdfBatch:
    """Many closed meshes voxelized to signed distance together on one GPU.

    meshes: list of (points (V,3), triangles (T,3)); resolutions / pads: one per mesh
    (cmg_plc_voxelize's resolution and pad). The points and triangles are copied into
    device tensors owned by the batch; `points` can be updated in place between steps
    (same shapes) to move vertices."""

    def __init__(self, meshes, resolutions, pads=None, device="cuda", stream=None):
        import torch
        _bind()
        n = len(meshes)
        if n == 0:
            raise ValueError("empty batch")
        pads = [0] * n if pads is None else list(pads)
        pts = [torch.as_tensor(p, dtype=torch.float64, device=device).reshape(-1, 3)
               for p, _ in meshes]
        tri = [torch.as_tensor(t, dtype=torch.int32, device=device).reshape(-1, 3)
               for _, t in meshes]
        self.point_off = np.zeros(n + 1, dtype=np.int64)
        self.tri_off = np.zeros(n + 1, dtype=np.int64)
        for m in range(n):
            self.point_off[m + 1] = self.point_off[m] + pts[m].shape[0]
            self.tri_off[m + 1] = self.tri_off[m] + tri[m].shape[0]
        self.points = torch.cat(pts).contiguous()
        self.tris = torch.cat(tri).contiguous()
        self.res = np.ascontiguousarray(resolutions, dtype=np.int32)
        self.pads = np.ascontiguousarray(pads, dtype=np.int32)
        h = _P()
        err = _err()
        st = lib().cmg_voxelize_batch_create(
            n, self.point_off.ctypes.data_as(_P), self.tri_off.ctypes.data_as(_P),
            self.res.ctypes.data_as(_P), self.pads.ctypes.data_as(_P),
            self.points.data_ptr(), self.tris.data_ptr(), _stream(stream), ctypes.byref(h),
            err, len(err))
        _check(st, err)
        self._h = h.value
        # held by the object: module globals may already be gone when it is collected at
        # interpreter exit
        self._destroy = lib().cmg_voxelize_batch_destroy
        self.n = n
        self.cells = int(lib().cmg_voxelize_batch_cells(self._h))
        self.grids = []
        for m in range(n):
            dims = np.zeros(3, dtype=np.int32)
            org = np.zeros(3)
            sp = ctypes.c_double()
            off = ctypes.c_int64()
            lib().cmg_voxelize_batch_grid(self._h, m, dims.ctypes.data_as(_P),
                                          org.ctypes.data_as(_P), ctypes.byref(sp),
                                          ctypes.byref(off))
            self.grids.append(Grid(dims, org, sp.value, off.value))
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def run(self, flags, sdf=None, grad_sdf=None, grad=None, stream=None):
        err = _err()
        _check(lib().cmg_voxelize_batch_run(
            self._h, flags, self.points.data_ptr(),
            sdf.data_ptr() if sdf is not None else None,
            grad_sdf.data_ptr() if grad_sdf is not None else None,
            grad.data_ptr() if grad is not None else None, _stream(stream), err, len(err)),
            err)

    def forward(self, out=None, graph=True, stream=None):
        """Signed distance of every grid, concatenated (float32, cells)."""
        import torch
        if out is None:
            out = torch.empty(self.cells, dtype=torch.float32, device=self.device)
        self.run(FORWARD | (GRAPH if graph else 0), sdf=out, stream=stream)
        return out

    def backward(self, grad_sdf, out=None, graph=True, stream=None):
        """dL/dpoints (float32, (V, 3)) for L = sum grad_sdf * sdf, after a forward."""
        import torch
        if out is None:
            out = torch.empty((self.points.shape[0], 3), dtype=torch.float32,
                              device=self.device)
        g = grad_sdf.contiguous()
        self.run(BACKWARD | (GRAPH if graph else 0), grad_sdf=g, grad=out, stream=stream)
        return out

    def step(self, grad_sdf, sdf=None, grad=None, graph=True, stream=None):
        """Forward and backward in one launch sequence (one graph when graph=True)."""
        import torch
        if sdf is None:
            sdf = torch.empty(self.cells, dtype=torch.float32, device=self.device)
        if grad is None:
            grad = torch.empty((self.points.shape[0], 3), dtype=torch.float32,
                               device=self.device)
        self.run(FORWARD | BACKWARD | (GRAPH if graph else 0), sdf=sdf,
                 grad_sdf=grad_sdf.contiguous(), grad=grad, stream=stream)
        return sdf, grad

    def autotune(self, grad_sdf, sdf=None, grad=None, stream=None) -> dict:
        """Times the six voxel-pass variants on this batch and keeps the fastest; returns
        {"vpl": .., "depth": .., "ms": {(vpl, depth): ms}}."""
        import torch
        if sdf is None:
            sdf = torch.empty(self.cells, dtype=torch.float32, device=self.device)
        if grad is None:
            grad = torch.empty((self.points.shape[0], 3), dtype=torch.float32,
                               device=self.device)
        vpl, depth = _i(), _i()
        ms = np.zeros(6, dtype=np.float32)
        err = _err()
        _check(lib().cmg_voxelize_batch_autotune(
            self._h, sdf.data_ptr(), grad_sdf.contiguous().data_ptr(), grad.data_ptr(),
            _stream(stream), ctypes.byref(vpl), ctypes.byref(depth), ms.ctypes.data_as(_P),
            err, len(err)), err)
        keys = [(1, 0), (1, 8), (2, 0), (2, 8), (4, 0), (4, 8)]
        self._variant = (vpl.value, depth.value)
        return {"vpl": vpl.value, "depth": depth.value,
                "ms": {f"{v}x{d}": float(t) for (v, d), t in zip(keys, ms)}}

    def set_variant(self, vpl: int, depth: int):
        _check(lib().cmg_voxelize_batch_set_variant(self._h, vpl, depth), None)
        self._variant = (vpl, depth)

    def variant_info(self) -> dict:
        """The voxel-pass variant in use: voxels per lane and LDS tile depth."""
        v, d = getattr(self, "_variant", (1, 0))
        return {"vpl": v, "depth": d}

    def vpl(self) -> int:
        return getattr(self, "_variant", (1, 0))[0]

    def status(self, stream=None) -> dict:
        ov = ctypes.c_uint()
        pairs = ctypes.c_uint64()
        err = _err()
        _check(lib().cmg_voxelize_batch_status(self._h, _stream(stream), ctypes.byref(ov),
                                               ctypes.byref(pairs), err, len(err)), err)
        return {"overflow": int(ov.value), "pairs": int(pairs.value)}

    def counters(self, stream=None) -> dict:
        """Work counters of the last step: per wavefront of the voxel passes, node box
        tests, triangle box tests and FP32 triangle distances; the overflow flags; and
        the occupancy pass's column-tile list entries."""
        out = np.zeros(5, dtype=np.uint64)
        err = _err()
        _check(lib().cmg_voxelize_batch_counters(self._h, _stream(stream),
                                                 out.ctypes.data_as(_P), err, len(err)), err)
        return {"node_tests": int(out[0]), "tri_tests": int(out[1]),
                "tri_evals": int(out[2]), "overflow": int(out[3]),
                "tile_entries": int(out[4])}

    def nearest(self, stream=None):
        """After a forward: per cell (device int32, sdf layout) the mesh-local index of
        the triangle the adjoint uses (-1 where none)."""
        import torch
        out = torch.empty(self.cells, dtype=torch.int32, device=self.device)
        err = _err()
        _check(lib().cmg_voxelize_batch_nearest(self._h, out.data_ptr(), _stream(stream),
                                                err, len(err)), err)
        return out

    def split(self, flat):
        """Per-mesh (nx, ny, nz) views of a concatenated cell array."""
        out = []
        for g in self.grids:
            nx, ny, nz = g.dims
            out.append(flat[g.offset:g.offset + g.ncell].reshape(nz, ny, nx).permute(2, 1, 0)
                       if hasattr(flat, "permute") else
                       flat[g.offset:g.offset + g.ncell].reshape(nz, ny, nx).transpose(2, 1, 0))
        return out

    def mesh_points(self, flat_grad):
        """Per-mesh slices of a (V, 3) point array."""
        return [flat_grad[self.point_off[m]:self.point_off[m + 1]] for m in range(self.n)]


def _make_function():
    import torch

    class SignedDistance(torch.autograd.Function):
        """SDF of one closed mesh on its voxelize grid, differentiable with respect to
        the vertex positions. forward: points (V, 3) float64, triangles (T, 3) -> grid
        (nx, ny, nz) float32 (grid[i, j, k] at origin + (i, j, k) * spacing). backward:
        the closest-feature adjoint; signs are held constant (zero gradient)."""

        @staticmethod
        def forward(ctx, points, tris, resolution, pad):
            b = SdfBatch([(points.detach(), tris)], [resolution], [pad],
                         device=points.device)
            sdf = b.forward(graph=False)
            ctx.batch = b
            ctx.dtype = points.dtype
            nx, ny, nz = b.grids[0].dims
            ctx.dims = (nx, ny, nz)
            return sdf.view(nz, ny, nx).permute(2, 1, 0)

        @staticmethod
        def backward(ctx, grad_out):
            g = grad_out.permute(2, 1, 0).contiguous().view(-1).to(torch.float32)
            grad = ctx.batch.backward(g, graph=False)
            return grad.to(ctx.dtype), None, None, None

    return SignedDistance


_fn = None


def signed_distance(points, tris, resolution: int = 64, pad: int = 1):
    """Differentiable signed-distance grid of one mesh (see SignedDistance)."""
    global _fn
    if _fn is None:
        _fn = _make_function()
    return _fn.apply(points, tris, resolution, pad)
