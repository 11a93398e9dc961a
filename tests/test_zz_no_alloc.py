"""The no-allocation orient3d entry point (cmg_hip_orient3d_signs_soa_noalloc), called
through the C ABI: its signs equal those of cmg_hip_orient3d_signs_soa on the same
batch, on two calls in a row (the second smaller, in the other mode, so stale workspace
contents of the first call must not leak into it), and sizes beyond its fixed workspace
are refused."""
import ctypes

import numpy as np
import pytest

from cybermeshgenerator_amd import workloads as W

pytestmark = pytest.mark.gpu

_P = ctypes.c_void_p


def _bind(lib):
    f = lib.cmg_hip_orient3d_signs_soa_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [_P, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, _P, _P,
                  ctypes.c_char_p, ctypes.c_size_t]
    g = lib.cmg_hip_orient3d_signs_soa
    g.restype = ctypes.c_int
    g.argtypes = [_P, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, _P, _P,
                  ctypes.c_size_t, _P, ctypes.c_char_p, ctypes.c_size_t]
    return f, g


def test_noalloc_matches_the_allocating_entry_point(gpu):
    import torch
    lib = gpu.lib()
    noalloc, soa_signs = _bind(lib)
    stream = torch.cuda.current_stream()
    st = _P(stream.cuda_stream)
    err = ctypes.create_string_buffer(256)
    for n, mode, seed in ((1 << 20, 0, 81), (12345, 1, 82), (1 << 20, 1, 83), (3, 0, 84)):
        print(f"noalloc: n={n} mode={mode} seed={seed}", flush=True)
        soa_np, _ = W.tuples_orient3d(n, seed=seed)
        soa = torch.from_numpy(soa_np).cuda()
        got = torch.full((n,), 9, dtype=torch.int8, device="cuda")
        want = torch.full((n,), 7, dtype=torch.int8, device="cuda")
        rc = noalloc(_P(soa.data_ptr()), n, n, mode, _P(got.data_ptr()), st, err, 256)
        assert rc == 0, err.value
        ws = torch.empty(gpu.pred_workspace_size(n, 4), dtype=torch.uint8, device="cuda")
        rc = soa_signs(_P(soa.data_ptr()), n, n, mode, _P(want.data_ptr()),
                       _P(ws.data_ptr()), ws.numel(), st, err, 256)
        assert rc == 0, err.value
        torch.cuda.synchronize()
        assert torch.equal(got, want), (n, mode, int((got != want).sum()))
    # beyond the fixed workspace: refused before any launch, output untouched
    big = (1 << 20) + 1
    out = torch.full((4,), 5, dtype=torch.int8, device="cuda")
    rc = noalloc(_P(soa.data_ptr()), big, big, 0, _P(out.data_ptr()), st, err, 256)
    assert rc == 1 and b"2^20" in err.value
    assert int(out.sum()) == 20
    # n == 0 is a no-op
    assert noalloc(None, 0, 0, 0, None, st, err, 256) == 0
    torch.cuda.synchronize()
