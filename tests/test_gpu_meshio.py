"""Parity of the GPU STL path with the CPU oracle: the binary decode and the merge of
coincident vertices are bit-exact (points, numbering, triangles) through the host drop-in
(cmg_read_plc) and the device path (cmg_hip_stl_decode, cmg_hip_weld)."""
import numpy as np
import pytest

import oracle
from tests.stlfiles import shared_surface, write_ascii_stl, write_binary_stl
from tests.test_oracle_meshio import same_bits, tetra_soup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 255, 256, 257, 3001, 1 << 16])
def test_read_plc_binary_stl(gpu, tmp_path, n):
    from cybermeshgenerator_amd import files
    print(f"stl binary: {n} triangles", flush=True)
    p = str(tmp_path / "b.stl")
    write_binary_stl(p, shared_surface(n, seed=n))
    plc = files.read_plc(p)
    pts, tris, binary = oracle.read_stl(p)
    assert binary
    assert same_bits(plc.points, pts)
    assert np.array_equal(plc.triangles, tris)


def test_read_plc_ascii_stl(gpu, tmp_path):
    from cybermeshgenerator_amd import files
    p = str(tmp_path / "a.stl")
    soup = shared_surface(4000, seed=5).astype(np.float64)
    soup[::7] += 1e-9  # values that need all 17 digits
    write_ascii_stl(p, soup)
    plc = files.read_plc(p)
    pts, tris, binary = oracle.read_stl(p)
    assert not binary
    assert same_bits(plc.points, pts) and np.array_equal(plc.triangles, tris)
    # the reference's own cases (tests/io/test_stl.cpp:35-109)
    P, F, tet = tetra_soup()
    write_ascii_stl(p, tet, fmt="{:g}")
    plc = files.read_plc(p)
    assert len(plc.points) == 4 and len(plc.triangles) == 4
    open(p, "w").write("solid bad\n facet normal 0 0 1\n outer loop\n vertex 0 0\n")
    with pytest.raises(RuntimeError, match="short vertex record"):
        files.read_plc(p)
    write_binary_stl(p, np.zeros((0, 3, 3), np.float32))
    with pytest.raises(RuntimeError, match="positive multiple of 3"):
        files.read_plc(p)


def test_device_decode_and_weld(gpu, tmp_path):
    """A 2^20-triangle binary STL decoded and merged on the device, then fed to the
    voxelizer's inputs unchanged: bit-exact against the oracle's reader."""
    import torch
    from cybermeshgenerator_amd import files
    n = 1 << 20
    p = str(tmp_path / "big.stl")
    write_binary_stl(p, shared_surface(n, seed=9))
    xyz, tris = files.read_stl_device(p)
    pts, otris, _ = oracle.read_stl(p)
    assert same_bits(xyz.cpu().numpy(), pts)
    assert np.array_equal(tris.cpu().numpy(), otris)
    # the merge alone on a device soup (with -0.0 / +0.0 pairs and repeats)
    soup = shared_surface(50000, seed=3).astype(np.float64).reshape(-1, 3)
    wp, wi = files.weld(torch.tensor(soup, device="cuda"))
    op, oi = oracle.weld(soup)
    assert same_bits(wp.cpu().numpy(), op) and np.array_equal(wi.cpu().numpy(), oi)
    # not a binary STL: rejected before any kernel
    raw = torch.zeros(100, dtype=torch.uint8, device="cuda")
    with pytest.raises(RuntimeError, match="not a binary STL"):
        files.stl_decode(raw)


def test_weld_nan_vertices_never_merge(gpu):
    """A vertex with a NaN coordinate is equal to no vertex, itself included, so each one
    becomes a point of its own; the other vertices merge by exact value (-0 == +0) in
    order of first occurrence. Expected numbering from a dict of first occurrences."""
    import math
    import torch
    from cybermeshgenerator_amd import files
    soup = shared_surface(3000, seed=11).astype(np.float64).reshape(-1, 3)
    soup[::97, 1] = np.nan                      # NaN vertices, some at repeated positions
    soup[5::211] = soup[4::211][: len(soup[5::211])]  # repeats next to them
    soup[7::131, 0] = -0.0
    soup[8::131, 0] = 0.0
    first, pts, idx = {}, [], []
    for v in soup:
        if any(math.isnan(c) for c in v):
            idx.append(len(pts))
            pts.append(v)
            continue
        key = tuple(0.0 if c == 0 else c for c in v)
        if key not in first:
            first[key] = len(pts)
            pts.append(v)
        idx.append(first[key])
    wp, wi = files.weld(torch.tensor(soup, device="cuda"))
    assert np.array_equal(wi.cpu().numpy().ravel(), np.array(idx))
    assert same_bits(wp.cpu().numpy(), np.array(pts))
    assert int(np.isnan(wp.cpu().numpy()).any(1).sum()) == int(np.isnan(soup).any(1).sum())


def test_bunny_published_occupancy(gpu, tmp_path):
    """The reference's published bunny occupancy (examples/bunny/README.md:70-72: 24,667
    cells, volume 280,129 at resolution 48, pad 1) through the drop-in path: cmg_read_plc
    (binary STL decoded and merged on the GPU), then cmg_plc_voxelize on the GPU,
    bit-identical to the oracle's grid."""
    from cybermeshgenerator_amd import files
    from tests.stlfiles import BUNNY_RES48_CELLS, BUNNY_RES48_VOLUME, bunny_stl
    path = bunny_stl(tmp_path)
    plc = files.read_plc(path)
    pts, tris, _ = oracle.read_stl(path)
    assert same_bits(plc.points, pts) and np.array_equal(plc.triangles, tris)
    vg = gpu.voxelize(plc, resolution=48, mode="occupancy", pad=1, flat=True)
    g = oracle.make_grid(pts, 48, 1)
    want = oracle.occupancy(pts, tris, g)
    assert np.array_equal(vg.grid, want)
    assert int(vg.grid.sum()) == BUNNY_RES48_CELLS
    assert round(float(vg.grid.sum()) * vg.spacing ** 3) == BUNNY_RES48_VOLUME
    # the reference's (nx, ny, nz) view of the same cells
    v3 = gpu.voxelize(plc, resolution=48, mode="occupancy")
    assert v3.grid.shape == (50, 42, 50) and int(v3.grid.sum()) == BUNNY_RES48_CELLS
