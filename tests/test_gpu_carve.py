"""Parity of the HIP PLC-carve classifier with the CPU oracle (oracle/carve.c): the rays
mode (the reference's CarveGrid, tetrahedralize_plc.cpp:60-174) and the parity mode (the
voxelizer's column parity at arbitrary points) are bit-exact, for points and for tet
centroids, through the device API (cmg_carve_*) and the host drop-in (cmg_plc_carve)."""
import numpy as np
import pytest

import oracle
from cybermeshgenerator_amd import workloads as W
from tests.meshes import subdivided_cube
from tests.test_oracle_carve import _l_prism, _near_and_far_points, cell_centres, grazing_cube

pytestmark = pytest.mark.gpu


def _cases():
    c1 = W.icosphere_c1(seed=1)
    c1s = W.icosphere_c1(seed=1, snap=2.0 ** -6)
    c3 = W.torus_c3(seed=3)
    c4 = W.harmonic_c4(seed=4)
    cube = subdivided_cube(4)
    return {"c1": (c1.points, c1.tris), "c1_snapped": (c1s.points, c1s.tris),
            "torus_c3": (c3.points, c3.tris), "harmonic_c4": (c4.points, c4.tris),
            "cube": cube, "l_prism": _l_prism(), "grazing_cube": grazing_cube()}


CASES = _cases()


def _grid(pts, tris, mode, res=0):
    import torch
    from cybermeshgenerator_amd.carve import CarveGrid
    P = torch.tensor(np.asarray(pts, np.float64), device="cuda")
    T = torch.tensor(np.asarray(tris, np.int32), device="cuda")
    return CarveGrid(P, T, mode=mode, resolution=res)


def _queries(pts, tris, seed, n):
    rng = np.random.default_rng(seed)
    q = _near_and_far_points(np.asarray(pts, np.float64), np.asarray(tris), rng, n)
    # points on a 2^-6 lattice (exact ties with snapped vertices and edges)
    lo, hi = np.min(pts, 0), np.max(pts, 0)
    lat = np.round(rng.uniform(lo, hi, (n // 4, 3)) * 64) / 64
    return np.concatenate([q, lat, pts[: n // 4]])


@pytest.mark.parametrize("name", list(CASES))
def test_rays_points_bit_exact(gpu, name):
    import torch
    pts, tris = CASES[name]
    print(f"carve rays points: {name} tris={len(tris)}", flush=True)
    q = _queries(pts, tris, 1, 20000)
    cg = _grid(pts, tris, "rays")
    got = cg.points_inside(torch.tensor(q, device="cuda")).cpu().numpy()
    want = oracle.Carve(pts, tris).points(q)
    assert np.array_equal(got, want), int((got != want).sum())
    assert 0 < want.sum() < len(q)


@pytest.mark.parametrize("name", list(CASES))
def test_parity_points_bit_exact(gpu, name):
    import torch
    pts, tris = CASES[name]
    print(f"carve parity points: {name} tris={len(tris)}", flush=True)
    q = _queries(pts, tris, 2, 20000)
    cg = _grid(pts, tris, "parity")
    got = cg.points_inside(torch.tensor(q, device="cuda")).cpu().numpy()
    r = oracle.carve_resolution(len(tris))
    g = oracle.make_grid(pts, r, 1)
    want = oracle.carve_parity_points(pts, tris, g, q)
    assert np.array_equal(got, want), int((got != want).sum())


@pytest.mark.parametrize("name", ["c1_snapped", "grazing_cube", "cube"])
def test_parity_at_cell_centres_is_the_occupancy(gpu, name):
    """At cell centres the parity classifier is the voxelizer (voxelize.cpp:169-195),
    nudges and the every-nudge-grazes partial count included."""
    import torch
    pts, tris = CASES[name]
    res = 16
    g = oracle.make_grid(pts, res, 1)
    q = cell_centres(g)
    cg = _grid(pts, tris, "parity", res)
    got = cg.points_inside(torch.tensor(q, device="cuda")).cpu().numpy()
    assert np.array_equal(got, oracle.occupancy(pts, tris, g))


@pytest.mark.parametrize("mode", ["rays", "parity"])
def test_tets_bit_exact_and_bad_indices(gpu, mode):
    import torch
    pts, tris = CASES["harmonic_c4"]
    rng = np.random.default_rng(7)
    allp = np.concatenate([pts, rng.uniform(0, 1, (5000, 3))])
    tets = rng.integers(0, len(allp), (50000, 4)).astype(np.int32)
    tets[::997, 2] = len(allp)       # out of range: rejected, flagged, 0
    tets[::1009, 0] = -1
    bad = (tets < 0).any(1) | (tets >= len(allp)).any(1)
    cg = _grid(allp, tris, mode)
    got = cg.tets_inside(torch.tensor(tets, device="cuda")).cpu().numpy()
    assert cg.bad_tets() == int(bad.sum())
    assert cg.bad_tets() == 0
    assert not got[bad].any()
    good = tets[~bad]
    if mode == "rays":
        want = oracle.Carve(allp, tris).tets(good)
    else:
        cen = np.array([oracle.tet_centroid(allp, t) for t in good])
        g = oracle.make_grid(allp, oracle.carve_resolution(len(tris)), 1)
        want = oracle.carve_parity_points(allp, tris, g, cen)
    assert np.array_equal(got[~bad], want)


@pytest.mark.parametrize("mode", ["rays", "parity"])
def test_bad_triangle_indices_rejected_at_setup(gpu, mode):
    """Device triangles with a vertex index outside the point array are counted on the
    device and refused before any kernel gathers points through them."""
    pts, tris = CASES["harmonic_c4"]
    for bad in (len(pts), -1):
        t = np.array(tris, np.int32)
        t[5, 1] = bad
        t[17, 2] = bad
        with pytest.raises(RuntimeError, match="2 triangle vertex indices outside"):
            _grid(pts, t, mode)
    _grid(pts, tris, mode)   # the same triangles with their indices in range


def test_plc_carve_host_drop_in(gpu):
    from cybermeshgenerator_amd.carve import plc_carve
    pts, tris = _l_prism()
    rng = np.random.default_rng(3)
    allp = np.concatenate([pts, rng.uniform(-0.2, 2.2, (400, 3))])
    # the PLC holds only the prism's facets; extra points are interior Steiner-like points
    plc = gpu.plc_from_mesh(allp, tris)
    tets = rng.integers(0, len(allp), (4000, 4))
    for mode in ("rays", "parity"):
        keep = plc_carve(plc, tets, mode)
        if mode == "rays":
            want = oracle.Carve(allp, tris).tets(tets)
        else:
            cen = np.array([oracle.tet_centroid(allp, t) for t in tets])
            g = oracle.make_grid(allp, oracle.carve_resolution(len(tris)), 1)
            want = oracle.carve_parity_points(allp, tris, g, cen)
        assert np.array_equal(keep, want), mode
    with pytest.raises(RuntimeError, match="out of range"):
        plc_carve(plc, [[0, 1, 2, len(allp)]])
    empty = gpu.PLC().add_points(allp)
    with pytest.raises(RuntimeError, match="no boundary triangles"):
        plc_carve(empty, tets)


def test_empty_query_and_full_size(gpu):
    """No queries is a no-op; 2^20 points on the C3 torus (r = 256 grids) stay bit-exact."""
    import torch
    pts, tris = CASES["torus_c3"]
    cg = _grid(pts, tris, "rays")
    out = cg.points_inside(torch.zeros((0, 3), dtype=torch.float64, device="cuda"))
    assert out.numel() == 0
    rng = np.random.default_rng(11)
    q = rng.uniform(0.1, 0.9, (1 << 20, 3))
    got = cg.points_inside(torch.tensor(q, device="cuda")).cpu().numpy()
    want = oracle.Carve(pts, tris).points(q)
    assert np.array_equal(got, want)


def _cylinder(n=2000, radius=0.5, height=0.05):
    """A closed cylinder whose two caps are fan-triangulated from one rim vertex: every
    cap triangle spans from that vertex to a far rim edge."""
    ang = 2 * np.pi * np.arange(n) / n
    ring = np.stack([radius * np.cos(ang), radius * np.sin(ang)], 1)
    pts = np.concatenate([np.c_[ring, np.zeros(n)], np.c_[ring, np.full(n, height)]])
    tris = []
    for i in range(2, n):
        tris.append((0, i, i - 1))                  # bottom cap, facing -z
        tris.append((n, n + i - 1, n + i))          # top cap, facing +z
    for i in range(n):
        j = (i + 1) % n
        tris += [(i, j, n + j), (i, n + j, n + i)]
    return pts, np.array(tris, np.int32)


def test_cell_lists_beyond_2_32_entries_rejected(gpu):
    """The parity grid's cell-list total is summed over the triangles' rectangles in 64
    bits and refused at 2^31 or more before any list is counted, scanned or filled: here
    the caps' 3,996 fan triangles cover more than 2^32 (column, triangle) entries on a
    4090 x 4090 grid, a total a 32-bit sum would wrap."""
    pts, tris = _cylinder()
    res = 4090
    g = oracle.make_grid(pts, res, 1)
    assert g.dims[0] * g.dims[1] <= 1 << 24
    cx = np.floor((pts[:, 0] - g.corner[0]) / g.spacing).astype(np.int64)
    cy = np.floor((pts[:, 1] - g.corner[1]) / g.spacing).astype(np.int64)
    w = cx[tris].max(1) - cx[tris].min(1) + 1
    h = cy[tris].max(1) - cy[tris].min(1) + 1
    assert int((w * h).sum()) > 1 << 32
    with pytest.raises(RuntimeError, match="exceed 2\\^31 entries"):
        _grid(pts, tris, "parity", res)
    _grid(pts, tris, "parity", 64)     # the same surface on a coarser grid
