"""Parity of the HIP voxelization with the CPU oracle: occupancy bit-exact, SDF signs
bit-exact, SDF magnitudes bit-exact within one cell of the surface (FP64 re-evaluation)
and within a stated FP32 tolerance elsewhere, SDF adjoint against the FP64 gradient."""
import numpy as np
import pytest

import oracle
from cybermeshgenerator_amd import workloads as W
from tests.meshes import box_on_cell_centres, brick, subdivided_cube, uv_sphere

pytestmark = pytest.mark.gpu


def gpu_occ(cm, pts, tris, res, pad):
    plc = cm.plc_from_mesh(pts, tris)
    return cm.voxelize(plc, resolution=res, mode="occupancy", pad=pad, flat=True).grid


@pytest.mark.parametrize("snap", [None, 2.0 ** -6, 2.0 ** -7])
def test_occupancy_icosphere_c1(gpu, snap):
    case = W.icosphere_c1(seed=1, snap=snap)
    got = gpu_occ(gpu, case.points, case.tris, 64, 0)
    g = oracle.make_grid(case.points, 64, 0)
    want, stats = oracle.occupancy(case.points, case.tris, g, with_stats=True)
    assert np.array_equal(got, want), int((got != want).sum())
    if snap:
        assert stats[0] > 0   # the nudge path really ran


def test_occupancy_cube_known_answer(gpu):
    pts, tris = subdivided_cube(4)
    got = gpu_occ(gpu, pts, tris, 16, 1).reshape(18, 18, 18)
    want = np.zeros((18, 18, 18), np.uint8)
    want[1:17, 1:17, 1:17] = 1
    assert np.array_equal(got, want)


def test_occupancy_noncubic_axes(gpu):
    for X, Y, Z in [(2.0, 1.0, 1.0), (1.0, 2.0, 1.0), (1.0, 1.0, 2.0), (3.0, 1.0, 2.0)]:
        pts, tris = brick(X, Y, Z)
        got = gpu_occ(gpu, pts, tris, 20, 1)
        g = oracle.make_grid(pts, 20, 1)
        assert np.array_equal(got, oracle.occupancy(pts, tris, g))


def test_occupancy_every_nudge_grazes(gpu):
    """A zero-area (collinear) triangle grazes every attempt of the columns it covers, so
    the reference keeps the crossings of the triangles listed before it
    (voxelize.cpp:163-164). The replay pass must reproduce that exactly."""
    pts, tris = subdivided_cube(4)
    extra = np.array([[0.3, 0.3, 0.5], [0.5, 0.5, 0.5], [0.7, 0.7, 0.5]])
    pts = np.concatenate([pts, extra])
    k = len(pts) - 3
    mid = tris.shape[0] // 2
    tris = np.concatenate([tris[:mid], [[k, k + 1, k + 2]], tris[mid:]]).astype(np.int32)
    g = oracle.make_grid(pts, 16, 1)
    want, stats = oracle.occupancy(pts, tris, g, with_stats=True)
    assert stats[1] > 0      # some columns exhausted every nudge
    got = gpu_occ(gpu, pts, tris, 16, 1)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("which", ["torus", "harmonic", "sphere"])
def test_occupancy_larger_meshes(gpu, which):
    if which == "torus":
        case = W.torus_c3(seed=3)          # 65,536 triangles, 256^3
    elif which == "harmonic":
        case = W.harmonic_c4(seed=4)       # 20,480 triangles, 128^3
    else:
        sp, st = uv_sphere(0.4, 40, 60)
        case = W.MeshCase("uv", W.anchored(sp + 0.5), st, 96, 0)
    got = gpu_occ(gpu, case.points, case.tris, case.resolution, case.pad)
    g = oracle.make_grid(case.points, case.resolution, case.pad)
    want = oracle.occupancy(case.points, case.tris, g)
    assert np.array_equal(got, want), int((got != want).sum())


# ---- signed distance ---------------------------------------------------------------

def check_sdf(cm, pts, tris, res, pad, tol=None):
    """GPU SDF vs the FP64 oracle: signs bit-exact everywhere, values bit-exact within
    one cell of the surface, FP32 error <= tol elsewhere. Returns the far-field error."""
    plc = cm.plc_from_mesh(pts, tris)
    got = cm.voxelize(plc, resolution=res, mode="sdf", pad=pad, flat=True).grid
    g = oracle.make_grid(pts, res, pad)
    occ = oracle.occupancy(pts, tris, g)
    want, best, _ = oracle.sdf_bvh(pts, tris, g, occ)
    assert np.array_equal(np.signbit(got), np.signbit(want))
    near = np.sqrt(best) <= g.spacing
    assert np.array_equal(got[near].view(np.uint32), want[near].view(np.uint32)), \
        int((got[near] != want[near]).sum())
    err = np.abs(got.astype(np.float64) - want.astype(np.float64))
    extent = max(g.dims) * g.spacing
    if tol is None:
        tol = 4e-6 * extent          # FP32 distance from the corner: a few ulps of extent
    assert err.max() <= tol, (err.max(), tol)
    return {"max": float(err.max()), "rms": float(np.sqrt((err ** 2).mean())),
            "near_cells": int(near.sum())}


def test_sdf_uv_sphere_reference_case(gpu):
    """The brute-force case of tests/test_voxelize.cpp:194-275 (16x16 UV sphere, R=20,
    pad=1), bit-exact near the surface."""
    sp, st = uv_sphere(1.0, 16, 16)
    check_sdf(gpu, sp, st, 20, 1)


def test_sdf_cube_signs_and_centre(gpu):
    pts, tris = subdivided_cube(4)
    plc = gpu.plc_from_mesh(pts, tris)
    d = gpu.voxelize(plc, resolution=16, mode="sdf", pad=2).grid
    assert d[10, 10, 10] < 0 and 0.3 < abs(d[10, 10, 10]) < 0.6
    assert d[0, 0, 0] > 0
    check_sdf(gpu, pts, tris, 16, 2)


@pytest.mark.parametrize("snap", [None, 2.0 ** -7])
def test_sdf_icosphere_c1(gpu, snap):
    case = W.icosphere_c1(seed=1, snap=snap)
    check_sdf(gpu, case.points, case.tris, 64, 0)


def test_sdf_harmonic_c4(gpu):
    case = W.harmonic_c4(seed=4)
    r = check_sdf(gpu, case.points, case.tris, 128, 0)
    print("C4 FP32 SDF error vs FP64:", r)


def test_sdf_torus_c3_full(gpu):
    """C3 at full size: 65,536 triangles, 256^3 cells, every cell compared."""
    case = W.torus_c3(seed=3)
    r = check_sdf(gpu, case.points, case.tris, 256, 0)
    print("C3 FP32 SDF error vs FP64:", r)


def _fp32_distance_margin(g):
    """The device's FP32 distance error margin for a grid: 64 u times the grid diagonal
    plus 1e-6 cell widths (u = 2^-24)."""
    return 64.0 * 2.0 ** -24 * max(g.dims) * g.spacing * 3 ** 0.5 + 1e-6 * g.spacing


def test_sdf_gradient_c4(gpu):
    """SDF adjoint at C4 size (20,480 triangles, 128^3 cells, N(0,1) upstream gradient).

    The adjoint follows the FP32 nearest triangle. Where two triangles are equidistant to
    within the FP32 distance error (next to a shared edge, or on the medial axis) that
    choice may differ from the FP64 one, and the cell's contribution lands on the other
    triangle's vertices. So the check has three parts:
      1. the triangle the adjoint uses is a nearest triangle to within twice the FP32
         distance margin (FP64 distances, every cell);
      2. given those triangles, the FP64 CPU adjoint equals the device adjoint to FP32
         accumulation accuracy (rel L2 < 1e-4, max < 1e-3 of the largest entry);
      3. against the FP64 adjoint of the FP64 nearest triangles: rel L2 < 2e-3 and every
         entry within 3e-2 of the largest; FP64 central differences (256 random vertices
         x 3 axes) pin the FP64 adjoint at h = 1e-7 to 2e-6 of the largest entry, plus
         the difference allowed to cells whose nearest triangle changes within the step,
         and are compared with the device adjoint at h = 1e-5 (beyond the tie allowance,
         within 1e-3 of the largest entry)."""
    case = W.harmonic_c4(seed=4)
    pts, tris = case.points, case.tris
    g = oracle.make_grid(pts, 128, 0)
    gout = W.upstream_grad(g.ncell, seed=40)
    plc = gpu.plc_from_mesh(pts, tris)
    sdf, grad = gpu.sdf_with_grad(plc, 128, 0, gout)
    near = gpu.sdf_nearest(plc, 128, 0)
    occ = oracle.occupancy(pts, tris, g)
    _, best, arg = oracle.sdf_bvh(pts, tris, g, occ)
    # 1. a nearest triangle, to within the FP32 margin
    assert (near >= 0).all() and (near < tris.shape[0]).all()
    d_used = np.sqrt(oracle.sdf_dist2_to(pts, tris, g, near))
    gap = d_used - np.sqrt(best)
    assert gap.min() >= 0 and gap.max() <= 2 * _fp32_distance_margin(g), gap.max()
    flips = (near != arg).mean()
    # of the cells on another triangle than the FP64 search's: exact ties (both at the
    # same FP64 distance: a closest point on a shared edge or vertex) and FP32 near-ties
    other = near != arg
    exact_tie = (oracle.sdf_dist2_to(pts, tris, g, near) == best) & other
    tie_split = (exact_tie.sum() / max(other.sum(), 1), (other & ~exact_tie).mean())
    # 2. the scatter itself
    same = oracle.sdf_grad(pts, tris, g, occ, gout, arg=near)
    scale = np.abs(same).max()
    rel_same = np.linalg.norm(grad - same) / np.linalg.norm(same)
    max_same = np.abs(grad - same).max() / scale
    # 3. the FP64 adjoint and finite differences
    ref = oracle.sdf_grad(pts, tris, g, occ, gout)
    rel = np.linalg.norm(grad - ref) / np.linalg.norm(ref)
    max_ref = np.abs(grad - ref).max() / scale
    print(f"C4 gradient: cells on another nearest triangle {flips:.2e} (of them "
          f"{tie_split[0]:.2%} exact FP64 ties; FP32 near-ties on {tie_split[1]:.2e} of all "
          f"cells; max distance gap "
          f"{gap.max():.2e}); vs FP64 adjoint of the same triangles rel L2 {rel_same:.2e}, "
          f"max {max_same:.2e}; vs FP64 adjoint rel L2 {rel:.2e}, max {max_ref:.2e}")
    assert np.abs(grad[-2:]).max() == 0            # anchors are not referenced
    assert rel_same < 1e-4 and max_same < 1e-3
    assert rel < 2e-3 and max_ref < 3e-2
    rng = np.random.default_rng(41)
    nv = pts.shape[0] - 2
    fd_err, tied = [], 0
    for v in rng.choice(nv, 256, replace=False):
        for axis in range(3):
            # ties: cells whose nearest triangle changes between -h, 0 and +h, where the
            # difference quotient mixes several triangles' derivatives. A cell whose
            # closest point crosses a face/edge/vertex boundary inside the step leaves an
            # O(h) difference, so the FP64 adjoint is pinned at h = 1e-7 (where that is
            # below 2e-6 of the largest entry); the device adjoint is compared at 1e-5.
            fd7, ties7 = oracle.sdf_fd(pts, tris, g, occ, gout, int(v), axis, h=1e-7,
                                       base=(best, arg), with_ties=True)
            assert abs(fd7 - ref[v, axis]) <= 2e-6 * scale + 1e-6 * abs(fd7) + ties7, (v, axis)
            fd, ties = oracle.sdf_fd(pts, tris, g, occ, gout, int(v), axis, h=1e-5,
                                     base=(best, arg), with_ties=True)
            tied += ties > 1e-9 * scale
            fd_err.append(max(0.0, abs(fd - grad[v, axis]) - ties) / scale)
    fd_err = np.array(fd_err)
    print(f"C4 gradient vs FD (beyond the tie allowance): max |err| / max|g| "
          f"{fd_err.max():.3e}, median {np.median(fd_err):.3e}; {tied} of 768 with ties")
    assert fd_err.max() < 1e-3 and np.median(fd_err) < 1e-4


def test_adjoint_with_cells_on_the_surface(gpu):
    """Cell centres exactly on faces, edges and corners: those cells carry no gradient,
    the others match the FP64 adjoint of the same triangles."""
    pts, tris = box_on_cell_centres()
    g = oracle.make_grid(pts, 4, 0)
    gout = W.upstream_grad(g.ncell, seed=12)
    plc = gpu.plc_from_mesh(pts, tris)
    _, grad = gpu.sdf_with_grad(plc, 4, 0, gout)
    near = gpu.sdf_nearest(plc, 4, 0)
    occ = oracle.occupancy(pts, tris, g)
    ref = oracle.sdf_grad(pts, tris, g, occ, gout, arg=near)
    assert np.abs(grad - ref).max() <= 1e-5 * np.abs(ref).max()



def _polygon_solids():
    """A unit cube with 6 quad facets, and a rotated hexagonal prism with hexagon caps and
    quad sides (facets of 4 and 6 vertices, outward order)."""
    cube = np.array([[x, y, z] for z in (0.0, 1.0) for y in (0.0, 1.0) for x in (0.0, 1.0)])
    cube_f = [[0, 2, 3, 1], [4, 5, 7, 6], [0, 1, 5, 4], [2, 6, 7, 3], [0, 4, 6, 2],
              [1, 3, 7, 5]]
    ang = np.arange(6) * np.pi / 3
    ring = np.stack([0.4 * np.cos(ang), 0.4 * np.sin(ang)], 1)
    hexp = np.concatenate([np.c_[ring, np.full(6, -0.3)], np.c_[ring, np.full(6, 0.3)]])
    hexp = hexp @ W.random_rotation(np.random.default_rng(8)).T + 0.5
    hex_f = [[5, 4, 3, 2, 1, 0], [6, 7, 8, 9, 10, 11]]
    hex_f += [[k, (k + 1) % 6, 6 + (k + 1) % 6, 6 + k] for k in range(6)]
    return [(cube, cube_f, 16, 1), (hexp, hex_f, 24, 2)]


@pytest.mark.parametrize("which", [0, 1])
def test_polygon_facets_fan_triangulated(gpu, which):
    """voxelize.cpp:29-38: facets of 4 and more vertices are fan-triangulated
    (v0, v[i-1], v[i]) before voxelization; through cmg_plc_add_facet the GPU path gives
    the oracle's triangles, occupancy and SDF."""
    pts, faces, res, pad = _polygon_solids()[which]
    plc = gpu.PLC().add_points(pts)
    for f in faces:
        plc.add_facet(f)
    tris = oracle.fan_triangulate([np.array(f, np.int32) for f in faces])
    assert tris.shape[0] == sum(len(f) - 2 for f in faces)
    from cybermeshgenerator_amd.files import plc_geometry
    lib_pts, lib_tris = plc_geometry(plc._handle)     # cmg_plc_points / cmg_plc_triangles
    assert np.array_equal(lib_pts, pts) and np.array_equal(lib_tris, tris)
    g = oracle.make_grid(pts, res, pad)
    occ = oracle.occupancy(pts, tris, g)
    got = gpu.voxelize(plc, resolution=res, mode="occupancy", pad=pad, flat=True).grid
    assert np.array_equal(got, occ)
    if which == 0:      # the cube's known answer (test_voxelize.cpp:140-163)
        want = np.zeros((18, 18, 18), np.uint8)
        want[1:17, 1:17, 1:17] = 1
        assert np.array_equal(got.reshape(18, 18, 18), want)
    sdf = gpu.voxelize(plc, resolution=res, mode="sdf", pad=pad, flat=True).grid
    ref, best, _ = oracle.sdf_bvh(pts, tris, g, occ)
    assert np.array_equal(np.signbit(sdf), occ == 1)
    near = np.sqrt(best) <= g.spacing
    assert np.array_equal(sdf[near], ref[near])
    assert np.abs(sdf.astype(np.float64) - ref).max() <= 4e-6 * max(g.dims) * g.spacing
