"""Parity of the GPU quality report (cmg_hip_quality_report) with the CPU oracle of
src/quality/metrics.cpp:81-169.

Bit-exact: tet count, min/max volume, min/max radius-edge ratio, the worst-tets list
(ratios and indices; equal ratios by ascending index). Within stated tolerances:
  - sums (total volume, mean ratio, mean dihedral): the GPU sums per workgroup then in
    workgroup order, the reference left to right; both errors are below n * 2^-53 times
    the sum of magnitudes, so |gpu - cpu| <= 2 n 2^-53 sum|x| (tolerance written below);
  - dihedral angles: the device acos and the host's differ by a few ulp; min/max dihedral
    agree to 1e-12 degrees and a histogram bucket may differ only by angles within 1e-12
    degrees of a bucket edge (counted from the oracle's angles)."""
import numpy as np
import pytest

import oracle
from cybermeshgenerator_amd import workloads as W

pytestmark = pytest.mark.gpu

U = 2.0 ** -53


def _gpu(pts, tets, k):
    from cybermeshgenerator_amd import quality
    return quality.report(pts, tets, worst_count=k)


def _compare(pts, tets, k):
    g = _gpu(pts, tets, k)
    c = oracle.quality_report(pts, tets, worst_count=k)
    n = len(tets)
    assert g["num_tets"] == c["num_tets"] == n
    for f in ("min_volume", "max_volume", "min_radius_edge", "max_radius_edge"):
        assert g[f] == c[f], f
    assert g["worst_tets"] == c["worst_tets"]
    vols = abs(c["total_volume"])
    assert abs(g["total_volume"] - c["total_volume"]) <= 2 * n * U * vols
    assert abs(g["mean_radius_edge"] - c["mean_radius_edge"]) <= \
        2 * n * U * abs(c["mean_radius_edge"]) + 1e-300
    assert abs(g["mean_dihedral"] - c["mean_dihedral"]) <= 2 * 6 * n * U * 180
    assert abs(g["min_dihedral"] - c["min_dihedral"]) <= 1e-12
    assert abs(g["max_dihedral"] - c["max_dihedral"]) <= 1e-12
    ang = oracle.dihedrals(pts, tets).ravel()
    edge = np.abs(ang - np.round(ang / 10) * 10) <= 1e-12
    moved = np.abs(np.array(g["dihedral_histogram"]) - np.array(c["dihedral_histogram"]))
    assert moved.sum() <= 2 * int(edge.sum()), (g["dihedral_histogram"],
                                                c["dihedral_histogram"])
    assert sum(g["dihedral_histogram"]) == 6 * n
    return g, c, int(moved.sum())


def test_jittered_lattice_with_degenerate_tets(gpu):
    pts, tets = W.lattice_tets(48, jitter=0.35, seed=1)
    # flat (coplanar) and repeated-vertex tets: sentinel ratios, 0-degree dihedrals
    extra = np.array([[0, 1, 2, 3], [5, 5, 6, 7], [10, 11, 12, 13]], np.int32)
    tets = np.concatenate([tets, extra])
    print(f"quality: {len(tets)} tets", flush=True)
    _compare(pts, tets, 10)
    _compare(pts, tets, 0)


def test_exact_lattice_angles_on_bucket_edges(gpu):
    """Kuhn tets of a dyadic lattice: every angle is 45, 60 or 90 degrees exactly."""
    pts, tets = W.lattice_tets(16)
    g, c, moved = _compare(pts, tets, -1)
    print(f"histogram entries moved across an edge: {moved}", flush=True)


def test_worst_list_selection_path(gpu):
    """Up to 64 worst tets come from a per-workgroup selection: same list as the oracle's
    full order, including all-equal ratios (ascending index) and partial last chunks."""
    pts, tets = W.lattice_tets(24, jitter=0.25, seed=6)
    for k in (1, 17, 64, 65):
        _compare(pts, tets, k)
    pts, tets = W.lattice_tets(16)          # every ratio equal
    g = _gpu(pts, tets, 64)
    assert g["worst_tets"] == list(range(64))
    _compare(pts, tets[:8193], 64)          # one full chunk and a single-tet chunk


def test_all_worst_and_small_meshes(gpu):
    pts, tets = W.lattice_tets(6, jitter=0.2, seed=4)
    _compare(pts, tets, -1)
    _compare(pts, tets[:1], 10)
    _compare(pts, tets[:7], 3)


def test_empty_and_bad_indices(gpu):
    pts, tets = W.lattice_tets(3)
    r = _gpu(pts, np.zeros((0, 4), np.int32), 10)
    assert r["num_tets"] == 0 and r["total_volume"] == 0 and r["worst_tets"] == []
    bad = tets.copy()
    bad[5, 2] = len(pts)
    with pytest.raises(RuntimeError, match="outside the point array"):
        _gpu(pts, bad, 10)
