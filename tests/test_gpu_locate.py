"""Parity of the batched GPU point location (cmg_hip_locate) with the CPU oracle of
src/delaunay/incremental.cpp:321-358: same tet and same walk length for every query,
from a common start and from per-query starts, inside and outside the hull, and through
the step-limit fallback; and of the batched Delaunay cavities (cmg_hip_cavity) with the
oracle's flood (:148-161): same tets in the same discovery order."""
import numpy as np
import pytest

import oracle
from tests.test_oracle_locate import cycling_mesh, lattice_mesh

pytestmark = pytest.mark.gpu


def _run(pts, tets, nbrs, q, start):
    from cybermeshgenerator_amd.locate import locate
    out, steps, brute = locate(pts, tets, nbrs, q, start)
    return out.cpu().numpy(), steps.cpu().numpy(), brute


@pytest.mark.parametrize("jitter", [0.0, 0.3])
def test_walks_match_the_oracle(gpu, jitter):
    pts, tets, nbrs = lattice_mesh(24, jitter=jitter, seed=3)
    rng = np.random.default_rng(4)
    q = np.concatenate([rng.uniform(0, 1, (30000, 3)),
                        rng.uniform(-0.5, 1.5, (2000, 3)),          # outside the hull
                        np.round(rng.uniform(0, 1, (2000, 3)) * 48) / 48])  # on faces
    print(f"locate: {len(tets)} tets, {len(q)} queries, jitter {jitter}", flush=True)
    for start in (0, rng.integers(0, len(tets), len(q)).astype(np.int32)):
        got, steps, brute = _run(pts, tets, nbrs, q, start)
        want, wsteps = oracle.locate(pts, tets, nbrs, q, start)
        assert np.array_equal(got, want)
        assert np.array_equal(steps, wsteps)
        assert brute == int((wsteps < 0).sum())


def test_step_limit_fallback_and_errors(gpu):
    pts, tets, nbrs, q = cycling_mesh()
    got, steps, brute = _run(pts, tets, nbrs, q, 0)
    assert got[0] == 0 and steps[0] == -1 and brute == 1
    pts, tets, nbrs = lattice_mesh(4)
    empty, _, _ = _run(pts, tets, nbrs, np.zeros((0, 3)), 0)
    assert empty.size == 0
    bad = nbrs.copy()
    bad[bad >= 0] = len(tets) + 5            # every neighbour out of range
    with pytest.raises(RuntimeError, match="out of range"):
        _run(pts, tets, bad, np.array([[0.9, 0.9, 0.9]]), 0)
    with pytest.raises(RuntimeError, match="start tet out of range"):
        _run(pts, tets, nbrs, np.array([[0.5, 0.5, 0.5]]), len(tets))


def test_query_count_limit(gpu):
    """nq >= 2^32 is rejected before any query is read (the Morton order carries 32-bit
    query indices); the buffers here are far smaller than nq and are never touched."""
    import ctypes
    import torch
    from cybermeshgenerator_amd import locate
    locate._bind()      # argument types of cmg_hip_locate (size_t counts)
    pts, tets, nbrs = lattice_mesh(2)
    P = torch.tensor(pts, dtype=torch.float64, device="cuda")
    T = torch.tensor(tets, dtype=torch.int32, device="cuda")
    N = torch.tensor(nbrs, dtype=torch.int32, device="cuda")
    q = torch.zeros((4, 3), dtype=torch.float64, device="cuda")
    out = torch.zeros(4, dtype=torch.int32, device="cuda")
    err = ctypes.create_string_buffer(128)
    v = ctypes.c_void_p
    st = gpu.lib().cmg_hip_locate(v(P.data_ptr()), P.shape[0], v(T.data_ptr()),
                                  v(N.data_ptr()), T.shape[0], v(q.data_ptr()), 1 << 32,
                                  None, 0, v(out.data_ptr()), None, None,
                                  v(torch.cuda.current_stream().cuda_stream), err, 128)
    assert st == 1 and b"2^32" in err.value
    assert int(out.abs().sum()) == 0


@pytest.mark.parametrize("lattice", [0, 24])
def test_cavities_match_the_oracle(gpu, lattice):
    """Queries located on the GPU in a Delaunay mesh (uniform points, or lattice points
    with many exactly cospherical tuples and queries on circumspheres, where the compat
    insphere falls back to its exact sign), then their cavities: counts and tets in
    discovery order identical to the oracle's; a small cap marks the larger cavities."""
    from cybermeshgenerator_amd import workloads as W
    from cybermeshgenerator_amd.locate import cavity, locate
    # lattice: 3000 of the 25^3 lattice points, queries on the lattice (about a quarter
    # of the boundary tests have the query exactly on a circumsphere)
    pts, tets, nbrs = W.delaunay_tets(3000 if lattice else 20000, seed=41, lattice=lattice)
    rng = np.random.default_rng(42)
    if lattice:
        q = rng.integers(1, lattice, size=(1 << 14, 3)).astype(np.float64) / lattice
    else:
        q = rng.uniform(0.02, 0.98, (1 << 14, 3))
    print(f"cavity: {len(tets)} tets, {len(q)} queries, lattice {lattice}", flush=True)
    loc, _, _ = locate(pts, tets, nbrs, q, 0)
    loc_np = loc.cpu().numpy()
    assert np.array_equal(loc_np, oracle.locate(pts, tets, nbrs, q, 0)[0])
    count, cav, bad = cavity(pts, tets, nbrs, q, loc, cap=256)
    want_count, want_cav = oracle.cavity(pts, tets, nbrs, q, loc_np, cap=256)
    count = count.cpu().numpy()
    assert bad == 0 and (count > 0).all() and count.max() > 12
    assert np.array_equal(count, want_count)
    assert np.array_equal(cav.cpu().numpy(), want_cav)
    print(f"cavity sizes: mean {count.mean():.1f}, max {count.max()}", flush=True)
    small, _, _ = cavity(pts, tets, nbrs, q, loc, cap=12)
    assert np.array_equal(small.cpu().numpy(), np.where(want_count > 12, -1, want_count))


def test_cavity_edge_cases(gpu):
    """A query located at -1 gets an empty cavity; a neighbour index out of range fails
    that query (count -1) and is counted, without a read through it."""
    from cybermeshgenerator_amd import workloads as W
    from cybermeshgenerator_amd.locate import cavity
    pts, tets, nbrs = W.delaunay_tets(500, seed=43)
    q = np.array([[0.5, 0.5, 0.5], [0.4, 0.6, 0.5]])
    loc = oracle.locate(pts, tets, nbrs, q, 0)[0]
    count, _, bad = cavity(pts, tets, nbrs, q, np.array([-1, loc[1]], np.int32))
    assert count.cpu().numpy()[0] == 0 and count.cpu().numpy()[1] > 0 and bad == 0
    broken = nbrs.copy()
    broken[broken >= 0] = len(tets) + 7
    count, _, bad = cavity(pts, tets, broken, q, loc.astype(np.int32))
    assert (count.cpu().numpy() == -1).all() and bad == 2
