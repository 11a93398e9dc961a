"""The register-only exact stages of the GPU predicate passes. insphere (expansion.hpp: fixed point with
31-, 62- and 126-bit differences, a double-word determinant with a rigorous error bound)
built for the host and checked against the oracle's exact insphere: every tuple a stage
decides must get the oracle's sign, and together they must decide the exactly
cospherical and 1-ulp workload tuples."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from cybermeshgenerator_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "native", "register_stages")


def _run(tuples, tmp_path):
    if not os.path.exists(BIN):
        pytest.skip("tests/native/register_stages not built (run the package build)")
    t = np.ascontiguousarray(tuples, dtype=np.float64).reshape(-1, 15)
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    t.tofile(src)
    subprocess.run([BIN, "insphere", str(src), str(dst)], check=True)
    out = np.fromfile(dst, dtype=np.int8).reshape(-1, 4, 2).astype(np.int64)
    ok, sign = out[:, :, 0] > 0, out[:, :, 1]
    # slots: fixed point with 62-bit differences, the double-word bound, fixed point with
    # 126-bit differences, fixed point with 31-bit differences; `stage` reports the first
    # two in that order (1, 2)
    stage = np.where(ok[:, 0], 1, np.where(ok[:, 1], 2, 0))
    first = np.where(ok[:, 0], sign[:, 0], sign[:, 1])
    _run.last = (ok, sign)
    return stage, first


def _oracle(tuples):
    soa = np.ascontiguousarray(np.asarray(tuples, dtype=np.float64).reshape(-1, 15).T)
    return oracle.insphere_batch(soa, oracle.EXACT).astype(np.int64)


def _check(tuples, tmp_path):
    stage, sign = _run(tuples, tmp_path)
    ok, each = _run.last
    want = _oracle(tuples)
    for k in range(4):  # every stage on its own agrees wherever it decides
        assert np.array_equal(each[ok[:, k], k], want[ok[:, k]]), k
    return stage, want


def test_workload_degenerate_tuples(tmp_path):
    soa, kind = W.tuples_insphere(20000, seed=77)
    tuples = soa.T.copy()
    stage, want = _check(tuples, tmp_path)
    assert (stage[kind == 1] == 1).all()        # exactly cospherical: fixed point, all 0
    assert (want[kind == 1] == 0).all()
    assert (stage[kind == 2] > 0).all()         # one coordinate 1 ulp off
    assert (stage[kind == 2] == 1).mean() > 0.9


def test_inline_order_on_the_workload(tmp_path):
    """The order the streaming filter drains its ring in (insphere_sign_inline): the
    31-bit fixed point decides every exactly cospherical workload tuple (differences of
    about 24 bits on the 2^-24 grid), the double-word bound about 95 % of the 1-ulp ones
    (a moved coordinate adds its ulp to the spread, beyond 31 bits), and the 62-bit fixed
    point the rest."""
    soa, kind = W.tuples_insphere(20000, seed=78)
    tuples = soa.T.copy()
    _check(tuples, tmp_path)
    ok, sign = _run.last
    assert ok[kind == 1, 3].all() and (sign[kind == 1, 3] == 0).all()
    ulp = kind == 2
    assert (ok[ulp, 3] | ok[ulp, 1] | ok[ulp, 0]).all()
    assert ok[ulp, 1].mean() > 0.9


def test_narrow_fixed_point_limit_and_extremes(tmp_path):
    """Differences spanning 31 bits are decided by the 31-bit stage and 32 bits are not;
    differences at the extremes of the range (+-(2^31 - 1) times a power of two) keep
    every intermediate in range: the sign equals the oracle's."""
    base = np.array([1.0, 0.25, 0.5, -0.5, 0.125, 0.75, 0.375, -0.25, 0.0, 0.0, 0.0, 0.0,
                     0.0, 0.0, 0.0])
    t31 = base.copy()
    t31[8] = 2.0 ** -30
    t32 = base.copy()
    t32[8] = 2.0 ** -31
    _run(np.stack([t31, t32]), tmp_path)
    ok, _ = _run.last
    assert ok[0, 3] and not ok[1, 3]
    rng = np.random.default_rng(31)
    top = float(2 ** 31 - 1)
    rows = []
    for scale in (-100, -31, 0, 20, 90):
        # coordinates drawn from {+-top, +-(top - small), small}; e = 0 so p - e = p
        q = rng.choice([top, -top, top - 1, -(top - 7), 5.0, -3.0, 0.0], size=(200, 15))
        q[:, 12:] = 0.0
        rows.append(np.ldexp(q, scale))
        # exactly cospherical at the extremes: e on the sphere too, four more points
        c = np.zeros(3)
        r = top / 2
        pts = np.array([[r, 0, 0], [-r, 0, 0], [0, r, 0], [0, -r, 0], [0, 0, r]])
        rows.append(np.ldexp(np.tile(pts.reshape(1, 15), (4, 1)) + np.tile(c, 5), scale))
    tuples = np.concatenate(rows)
    _check(tuples, tmp_path)
    ok, sign = _run.last
    assert ok[:, 3].mean() > 0.95
    want = _oracle(tuples)
    assert (want[ok[:, 3]] == 0).any() and (want[ok[:, 3]] > 0).any() and (want[ok[:, 3]] < 0).any()


def test_integer_grids_scaled_by_powers_of_two(tmp_path):
    rng = np.random.default_rng(5)
    vecs = W._norm_vectors(81)
    rows = []
    for bits in (3, 20, 40, 55, 61):
        for scale in (-140, -60, 0, 37, 85):
            # random integer tuples (nonzero determinants, both signs)
            q = rng.integers(-(2 ** bits), 2 ** bits, size=(40, 15)).astype(np.float64)
            rows.append(np.ldexp(q, scale))
            # exactly cospherical: integer centre + distinct vectors of one norm times m
            c = rng.integers(-(2 ** min(bits, 40)), 2 ** min(bits, 40), size=(40, 1, 3))
            m = rng.integers(1, 2 ** max(1, min(bits, 48) - 5), size=(40, 1, 1))
            pick = np.argsort(rng.random((40, len(vecs))), axis=1)[:, :5]
            pts = (c + vecs[pick] * m).astype(np.float64)
            rows.append(np.ldexp(pts.reshape(40, 15), scale))
    tuples = np.concatenate(rows)
    stage, want = _check(tuples, tmp_path)
    dec = stage == 1
    assert dec.mean() > 0.95
    assert (want[dec] == 0).any() and (want[dec] > 0).any() and (want[dec] < 0).any()


def test_spread_limit(tmp_path):
    # differences spanning 62 bits are decided, 63 bits are left to the expansions
    base = np.array([1.0, 0.25, 0.5, -0.5, 0.125, 0.75, 0.375, -0.25, 0.0, 0.0, 0.0, 0.0,
                     0.0, 0.0, 0.0])
    t62 = base.copy()
    t62[8] = 2.0 ** -61
    t63 = base.copy()
    t63[8] = 2.0 ** -62
    stage, sign = _run(np.stack([t62, t63]), tmp_path)
    assert stage[0] == 1 and stage[1] != 1
    assert np.array_equal(sign[stage > 0], _oracle(np.stack([t62, t63]))[stage > 0])


def test_uniform_tuples_decided_ones_agree(tmp_path):
    rng = np.random.default_rng(9)
    tuples = rng.uniform(-1, 1, size=(5000, 15))
    # a quarter on a coarse grid (exact differences)
    tuples[:1250] = np.round(tuples[:1250] * 2 ** 20) / 2 ** 20
    stage, _ = _check(tuples, tmp_path)
    assert (stage[:1250] == 1).all()


def test_extreme_magnitudes_left_to_expansions(tmp_path):
    # outside [2^-150, 2^150] the stage declines (products of five differences would
    # leave the normal range, where the expansion arithmetic defines the reference sign)
    rng = np.random.default_rng(3)
    q = rng.integers(-1000, 1000, size=(8, 15)).astype(np.float64)
    stage, _ = _run(np.concatenate([np.ldexp(q, -1000), np.ldexp(q, 900)]), tmp_path)
    assert not stage.any()


def test_double_word_stage_on_widely_spread_tuples(tmp_path):
    # tuples of mixed magnitudes (differences often span more than 62 bits) and the
    # workload's 1-ulp tuples whose moved coordinate is small (so its ulp is tiny and the
    # differences span more than 62 bits): what the fixed-point stage declines, the
    # double-word bound decides, with the oracle's sign
    rng = np.random.default_rng(12)
    uni = rng.uniform(-1, 1, size=(4000, 15)) * 2.0 ** rng.integers(-40, 1, size=(4000, 15))
    soa, kind = W.tuples_insphere(200000, seed=31)
    ulp = soa.T[kind == 2]
    stage, want = _check(np.concatenate([uni, ulp]), tmp_path)
    n = len(uni)
    assert (stage[:n] > 0).all() and (stage[:n] == 2).mean() > 0.2
    wide = stage[n:] != 1
    assert wide.sum() > 0                      # a few in 10^4
    assert (stage[n:][wide] == 2).all()


def test_wide_fixed_point_stage(tmp_path):
    # differences spanning 63..126 bits: the narrow stage declines, the wide one decides
    # (exactly cospherical integer tuples scaled apart by 2^k and then shifted by a tiny
    # common offset on one vertex pair would not stay cospherical, so this uses random
    # integer tuples of 64..120-bit spread, including ones with repeated points: det 0)
    rng = np.random.default_rng(21)
    rows = []
    for spread in (64, 80, 100, 120):
        big = rng.integers(-(2 ** 40), 2 ** 40, size=(60, 15)).astype(np.float64)
        big = np.ldexp(big, spread - 41)          # top bits near 2^spread
        small = rng.integers(1, 2 ** 20, size=(60, 15)).astype(np.float64)
        t = big + np.where(rng.random((60, 15)) < 0.3, small, 0.0)
        t[:10, 3:6] = t[:10, 0:3]                 # a repeated point: det exactly 0
        rows.append(t)
    tuples = np.concatenate(rows)
    stage, want = _check(tuples, tmp_path)
    ok, _ = _run.last
    assert ok[:, 2].mean() > 0.95
    assert (want[ok[:, 2]] == 0).any()


def _run_o3(tuples, tmp_path):
    if not os.path.exists(BIN):
        pytest.skip("tests/native/register_stages not built (run the package build)")
    t = np.ascontiguousarray(tuples, dtype=np.float64).reshape(-1, 12)
    src, dst = tmp_path / "o3.bin", tmp_path / "o3out.bin"
    t.tofile(src)
    subprocess.run([BIN, "orient3d", str(src), str(dst)], check=True)
    out = np.fromfile(dst, dtype=np.int8).reshape(-1, 4, 2).astype(np.int64)
    return out[:, :, 0] > 0, out[:, :, 1]


def test_orient3d_stages_against_oracle(tmp_path):
    # orient3d: fixed point with 31-bit differences and the double-word bound (inside the
    # streaming kernel) and fixed point with 62- and 126-bit differences (exact pass); each
    # agrees with the oracle wherever it decides, and together they decide every workload
    # tuple, including the exact zeros with inexact differences (a point moved by 1 ulp
    # onto another). The 31-bit stage decides every exactly coplanar workload tuple (their
    # differences lie on the 2^-28 grid below 2^3), the double-word bound every 1-ulp one
    # with a nonzero determinant.
    soa, kind = W.tuples_orient3d(1 << 18, seed=2)
    t = soa.T.copy()
    ok, sign = _run_o3(t, tmp_path)
    want = oracle.orient3d_batch(soa, oracle.EXACT).astype(np.int64)
    for k in range(4):
        assert np.array_equal(sign[ok[:, k], k], want[ok[:, k]]), k
    assert ok[kind == 1, 0].all() and (want[kind == 1] == 0).all()
    assert ok[(kind == 2) & (want != 0), 1].all()
    left = ~ok[:, 0] & ~ok[:, 1]
    assert left.sum() > 0 and (want[left] == 0).all()
    assert ok[left, 2].mean() > 0.9              # the rest span more than 62 bits
    assert (ok[left, 2] | ok[left, 3]).all()     # and take the 126-bit stage
    # random integer tuples of up to 62-bit spread, scaled by powers of two
    rng = np.random.default_rng(8)
    rows = []
    for bits in (5, 30, 61):
        for scale in (-250, 0, 200):
            q = rng.integers(-(2 ** bits), 2 ** bits, size=(50, 12)).astype(np.float64)
            q[:10, 3:6] = q[:10, 0:3]                       # repeated point: det 0
            rows.append(np.ldexp(q, scale))
    t = np.concatenate(rows)
    ok, sign = _run_o3(t, tmp_path)
    want = oracle.orient3d_batch(np.ascontiguousarray(t.T), oracle.EXACT).astype(np.int64)
    assert np.array_equal(sign[ok[:, 2], 2], want[ok[:, 2]])
    assert ok[:, 2].mean() > 0.95 and (want[ok[:, 2]] == 0).any()
    # 64..120-bit spreads: only the 126-bit stage decides
    rows = []
    for spread in (64, 90, 120):
        big = np.ldexp(rng.integers(-(2 ** 40), 2 ** 40, size=(60, 12)).astype(np.float64),
                       spread - 41)
        small = rng.integers(1, 2 ** 20, size=(60, 12)).astype(np.float64)
        q = big + np.where(rng.random((60, 12)) < 0.3, small, 0.0)
        q[:10, 3:6] = q[:10, 0:3]
        rows.append(q)
    t = np.concatenate(rows)
    ok, sign = _run_o3(t, tmp_path)
    want = oracle.orient3d_batch(np.ascontiguousarray(t.T), oracle.EXACT).astype(np.int64)
    assert np.array_equal(sign[ok[:, 3], 3], want[ok[:, 3]])
    assert ok[:, 3].mean() > 0.95 and (want[ok[:, 3]] == 0).any()


def test_orient3d_fixed31_limits(tmp_path):
    """The 31-bit orient3d stage (the ring drain's first, slot 0): with the oracle's exact
    sign wherever it decides; decides every tuple whose differences are exact, lie on one
    grid and stay below 2^31 grid steps, at any power-of-two scale in range (exact zeros
    from repeated and coplanar points included); declines inexact differences, spreads
    beyond 31 bits and largest differences outside [2^-260, 2^31)."""
    rng = np.random.default_rng(12)
    rows, expect = [], []

    def add(q, decided):
        rows.append(q)
        expect.append(np.full(len(q), decided))

    for bits in (3, 20, 28):
        for scale in (-240, -40, 0, 1):
            q = rng.integers(-(2 ** bits), 2 ** bits, size=(60, 12)).astype(np.float64)
            q[:10, 3:6] = q[:10, 0:3]                                  # repeated point
            q[10:20, 9:12] = q[10:20, 3:6] + q[10:20, 6:9] - q[10:20, 0:3]  # coplanar
            add(np.ldexp(q, scale), True)
    # at the width limit: largest difference 2^31 - 1 grid steps, smallest 1
    lim = np.zeros((4, 12))
    lim[:, 0] = 2.0 ** 31 - 1
    lim[:, 4] = 1.0
    lim[:, 8] = rng.integers(1, 2 ** 20, size=4)
    lim[1:, 9] = [1.0, 0.0, 3.0]
    add(lim, True)
    wide = lim.copy()
    wide[:, 0] = 2.0 ** 31 + wide[:, 9]                                # 32 bits
    add(wide, False)
    # 41-bit spread with a small largest difference: not on a 31-bit grid
    q = rng.integers(2 ** 39, 2 ** 40, size=(20, 12)).astype(np.float64)
    q[:, 1] += 1.0
    add(np.ldexp(q, -20), False)
    # inexact difference: 1 - 2^-60
    q = rng.integers(-4, 5, size=(20, 12)).astype(np.float64)
    q[:, 0] = 1.0
    q[:, 9] = 2.0 ** -60
    add(q, False)
    # below range: every difference under 2^-260
    add(np.ldexp(rng.integers(-8, 9, size=(20, 12)).astype(np.float64), -275), False)
    t = np.concatenate(rows)
    expect = np.concatenate(expect)
    ok, sign = _run_o3(t, tmp_path)
    want = oracle.orient3d_batch(np.ascontiguousarray(t.T), oracle.EXACT).astype(np.int64)
    assert np.array_equal(sign[ok[:, 0], 0], want[ok[:, 0]])
    assert np.array_equal(ok[:, 0], expect)
    assert (want[expect] == 0).sum() >= 200 and (want[expect] != 0).sum() >= 200


def _run_exp(mode, tuples, tmp_path):
    """Host replay of the exact pass (register_stages <mode>-exp): decided flags and
    signs of fixed point (62-, 126-bit differences), then the expansions with the first
    pass's capacities and with the wide re-run's."""
    if not os.path.exists(BIN):
        pytest.skip("tests/native/register_stages not built (run the package build)")
    per = 12 if mode == "orient3d" else 15
    t = np.ascontiguousarray(tuples, dtype=np.float64).reshape(-1, per)
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    t.tofile(src)
    subprocess.run([BIN, mode + "-exp", str(src), str(dst)], check=True)
    out = np.fromfile(dst, dtype=np.int8).reshape(-1, 4, 2).astype(np.int64)
    return out[:, :, 0] > 0, out[:, :, 1]


@pytest.mark.parametrize("npts", [4, 5])
def test_expansion_passes_on_wide_spread_tuples(npts, tmp_path):
    """predicates.cpp:1609-1777 (orient3dexact) and :3237-3488 (insphereexact), restated as
    the device's expansion passes and run on the host: coordinates with independent
    exponents over +-300 (orient3d) / +-150 (insphere) and a repeated point (exact zeros,
    then 1-ulp moves). The fixed-point stages decline every tuple; the expansions decide
    every one with the oracle's sign; for orient3d some outgrow the first pass's capacity
    and are decided by the wide re-run."""
    soa = W.tuples_wide_spread(3000, npts, seed=40 + npts)
    ok, sign = _run_exp("orient3d" if npts == 4 else "insphere", soa.T, tmp_path)
    want = (oracle.orient3d_batch(soa, oracle.EXACT) if npts == 4
            else oracle.insphere_batch(soa, oracle.EXACT)).astype(np.int64)
    assert not ok[:, 0].any() and not ok[:, 1].any()      # beyond the fixed-point tiers
    assert ok[:, 3].all()                                 # the wide capacities suffice
    for k in (2, 3):
        assert np.array_equal(sign[ok[:, k], k], want[ok[:, k]]), k
    assert (want == 0).sum() >= 1500 and (want > 0).any() and (want < 0).any()
    if npts == 4:
        assert 0.02 < (~ok[:, 2]).mean() < 0.5            # some lanes need the wide pass
