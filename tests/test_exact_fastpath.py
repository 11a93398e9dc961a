"""The exact stage's register-only path (expansion.hpp, orient3d_sign_exact_diffs and
sum_sign): replayed in Python float64 (error-free Two-Sum, Dekker Two-Product, which give
the same exact pairs as the FMA form) and checked against exact rational arithmetic.
Decided lanes must carry the exact sign; most coplanar grid tuples must be decided."""
import numpy as np

from cybermeshgenerator_amd import workloads as W
from tests.meshes import frac_orient3d

SPLIT = 134217729.0


def two_sum(a, b):
    s = a + b
    bv = s - a
    av = s - bv
    return s, (a - av) + (b - bv)


def two_prod(a, b):
    p = a * b
    c = SPLIT * a
    ah = c - (c - a)
    al = a - ah
    c = SPLIT * b
    bh = c - (c - b)
    bl = b - bh
    return p, al * bl - (((p - ah * bh) - al * bh) - ah * bl)


def sum_sign(v, maxpass=6):
    v = list(v)
    n = len(v)
    for _ in range(maxpass):
        for i in range(1, n):
            v[i], v[i - 1] = two_sum(v[i], v[i - 1])
        tail = 0.0
        for x in v[:-1]:
            tail += abs(x)
        top = v[-1]
        if abs(top) > tail * (1.0 + 2.0 ** -40):
            return 1 if top > 0 else -1
        if top == 0.0 and tail == 0.0:
            return 0
    return None


def fast_sign(p):
    a, b, c, d = p
    diffs = []
    for q in (a, b, c):
        for k in range(3):
            x, y = two_sum(float(q[k]), -float(d[k]))
            if y != 0.0:
                return None
            diffs.append(x)
    adx, ady, adz, bdx, bdy, bdz, cdx, cdy, cdz = diffs
    terms = [(bdx, cdy, adz), (-cdx, bdy, adz), (cdx, ady, bdz),
             (-adx, cdy, bdz), (adx, bdy, cdz), (-bdx, ady, cdz)]
    v = []
    for x, y, z in terms:
        pp, e = two_prod(x, y)
        v.extend(two_prod(pp, z))
        v.extend(two_prod(e, z))
    return sum_sign(v)


def test_fast_path_signs_are_exact():
    soa, kind = W.tuples_orient3d(6000, seed=31, degenerate=0.5)
    decided = 0
    coplanar_decided = 0
    for i in range(soa.shape[1]):
        p = soa[:, i].reshape(4, 3)
        s = fast_sign(p)
        if s is None:
            continue
        decided += 1
        coplanar_decided += kind[i] == 1
        assert s == frac_orient3d(*p), i
    assert coplanar_decided >= 0.9 * (kind == 1).sum()
    # small-integer tuples: every difference exact, all decided
    rng = np.random.default_rng(32)
    for _ in range(500):
        p = rng.integers(-50, 51, size=(4, 3)).astype(float)
        s = fast_sign(p)
        assert s == frac_orient3d(*p)


def fma(a, b, c):
    from fractions import Fraction
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def fast_two_sum(a, b):
    s = a + b
    return s, b - (s - a)


def dw_mul(x, y):
    ch, cl1 = two_prod(x[0], y[0])
    tl0 = x[1] * y[1]
    tl1 = fma(x[0], y[1], tl0)
    cl2 = fma(x[1], y[0], tl1)
    return fast_two_sum(ch, cl1 + cl2)


def dw_add(x, y):
    sh, sl = two_sum(x[0], y[0])
    th, tl = two_sum(x[1], y[1])
    sl += th
    vh, vl = fast_two_sum(sh, sl)
    vl += tl
    return fast_two_sum(vh, vl)


def dw_sign(p):
    a, b, c, d = p
    D = []
    for q in (a, b, c):
        for k in range(3):
            D.append(two_sum(float(q[k]), -float(d[k])))
    adx, ady, adz, bdx, bdy, bdz, cdx, cdy, cdz = D
    neg = lambda x: (-x[0], -x[1])
    m1 = dw_add(dw_mul(bdx, cdy), neg(dw_mul(cdx, bdy)))
    m2 = dw_add(dw_mul(cdx, ady), neg(dw_mul(adx, cdy)))
    m3 = dw_add(dw_mul(adx, bdy), neg(dw_mul(bdx, ady)))
    det = dw_add(dw_add(dw_mul(adz, m1), dw_mul(bdz, m2)), dw_mul(cdz, m3))
    A = [abs(x[0]) for x in D]
    perm = ((A[3] * A[7] + A[6] * A[4]) * A[2] + (A[6] * A[1] + A[0] * A[7]) * A[5] +
            (A[0] * A[4] + A[3] * A[1]) * A[8])
    bound = 2.0 ** -96 * perm * (1.0 + 2.0 ** -40)
    v = det[0] + det[1]
    if v > bound:
        return 1
    if -v > bound:
        return -1
    return None


def test_double_word_stage_signs_are_exact():
    soa, kind = W.tuples_orient3d(3000, seed=33, degenerate=1.0)
    decided_ulp = 0
    for i in range(soa.shape[1]):
        p = soa[:, i].reshape(4, 3)
        s = dw_sign(p)
        if s is None:
            continue
        assert s == frac_orient3d(*p), i
        decided_ulp += kind[i] == 2
    assert decided_ulp >= 0.9 * (kind == 2).sum()
