"""Pins of MOODu2's smooth-extremum override (MOOD.jl:108-133, 223-276) in the oracle, with MUSCL3's curvature
coefficients: the reference's own unit-test cases (unit_core.jl:781-874), on the same 20 x 20 periodic lattice with
hand-set states and curvature coefficients, so the criterion is tested on its own."""
import numpy as np
import pytest

from hcc_hip import _lib as L
from hcc_hip import workloads as W
from oracle import oracle as O

N_SIDE = 20
DX = 1.0 / N_SIDE
DELTA = 1e-8


@pytest.fixture(scope="module")
def lattice():
    dom = W.domain(2, (0.0, 0.0), (1.0, 1.0), (True, True), DX, 2.5, 0.0, 0)
    c = O.cloud_generate(dom)
    sch = W.scheme(2, L.HCC_PDE_LINADV, adv_speed=[[1.0, 1.0], [1.0, 1.0]], limiter=L.HCC_LIM_NONE,
                   strategy=L.HCC_STRAT_EPD2, criterion=L.HCC_CRIT_U2, order=3)
    o = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"], sch)
    r, idx, _, _ = o.neighbors()
    return o, c["pos"], r, idx


def _fields(x, f1, d1, f2, d2):
    """States (N, 2) and gradient records (N, 5, 2): the curvature terms x^2, y^2 (entries D, D + 1) set, the rest 0."""
    U = np.column_stack([f1(x), f2(x)])
    G = np.zeros((len(x), 5, 2))
    c1, c2 = d1(x), d2(x)
    G[:, 2, 0], G[:, 3, 0] = c1
    G[:, 2, 1], G[:, 3, 1] = c2
    return U, G


def _flags(o, U, G, p, cand, plateau=1e-3, ratio=0.5, orders=None):
    o_ = np.full(o.N, 3, np.int32) if orders is None else orders
    return o.u2_flags3(U, p, np.asarray(cand, float), DELTA, plateau, ratio, G, o_)


def _nearest(x, c):
    return int(np.argmin(np.sum((x - np.asarray(c)) ** 2, axis=1)))


def _local_extrema(U, r, idx, p, m):
    v = np.concatenate([[U[p, m]], U[idx[r[p]:r[p + 1]], m]])
    return v.min(), v.max()


W2 = 0.1
C0 = np.array([0.5, 0.5])


def gauss(x):
    return np.exp(-np.sum((x - C0) ** 2, axis=1) / W2)


def test_smooth_minimum_next_to_an_inflection_line(lattice):
    o, x, r, idx = lattice
    f1 = lambda x: 1 - 0.5 * gauss(x)
    d1 = lambda x: (-0.5 * gauss(x) * (4 * (x[:, 0] - C0[0]) ** 2 / W2 ** 2 - 2 / W2),
                    -0.5 * gauss(x) * (4 * (x[:, 1] - C0[1]) ** 2 / W2 ** 2 - 2 / W2))
    f2 = lambda x: np.sin(2 * np.pi * x[:, 0])
    d2 = lambda x: (-(2 * np.pi) ** 2 * np.sin(2 * np.pi * x[:, 0]), np.zeros(len(x)))
    U, G = _fields(x, f1, d1, f2, d2)
    p = _nearest(x, C0)
    lo1, _ = _local_extrema(U, r, idx, p, 0)
    cand = [lo1 - 10 * DELTA, U[p, 1]]
    assert _flags(o, U, G, p, cand) is False                 # a smooth minimum undershot: accepted
    lo2, _ = _local_extrema(U, r, idx, p, 1)
    assert _flags(o, U, G, p, [U[p, 0], lo2 - 10 * DELTA]) is True   # the inflection line: the curvature changes sign
    orders = np.full(o.N, 3, np.int32)
    orders[p] = 2
    assert _flags(o, U, G, p, cand, orders=orders) is True   # below order 3: the plain DMP


def test_plateau(lattice):
    o, x, r, idx = lattice
    q = (0.1, 0.1)
    narrow = lambda x: np.exp(-np.sum((x - C0) ** 2, axis=1) / 0.01)
    ripple = lambda x: 1e-6 * np.sin(20 * np.pi * x[:, 0] + 0.7)
    g1 = lambda x: 1 - 0.5 * narrow(x) + ripple(x)
    dg1 = lambda x: (-1e-6 * (20 * np.pi) ** 2 * np.sin(20 * np.pi * x[:, 0] + 0.7), np.zeros(len(x)))
    zero = lambda x: np.zeros(len(x))
    U, G = _fields(x, g1, dg1, zero, lambda x: (zero(x), zero(x)))
    p = _nearest(x, q)
    lo, hi = _local_extrema(U, r, idx, p, 0)
    assert hi - lo < 1e-5
    cand = [hi + 5 * DELTA, 0.0]
    assert _flags(o, U, G, p, cand, plateau=0.0) is True     # the ripple's curvature changes sign
    assert _flags(o, U, G, p, cand, plateau=1e-3) is False   # local range 2e-6 < 1e-3 x global range 0.5
    orders = np.full(o.N, 3, np.int32)
    orders[p] = 1
    assert _flags(o, U, G, p, cand, plateau=1e-3, orders=orders) is False   # a plateau needs no curvature


def test_a_jump_is_flagged_and_the_corner_needs_the_ratio_test(lattice):
    o, x, r, idx = lattice
    zero = lambda x: np.zeros(len(x))
    step = lambda x: np.where(x[:, 0] < 0.5, 0.0, 1.0)
    dstep = lambda x: (np.where(np.abs(x[:, 0] - 0.5) < 2.5 * DX, np.where(x[:, 0] < 0.5, 1 / DX ** 2, -1 / DX ** 2),
                                0.0), zero(x))
    U, G = _fields(x, step, dstep, zero, lambda x: (zero(x), zero(x)))
    p = _nearest(x, (0.5 - DX / 2, 0.5))
    _, hi = _local_extrema(U, r, idx, p, 0)
    cand = [hi + 1e-3, 0.0]
    for plateau in (0.0, 1e-3, 1e-2):
        assert _flags(o, U, G, p, cand, plateau=plateau) is True
    corner = lambda x: (np.where((0.5 <= x[:, 0]) & (x[:, 0] < 0.5 + DX / 2), -1 / DX ** 2, 0.0), zero(x))
    U, G = _fields(x, step, corner, zero, lambda x: (zero(x), zero(x)))
    p = _nearest(x, (0.5 + DX, 0.5))
    _, hi = _local_extrema(U, r, idx, p, 0)
    cand = [hi + 1e-3, 0.0]
    assert _flags(o, U, G, p, cand) is True
    assert _flags(o, U, G, p, cand, ratio=0.0) is False      # without the ratio test the collapse is not seen


def test_muscl2_keeps_the_plain_dmp(lattice):
    """MUSCL order 2 has no curvature coefficients: a violating component is flagged whatever the gradients hold."""
    o3, x, r, idx = lattice
    dom = W.domain(2, (0.0, 0.0), (1.0, 1.0), (True, True), DX, 2.5, 0.0, 0)
    c = O.cloud_generate(dom)
    o2 = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"],
                     W.scheme(2, L.HCC_PDE_LINADV, adv_speed=[[1.0, 1.0], [1.0, 1.0]], criterion=L.HCC_CRIT_U2))
    f1 = lambda x: 1 - 0.5 * gauss(x)
    U = np.column_stack([f1(x), np.zeros(len(x))])
    p = _nearest(x, C0)
    lo1, _ = _local_extrema(U, r, idx, p, 0)
    G2 = np.zeros((len(x), 2, 2))
    assert o2.u2_flags3(U, p, np.array([lo1 - 10 * DELTA, 0.0]), DELTA, 1e-3, 0.5, G2,
                        np.full(o2.N, 2, np.int32)) is True
