"""Pins of the oracle's IMEX stage (implicit part of the tableau, diagonal solve, K_I) and its source terms against
the reference's own scheme tests (test/unit_schemes.jl:275-346) and closed forms."""
import math

import numpy as np
import pytest

from hcc_hip import _lib as L
from hcc_hip import workloads as W
from oracle import oracle as O


def _line(n, tableau, c=1.0):
    """make_pg_1d(n) and make_advection_1d(c) of the reference tests, MUSCL2 without limiter, no MOOD."""
    dom = W.domain(1, (-1,), (1,), (True,), 2 / n, 2.5, 0.0, 0)
    cl = O.cloud_generate(dom)
    sch = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[c]], limiter=L.HCC_LIM_NONE, strategy=L.HCC_STRAT_NONE,
                   criterion=L.HCC_CRIT_NONE, tableau_name=tableau)
    o = O.OracleSim(dom, cl["pos"], cl["is_boundary"], cl["tags"], sch)
    return o, cl["pos"][:, 0]


def test_implicit_stage_rhs_from_the_solve():
    """unit_schemes.jl:275-300: with one implicit source, K_I of a diagonal stage is (U_out - U_in) / (dt a_ii) = S(U_out)
    = -lambda Y_i; a constant state decays to exp(-lambda dt) to second order (ARS222)."""
    o, x = _line(40, "IMEX_ARS222")
    lam, dt = 3.0, 0.1
    o.set_source(2, np.full(o.N, lam))
    o.set_state(np.ones((o.N, 1)))
    o.step(0.0, dt)
    tb = W.tableau("IMEX_ARS222")
    checked = 0
    for i in range(3):
        if abs(tb["a_im"][i][i]) <= 1e-14:
            continue
        Y, K = o.stage("Y", i), o.stage("K_I", i)
        assert np.allclose(K, -lam * Y, rtol=1e-12, atol=0)
        checked += 1
    assert checked >= 1
    assert np.max(np.abs(o.state() - math.exp(-lam * dt))) < 5e-3


@pytest.mark.parametrize("name,order", [("IMEX_ARS222", 2), ("IMEX_ARS233", 3), ("IMEX_PRSSP3", 2),
                                        ("IMEX_SSP2332", 2)])
def test_imex_tableaus_observed_order(name, order):
    """unit_schemes.jl:302-324: advection (explicit) with a decay of rate 3 (1 + 0.8 sin(pi x)) (implicit) on a
    periodic 20-particle line; the error against a fine-dt run shrinks at the tableau's order."""
    def run(n, tend=0.4):
        o, x = _line(20, name)
        o.set_source(2, 3.0 * (1 + 0.8 * np.sin(np.pi * x)))
        o.set_state(np.sin(np.pi * x)[:, None])
        dt = tend / n
        for k in range(n):
            o.step(k * dt, dt)
        return o.state()[:, 0]
    ref = run(640)
    e1, e2 = np.max(np.abs(run(20) - ref)), np.max(np.abs(run(40) - ref))
    assert math.log2(e1 / e2) > order - 0.2


def test_implicit_source_with_an_explicit_tableau_is_applied_explicitly():
    """unit_schemes.jl:326-346: with RK3 (no implicit part) the implicit decay is evaluated explicitly with the
    explicit weights: one step of u' = -lambda u from u = 1 is the cubic Taylor polynomial of exp(-lambda dt). An
    explicit source gives the same step."""
    lam, dt = 3.0, 0.1
    z = -lam * dt
    for mode in (2, 1):
        o, x = _line(20, "RK3_SSP")
        o.set_source(mode, np.full(o.N, lam))
        o.set_state(np.ones((o.N, 1)))
        o.step(0.0, dt)
        assert np.allclose(o.state(), 1 + z + z ** 2 / 2 + z ** 3 / 6, rtol=1e-13, atol=0)
    assert W.tableau("RK3_SSP")["substep_ratio"] == pytest.approx(1.0, abs=1e-9)


def test_imex_tableaus_order_conditions():
    """ButcherTableaus.jl:183-264: both parts of each IMEX tableau are consistent (row sums = c, sum b = 1) and
    second order (b . c = 1/2); ARS233's parts are third order (b . c^2 = 1/3, b . A c = 1/6)."""
    for name, t in W.IMEX_TABLEAUS.items():
        for a, b, c in ((t["a_t"], t["b_t"], t["c_t"]), (t["a_im"], t["b_im"], t["c_im"])):
            a, b, c = np.array(a), np.array(b), np.array(c)
            assert np.allclose(a.sum(1), c, atol=1e-15) and abs(b.sum() - 1) < 1e-15
            if name != "IMEX_Euler":
                assert abs(b @ c - 0.5) < 1e-15
            if name == "IMEX_ARS233":
                assert abs(b @ c ** 2 - 1 / 3) < 1e-15 and abs(b @ a @ c - 1 / 6) < 1e-15
