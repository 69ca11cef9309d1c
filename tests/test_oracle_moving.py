"""Pins of the oracle's moving grids (GridMovers.jl): ALE relative flux, stage positions, frozen-topology refresh and
the end-of-step search, against the reference's own moving-grid tests (test/unit_schemes.jl:774-860) and closed
forms."""
import math

import numpy as np
import pytest

from hcc_hip import _lib as L
from hcc_hip import workloads as W
from oracle import oracle as O


def _line_sim(n, c, mover, alpha=1.0, limiter=L.HCC_LIM_VK, tableau="RK2_SSP"):
    dom = W.domain(1, (-1,), (1,), (True,), 2 / n, 2.5, 0.0, 0)
    cl = O.cloud_generate(dom)
    sch = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[c]], limiter=limiter, strategy=L.HCC_STRAT_NONE,
                   criterion=L.HCC_CRIT_NONE, tableau_name=tableau, mover=mover, mover_alpha=alpha)
    return O.OracleSim(dom, cl["pos"], cl["is_boundary"], cl["tags"], sch), cl["pos"][:, 0]


def test_refresh_geometry_matches_the_search_and_respects_periodicity():
    """unit_schemes.jl:787-806: a refresh at unchanged positions reproduces the search's distances and weights
    bit for bit; after a rigid translation that leaves the canvas the minimal-image distances and the weights are
    unchanged to 1e-12; a particle moved beyond R keeps its slots with weight 0; a full search re-wraps positions."""
    o, x0 = _line_sim(60, 0.7, L.HCC_MOVER_LAGRANGIAN)
    r, idx, w0, d0 = o.neighbors()
    o.set_positions(o.positions())
    _, idx1, w1, d1 = o.neighbors()
    assert np.array_equal(idx1, idx) and np.array_equal(w1, w0) and np.array_equal(d1, d0)
    o.set_positions(o.positions() + 0.37 * 2.0)
    _, idx2, w2, d2 = o.neighbors()
    assert np.array_equal(idx2, idx)
    assert np.max(np.abs(d2 - d0)) < 1e-12 and np.max(np.abs(w2 - w0)) < 1e-12
    x = o.positions()
    x[0, 0] += 0.5
    o.set_positions(x)
    _, _, w3, _ = o.neighbors()
    assert np.all(w3[r[0]:r[1]] == 0.0)
    far = idx[r[2]:r[3]] != 0
    assert np.all(w3[r[2]:r[3]][far] == w2[r[2]:r[3]][far])   # pairs without particle 0 are untouched
    o.set_positions(o.positions(), full_search=True)
    xs = o.positions()[:, 0]
    assert np.all((xs >= -1.0) & (xs < 1.0))
    _, idx4, w4, _ = o.neighbors()
    assert np.all(w4 > 0)                                        # the new lists hold only pairs within R


@pytest.mark.parametrize("limiter", [L.HCC_LIM_NONE, L.HCC_LIM_VK, L.HCC_LIM_BJ])
def test_lagrangian_transport_is_exact(limiter):
    """unit_schemes.jl:826-845: w = u freezes the relative flux F - w rho to zero, so rho is carried unchanged by the
    particles, which move with the flow (positions wrapped into the periodic domain)."""
    N, c = 60, 0.7
    o, x0 = _line_sim(N, c, L.HCC_MOVER_LAGRANGIAN, limiter=limiter)
    r0 = 0.5 + 0.3 * np.sin(np.pi * x0)
    o.set_state(r0[:, None])
    dt = 0.4 * (2 / N) / c
    for k in range(12):
        o.step(k * dt, dt)
    t = 12 * dt
    wrap = lambda x: -1.0 + np.mod(x + 1.0, 2.0)
    assert np.max(np.abs(o.state()[:, 0] - r0)) < 1e-10
    assert np.max(np.abs(o.positions()[:, 0] - wrap(x0 + c * t))) < 1e-10


def test_scaled_lagrangian_motion_converges():
    """unit_schemes.jl:847-860 with ScaledLagrangianMover(0.5) in place of the prescribed field: the grid moves at
    half the advection speed, the scheme sees the relative speed, and the error against the exact transport
    ic(x - c t) converges at second order (MUSCL2, RK3)."""
    c = 0.7
    errs = []
    for n in (60, 120):
        o, x0 = _line_sim(n, c, L.HCC_MOVER_SCALED, alpha=0.5, limiter=L.HCC_LIM_NONE, tableau="RK3_SSP")
        o.set_state((0.5 + 0.3 * np.sin(np.pi * x0))[:, None])
        dt, t = 0.3 * (2 / n) / c, 0.0
        while t < 0.5 - 1e-12:
            h = min(dt, 0.5 - t)
            o.step(t, h)
            t += h
        x = o.positions()[:, 0]
        exact = 0.5 + 0.3 * np.sin(np.pi * (x - c * t))
        errs.append(math.sqrt(np.mean((o.state()[:, 0] - exact) ** 2)))
        assert np.max(np.abs(x - (-1.0 + np.mod(x0 + 0.5 * c * t + 1.0, 2.0)))) < 1e-10
    assert math.log2(errs[0] / errs[1]) > 1.5


def test_lagrangian_euler_is_galilean_invariant():
    """Euler with a uniform velocity u0 and pressure and a density wave on a periodic cloud, Lagrangian mover (w = u):
    the relative flux F - w U and the shifted wave speed |u - w| + c make the run equal, to rounding, the fixed-grid
    run of the same density and pressure at rest, seen from a frame moving with u0 (rho equal, rho u = u0 rho,
    E = E_rest + |u0|^2 rho / 2), while the particles move with u0. A wrong sign of w or an unshifted wave speed
    fails it."""
    h = 1 / 24
    dom = W.domain(2, (0, 0), (1, 1), (True, True), h, 2.5, 0.0, 0)
    cl = O.cloud_generate(dom)
    x = cl["pos"]
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * x[:, 0]) * np.cos(2 * np.pi * x[:, 1])
    u0 = np.array([0.6, -0.4])
    sims = []
    for mover, vel in ((L.HCC_MOVER_LAGRANGIAN, u0), (L.HCC_MOVER_NONE, np.zeros(2))):
        sch = W.scheme(4, L.HCC_PDE_EULER, delta=1e-4 * h * h, limiter=L.HCC_LIM_NONE, strategy=L.HCC_STRAT_NONE,
                       criterion=L.HCC_CRIT_NONE, mover=mover)
        o = O.OracleSim(dom, cl["pos"], cl["is_boundary"], cl["tags"], sch)
        o.set_state(W.euler_conserved(rho, np.tile(vel, (len(x), 1)), np.ones(len(x)), 1.4))
        sims.append(o)
    mv, rest = sims
    dt, t = 0.3 * rest.time_step(), 0.0
    for k in range(5):
        mv.step(t, dt)
        rest.step(t, dt)
        t += dt
    U, R = mv.state(), rest.state()
    assert np.max(np.abs(R[:, 0] - rho)) > 1e-4          # the density did change (Rusanov diffusion of the wave)
    assert np.max(np.abs(U[:, 0] - R[:, 0])) < 1e-12
    assert np.max(np.abs(U[:, 1:3] - u0 * R[:, :1])) < 1e-12
    assert np.max(np.abs(U[:, 3] - (R[:, 3] + 0.5 * (u0 @ u0) * R[:, 0]))) < 1e-12
    moved = np.mod(x + u0 * t, 1.0)
    d = np.abs(mv.positions() - moved)
    assert np.max(np.minimum(d, 1 - d)) < 1e-12
