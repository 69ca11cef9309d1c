"""Parity of the IMEX stage and the source terms (through the C ABI) with the oracle: every step bit-exact."""
import numpy as np
import pytest

from hcc_hip import _lib as L
from hcc_hip import workloads as W
from helpers import build_pair, first_diff, rel_linf

pytestmark = pytest.mark.gpu


def _steps(st, o, cfl, n):
    t = 0.0
    for k in range(n):
        dt_g, dt_o = st.time_step(), o.time_step()
        assert dt_g == dt_o
        dt = cfl * dt_g
        st.step(t, dt)
        o.step(t, dt)
        t += dt
        U, U_o = st.download(), o.state()
        assert np.array_equal(st.orders(), o.orders()), f"step {k}: orders"
        assert np.array_equal(st.mood_flags(), o.mood_flags()), f"step {k}: MOOD flags"
        assert rel_linf(U, U_o) <= 1e-12
        assert np.array_equal(U, U_o), f"step {k}: {first_diff(U, U_o)}"


def _advection(tableau, mode, limiter=L.HCC_LIM_VK):
    wl = W.advection_2d(64, 0)
    sch = wl["scheme"]
    wl["scheme"] = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=sch["adv_speed"], delta=sch["delta"], limiter=limiter,
                            tableau_name=tableau)
    x = wl["cloud"]["pos"]
    rates = 3.0 * (1 + 0.8 * np.sin(np.pi * x[:, 0]))
    return wl, rates


@pytest.mark.parametrize("tableau", ["IMEX_ARS222", "IMEX_ARS233", "IMEX_PRSSP3", "IMEX_SSP2332", "IMEX_Euler"])
def test_imex_decay_advection(ctx, tableau):
    """IMEX tableaus (ButcherTableaus.jl:183-264) with the implicit decay of rate 3 (1 + 0.8 sin(pi x)) on the 64^2
    periodic advection cloud with MOOD (stage candidates checked): 6 steps bit-exact."""
    wl, rates = _advection(tableau, L.HCC_SRC_DECAY_IMPLICIT)
    g, st, o = build_pair(ctx, wl)
    st.set_source(L.HCC_SRC_DECAY_IMPLICIT, rates)
    o.set_source(2, rates)
    _steps(st, o, wl["cfl"], 6)


def test_imex_tableau_without_source(ctx):
    """An implicit tableau part with no source: K_I = 0, its terms still enter the stages (signed zeros included)."""
    wl, _ = _advection("IMEX_ARS222", L.HCC_SRC_NONE)
    g, st, o = build_pair(ctx, wl)
    _steps(st, o, wl["cfl"], 4)


@pytest.mark.parametrize("mode", [L.HCC_SRC_DECAY_EXPLICIT, L.HCC_SRC_DECAY_IMPLICIT])
def test_decay_with_rk3_euler(ctx, mode):
    """SSP-RK3 on the four-quadrant Euler problem with a decay of rate 0.5 (1 + x): explicit (added to K_E in the
    fused flux passes) or implicit applied explicitly (_explicit_implicit_part; MOOD checks Y + h (K + K_I))."""
    wl = W.euler_2d(n=48, seed=3)
    g, st, o = build_pair(ctx, wl)
    rates = 0.5 * (1 + wl["cloud"]["pos"][:, 0])
    st.set_source(mode, rates)
    o.set_source(int(mode), rates)
    _steps(st, o, wl["cfl"], 6)


def test_imex_euler_ars233_unlimited_mood(ctx):
    """ARS233 with the implicit decay on the Euler quadrants: the stage-candidate MOOD fixed point with K_I."""
    wl = W.euler_2d(n=48, seed=3)
    sch = wl["scheme"]
    wl["scheme"] = W.scheme(4, L.HCC_PDE_EULER, delta=sch["delta"], bcs=sch["bcs"], tableau_name="IMEX_ARS233")
    g, st, o = build_pair(ctx, wl)
    rates = np.full(g.N, 0.8)
    st.set_source(L.HCC_SRC_DECAY_IMPLICIT, rates)
    o.set_source(2, rates)
    _steps(st, o, wl["cfl"], 5)
