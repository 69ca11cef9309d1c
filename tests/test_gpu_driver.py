"""Parity of the particle volumes, the conservation diagnostics and the device driver loop (through the C ABI) with
the oracle on identical seeded inputs: volumes, masses, snapshot times, snapshot states and step counts bit-exact."""
import numpy as np
import pytest

from hcc_hip import Grid, Stepper
from hcc_hip import _lib as L
from hcc_hip import workloads as W
from helpers import build_pair, first_diff
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["adv2d_n64", "euler2d_n48"])
def test_volumes_bit_exact(ctx, case):
    """ShepardVolumes at grid creation (one sweep) and recomputed with three sweeps (Volumes.jl:183-231)."""
    wl = W.advection_2d(64, 0) if case == "adv2d_n64" else W.euler_2d(n=48, seed=3)
    g, st, o = build_pair(ctx, wl)
    v, v_o = g.volumes(), o.volumes()
    assert np.array_equal(v, v_o), first_diff(v, v_o)
    v3, v3_o = g.volumes(3), o.volumes(3)
    assert np.array_equal(v3, v3_o), first_diff(v3, v3_o)
    assert abs(v3.sum() - 1.0) < 1e-13


def test_masses_bit_exact_after_steps(ctx):
    """total_mass / interior_mass (Volumes.jl:103-117) of the device state after three Euler steps with MOOD firing."""
    wl = W.euler_2d(n=48, seed=3)
    g, st, o = build_pair(ctx, wl)
    t = 0.0
    for _ in range(3):
        dt = wl["cfl"] * st.time_step()
        st.step(t, dt)
        o.step(t, dt)
        t += dt
    for interior in (False, True):
        m, m_o = st.mass(interior), o.mass(interior)
        assert np.array_equal(m, m_o), (interior, m, m_o)


def test_solve_matches_oracle_driver_advection(ctx):
    """hcc_solve against the oracle's solve_equation (test/time_integration.jl:72-132) on the 64^2 periodic advection
    cloud to t = 0.2 with 5 snapshots: times, states, masses and the step count bit-exact."""
    wl = W.advection_2d(64, 0)
    g, st, o = build_pair(ctx, wl)
    xs, us, ts, steps, el, ms = st.solve(0.2, wl["cfl"], True, 5)
    xs_o, us_o, ts_o, k_o, ms_o = O.solve_equation(o, 0.2, wl["cfl"], True, 5, pos=wl["cloud"]["pos"])
    assert steps == k_o and list(ts) == ts_o
    for c in range(6):
        assert np.array_equal(us[c], us_o[c]), (c, first_diff(us[c], us_o[c]))
        assert np.array_equal(xs[c], xs_o[c])
        assert np.array_equal(ms[c, 0], ms_o[c][0]) and np.array_equal(ms[c, 1], ms_o[c][1])
    assert el > 0


def test_solve_matches_oracle_driver_euler_without_ghosts(ctx):
    """The four-quadrant Euler problem with outflow ghosts, fixed dt (is_cfl = false) and remove_ghosts: the snapshots
    hold the interior particles only, in order."""
    wl = W.euler_2d(n=48, seed=3)
    g, st, o = build_pair(ctx, wl)
    dt = 0.4 * o.time_step()
    isb = wl["cloud"]["is_boundary"]
    xs, us, ts, steps, el, ms = st.solve(0.01, dt, False, 3, remove_ghosts=True)
    xs_o, us_o, ts_o, k_o, ms_o = O.solve_equation(o, 0.01, dt, False, 3, remove_ghosts=True, isb=isb,
                                                   pos=wl["cloud"]["pos"])
    assert steps == k_o and list(ts) == ts_o
    assert us.shape[1] == int((~isb.astype(bool)).sum())
    for c in range(4):
        assert np.array_equal(us[c], us_o[c]), (c, first_diff(us[c], us_o[c]))
        assert np.array_equal(xs[c], xs_o[c])
        assert np.array_equal(ms[c], np.array(ms_o[c]))
