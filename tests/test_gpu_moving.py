"""Parity of moving grids (GridMovers.jl) through the C ABI with the oracle: ALE relative flux, stage positions with
the frozen-topology stage geometry, and the end-of-step search with the periodic wrap and ShepardVolumes. After every
step the CFL step, orders, MOOD flags, states, positions, neighbour lists and volumes are bit-exact."""
import numpy as np
import pytest

from hcc_hip import Grid, Stepper
from hcc_hip import _lib as L
from hcc_hip import workloads as W
from helpers import build_pair, first_diff
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _check_geometry(g, o, k):
    x, x_o = g.positions(), o.positions()
    assert np.array_equal(x, x_o), f"step {k}: positions {first_diff(x, x_o)}"
    r, idx, w, _ = g.neighbors(distances=False)
    r_o, idx_o, w_o, _ = o.neighbors()
    assert np.array_equal(r, r_o) and np.array_equal(idx, idx_o), f"step {k}: neighbour lists"
    assert np.array_equal(w, w_o), f"step {k}: weights {first_diff(w, w_o)}"
    v, v_o = g.volumes(), o.volumes()
    assert np.array_equal(v, v_o), f"step {k}: volumes {first_diff(v, v_o)}"


def _run(g, st, o, n, cfl=None, dt=None):
    t = 0.0
    for k in range(n):
        if cfl is not None:
            dt_g, dt_o = st.time_step(), o.time_step()
            assert dt_g == dt_o, f"step {k}: CFL step {dt_g} vs {dt_o}"
            h = cfl * dt_g
        else:
            h = dt
        print(f"step {k} dt {h}", flush=True)
        st.step(t, h)
        o.step(t, h)
        t += h
        U, U_o = st.download(), o.state()
        assert np.array_equal(st.orders(), o.orders()), f"step {k}: orders"
        assert np.array_equal(st.mood_flags(), o.mood_flags()), f"step {k}: MOOD flags"
        assert np.array_equal(U, U_o), f"step {k}: {first_diff(U, U_o)}"
        _check_geometry(g, o, k)
    return t


def _line(n, c, mover, alpha=1.0, limiter=L.HCC_LIM_VK, tableau="RK2_SSP"):
    dom = W.domain(1, (-1,), (1,), (True,), 2 / n, 2.5, 0.0, 0)
    cloud = L.cloud_generate(W.domain_desc(dom))
    x = cloud["pos"][:, 0]
    sch = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[c]], limiter=limiter, strategy=L.HCC_STRAT_NONE,
                   criterion=L.HCC_CRIT_NONE, tableau_name=tableau, mover=mover, mover_alpha=alpha)
    return {"domain": dom, "scheme": sch, "cloud": cloud, "U0": (0.5 + 0.3 * np.sin(np.pi * x))[:, None]}


@pytest.mark.parametrize("limiter", [L.HCC_LIM_NONE, L.HCC_LIM_VK, L.HCC_LIM_BJ])
def test_lagrangian_line_transport(ctx, limiter):
    """unit_schemes.jl:826-845 on the device: LagrangianMover on the periodic 60-particle line, RK2, 12 steps of
    0.4 dx / c: bit-exact with the oracle, rho carried unchanged and the particles moved by c t (wrapped)."""
    wl = _line(60, 0.7, L.HCC_MOVER_LAGRANGIAN, limiter=limiter)
    g, st, o = build_pair(ctx, wl)
    x0 = wl["cloud"]["pos"][:, 0].copy()
    t = _run(g, st, o, 12, dt=0.4 * (2 / 60) / 0.7)
    assert np.max(np.abs(st.download() - wl["U0"])) < 1e-10
    wrap = -1.0 + np.mod(x0 + 0.7 * t + 1.0, 2.0)
    assert np.max(np.abs(g.positions()[:, 0] - wrap)) < 1e-10


def test_scaled_lagrangian_line_rk3(ctx):
    """unit_schemes.jl:847-860 with ScaledLagrangianMover(0.5) (the grid moves at half the advection speed), MUSCL
    unlimited, RK3: every step bit-exact, to t = 0.5."""
    wl = _line(120, 0.7, L.HCC_MOVER_SCALED, alpha=0.5, limiter=L.HCC_LIM_NONE, tableau="RK3_SSP")
    g, st, o = build_pair(ctx, wl)
    _run(g, st, o, 30, dt=0.3 * (2 / 120) / 0.7)


def _advection_moving(alpha, n=48):
    wl = W.advection_2d(n, 0)
    h = 1.0 / n
    wl["scheme"] = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0], [1.0]], delta=1e-4 * h * h,
                            mover=L.HCC_MOVER_SCALED, mover_alpha=alpha)
    return wl


def test_scaled_lagrangian_advection_2d_mood(ctx):
    """The jittered periodic 48^2 advection cloud with ScaledLagrangianMover(0.3): MUSCL-VK, MOOD (EPD1, MOODu1 on
    the Euler substeps), RK3, CFL steps; the cloud drifts across the periodic boundary and the lists change."""
    wl = _advection_moving(0.3)
    g, st, o = build_pair(ctx, wl)
    _run(g, st, o, 6, cfl=wl["cfl"])


def _euler_moving(variant, n=40):
    h = 1.0 / n
    dom = W.domain(2, (0.0, 0.0), (1.0, 1.0), (True, True), h, 2.5, 0.25, 5)
    cloud = L.cloud_generate(W.domain_desc(dom))
    x = cloud["pos"]
    inside = (np.abs(x[:, 0] - 0.45) < 0.15) & (np.abs(x[:, 1] - 0.5) < 0.15)
    rho = np.where(inside, 2.0, 1.0)
    u = np.column_stack([0.5 + 0.2 * np.sin(2 * np.pi * x[:, 1]), -0.3 + 0.2 * np.sin(2 * np.pi * x[:, 0])])
    p = np.where(inside, 1.5, 1.0)
    kw = dict(delta=1e-4 * h * h, mover=L.HCC_MOVER_LAGRANGIAN)
    kw.update({
        "epd1_u1": {},
        "epd2_u2_plateau": dict(strategy=L.HCC_STRAT_EPD2, criterion=L.HCC_CRIT_U2, plateau=0.05),
        "strict_epd0_u1_aligned": dict(strategy=L.HCC_STRAT_STRICT_EPD0,
                                        flux=L.HCC_FLUX_ALIGNED_RUSANOV),
        "epd0_u1_stored_streamed": dict(strategy=L.HCC_STRAT_EPD0, mood_substeps=False, streamed=True),
        "scaled_bj_only": dict(limiter=L.HCC_LIM_BJ, criterion=L.HCC_CRIT_ONLY, mover=L.HCC_MOVER_SCALED),
    }[variant])
    if variant == "scaled_bj_only":
        kw["mover_alpha"] = 0.6
    sch = W.scheme(4, L.HCC_PDE_EULER, **kw)
    return {"domain": dom, "scheme": sch, "cloud": cloud, "U0": W.euler_conserved(rho, u, p, 1.4), "cfl": 0.4}


@pytest.mark.parametrize("variant", ["epd1_u1", "epd2_u2_plateau", "strict_epd0_u1_aligned", "epd0_u1_stored_streamed",
                                     "scaled_bj_only"])
def test_lagrangian_euler_2d_mood(ctx, variant):
    """Euler on a jittered periodic 40^2 cloud with a sheared flow and a density/pressure square, LagrangianMover
    (ScaledLagrangianMover(0.6) for the last case): the MOOD variants, the streamed and aligned flux loops and the
    unfused fixed point on the stage geometry, five CFL steps, every one bit-exact with the oracle."""
    wl = _euler_moving(variant)
    g, st, o = build_pair(ctx, wl)
    _run(g, st, o, 5, cfl=wl["cfl"])
    if variant == "epd1_u1":
        assert np.sum(st.mood_flags()) > 0   # MOOD fired on the moving grid


def test_imex_decay_on_a_moving_grid(ctx):
    """The IMEX stage (ARS233, implicit decay) with ScaledLagrangianMover(0.5) on the 48^2 advection cloud: the
    positions follow the explicit part of the tableau."""
    wl = _advection_moving(0.5)
    h = 1.0 / 48
    wl["scheme"] = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0], [1.0]], delta=1e-4 * h * h,
                            tableau_name="IMEX_ARS233", mover=L.HCC_MOVER_SCALED, mover_alpha=0.5)
    g, st, o = build_pair(ctx, wl)
    rates = 3.0 * (1 + 0.8 * np.sin(np.pi * wl["cloud"]["pos"][:, 0]))
    st.set_source(L.HCC_SRC_DECAY_IMPLICIT, rates)
    o.set_source(L.HCC_SRC_DECAY_IMPLICIT, rates)
    _run(g, st, o, 4, cfl=wl["cfl"])


def test_solve_on_a_moving_grid(ctx):
    """hcc_solve with ScaledLagrangianMover(0.5) against the oracle's solve_equation: snapshot times, positions,
    states and masses bit-exact, the step count equal."""
    wl = _advection_moving(0.5, n=32)
    g, st, o = build_pair(ctx, wl)
    xs, us, ts, steps, el, ms = st.solve(0.1, wl["cfl"], True, 3)
    xs_o, us_o, ts_o, k_o, ms_o = O.solve_equation(o, 0.1, wl["cfl"], True, 3)
    assert steps == k_o and list(ts) == ts_o
    for c in range(4):
        assert np.array_equal(xs[c], xs_o[c]), (c, first_diff(xs[c], xs_o[c]))
        assert np.array_equal(us[c], us_o[c]), (c, first_diff(us[c], us_o[c]))
        assert np.array_equal(ms[c, 0], ms_o[c][0]) and np.array_equal(ms[c, 1], ms_o[c][1])
    assert not np.array_equal(xs[0], xs[3])


def test_moving_grid_rejects_unsupported_setups(ctx):
    """Outflow boundary conditions and multi-component advection on a moving grid are refused at creation."""
    wl = W.euler_2d(n=24, seed=1)
    c = wl["cloud"]
    g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
    sch = dict(wl["scheme"], mover=L.HCC_MOVER_LAGRANGIAN)
    with pytest.raises(L.HccError, match="outflow"):
        Stepper(ctx, g, sch)
    wl2 = W.advection_2d(16, 0)
    c2 = wl2["cloud"]
    g2 = Grid(ctx, W.domain_desc(wl2["domain"]), c2["pos"], c2["is_boundary"], c2["tags"])
    sch2 = W.scheme(2, L.HCC_PDE_LINADV, adv_speed=[[1.0, 0.5], [1.0, 0.5]], mover=L.HCC_MOVER_LAGRANGIAN)
    with pytest.raises(L.HccError, match="M = 1"):
        Stepper(ctx, g2, sch2)


def test_indicator_on_a_moving_grid_leaves_the_step_unchanged(ctx):
    """The indicator's feature gather on a moving grid runs on the stage geometry (weight 0 beyond R, as the moments):
    Lagrangian Euler with a network that never fires and with training-data collection on is bit-exact with the
    oracle's plain moving step, five CFL steps."""
    from hcc_hip import MLP
    from oracle import mlp as NN
    wl = _euler_moving("epd1_u1")
    g, st, o = build_pair(ctx, wl)
    p = np.zeros(NN.P, np.float32)
    p[-1] = -100.0                      # logit -100 everywhere: never above the threshold
    net = MLP(ctx, p)
    st.set_indicator(net, 0.0)
    st.collect(True)
    _run(g, st, o, 5, cfl=wl["cfl"])
    assert st.indicator_count() == 0
    assert np.sum(st.mood_flags()) > 0
    net.close()
