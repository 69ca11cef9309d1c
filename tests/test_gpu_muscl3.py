"""MUSCL order 3 on the device against the oracle, bit for bit: the geometry caches of the degree-2 reconstruction and
the safeguarded degree-3 divergence, one right-hand side (gradients with their degree-2 terms, K), and multi-step runs
in which MOOD drops particles through three orders, for advection in 1-D and 2-D, a two-equation system and 2-D Euler.
The cases follow the reference's own MUSCL3 GPU parity list (test/gpu_tests.jl:14-22: NoLimiter RK3, VK RK2,
StrictEPD0 + MOODu1 RK2 on a box) and its published step benchmark (benchmark/step.jl)."""
import numpy as np
import pytest

from hcc_hip import Grid, Stepper
from hcc_hip import _lib as L
from hcc_hip import workloads as W
from helpers import build_pair, first_diff, rel_linf

pytestmark = pytest.mark.gpu

STATE_RTOL = 1e-12


def _run_steps(st, o, nsteps, cfl=None, dt=None, check_each=True):
    """Steps both sides: with cfl, dt = cfl times the CFL step of each side (asserted equal); else the fixed dt."""
    t, log = 0.0, []
    for k in range(nsteps):
        if cfl is not None:
            dt_g, dt_o = st.time_step(), o.time_step()
            assert dt_g == dt_o, f"step {k}: CFL dt {dt_g} vs {dt_o}"
            h = cfl * dt_g
        else:
            h = dt
        st.step(t, h)
        o.step(t, h)
        t += h
        if check_each or k == nsteps - 1:
            U, U_o = st.download(), o.state()
            f, f_o = st.mood_flags(), o.mood_flags()
            assert np.array_equal(st.orders(), o.orders()), f"step {k}: particle orders differ"
            assert np.array_equal(f, f_o), f"step {k}: MOOD flags differ ({int(f.sum())} vs {int(f_o.sum())})"
            assert st.inadmissible() == o.inadmissible(), f"step {k}: inadmissible count differs"
            assert rel_linf(U, U_o) <= STATE_RTOL, f"step {k}: rel Linf {rel_linf(U, U_o)}"
            assert np.array_equal(U, U_o), f"step {k}: states not bit-identical: {first_diff(U, U_o)}"
            log.append((k, int(f.sum()), int(np.sum(st.orders() == 1)), int(np.sum(st.orders() == 2))))
    return log


def _adv(D, n, R, jitter, U0fn, periodic=True, M=1, **kw):
    h = 1.0 / n
    dom = W.domain(D, (0.0,) * D, (1.0,) * D, (periodic,) * D, h, R, jitter, 3)
    speeds = [[1.0, -0.6][:M], [0.5, 0.8][:M], [-0.7, 0.3][:M]][:D]
    sch = W.scheme(M, L.HCC_PDE_LINADV, adv_speed=speeds, order=3, **kw)
    cloud = L.cloud_generate(W.domain_desc(dom))
    U0 = U0fn(cloud["pos"])
    U0 = U0.reshape(len(U0), -1)
    return {"domain": dom, "scheme": sch, "cloud": cloud, "U0": U0}


GAUSS = lambda x: np.exp(-np.sum((x - 0.5) ** 2, axis=1) / 0.02)
BOX = lambda x: np.all((x > 0.3) & (x < 0.6), axis=1).astype(float)


@pytest.mark.parametrize("D,R,jitter", [(1, 2.5, 0.0), (1, 3.5, 0.2), (2, 2.2, 0.25), (2, 2.5, 0.0), (3, 2.0, 0.25)])
def test_geometry_caches_bit_exact(ctx, D, R, jitter):
    """k_moments3: the degree-2 reconstruction inverse, the divergence inverse of each stencil's degree (3, or 2 where
    the safeguard lowers it; zero-padded), its derivative rows and the per-slot derivative weights."""
    wl = _adv(D, {1: 64, 2: 48, 3: 12}[D], R, jitter, GAUSS)
    g, st, o = build_pair(ctx, wl)
    ig, og, idv, od, dw = st.moments()
    ig_o, og_o, idv_o, od_o, dw_o = o.moments()
    assert np.array_equal(og, og_o) and np.array_equal(od, od_o)
    assert np.array_equal(ig, ig_o), first_diff(ig, ig_o)
    assert np.array_equal(idv[:, 1:D + 1], idv_o[:, 1:D + 1]), first_diff(idv[:, 1:D + 1], idv_o[:, 1:D + 1])
    assert np.array_equal(dw, dw_o), first_diff(dw, dw_o)


RHS_CASES = {
    "adv2d_quadratic_none": (2, L.HCC_LIM_NONE, L.HCC_MODE_HARD, lambda x: (x @ [1.7, -0.6]) + 0.5 * x[:, 0] ** 2 - x[:, 0] * x[:, 1]),
    "adv2d_box_vk_hard": (2, L.HCC_LIM_VK, L.HCC_MODE_HARD, BOX),
    "adv2d_box_bj_soft": (2, L.HCC_LIM_BJ, L.HCC_MODE_SOFT, BOX),
    "adv2d_gauss_vk_soft": (2, L.HCC_LIM_VK, L.HCC_MODE_SOFT, GAUSS),
    "adv1d_box_vk_hard": (1, L.HCC_LIM_VK, L.HCC_MODE_HARD, BOX),
}


@pytest.mark.parametrize("case", list(RHS_CASES))
def test_rhs_bit_exact(ctx, case):
    D, lim, mode, f = RHS_CASES[case]
    wl = _adv(D, 40 if D == 2 else 80, 2.5, 0.25, f, limiter=lim, mode=mode)
    g, st, o = build_pair(ctx, wl)
    K, G = st.rhs(wl["U0"])
    K_o, G_o = o.rhs(wl["U0"])
    assert np.array_equal(G, G_o), first_diff(G, G_o)
    assert np.array_equal(K, K_o), first_diff(K, K_o)


def test_rhs_two_equations_and_euler(ctx):
    wl = _adv(2, 40, 2.5, 0.25, lambda x: np.column_stack([GAUSS(x), BOX(x)]), M=2)
    g, st, o = build_pair(ctx, wl)
    K, G = st.rhs(wl["U0"])
    K_o, G_o = o.rhs(wl["U0"])
    assert np.array_equal(G, G_o) and np.array_equal(K, K_o), first_diff(K, K_o)
    e = W.euler_2d(n=40, seed=2)
    e["scheme"] = dict(e["scheme"], muscl_order=3)
    g, st, o = build_pair(ctx, e)
    K, G = st.rhs(e["U0"])
    K_o, G_o = o.rhs(e["U0"])
    assert np.array_equal(G, G_o), first_diff(G, G_o)
    assert np.array_equal(K, K_o), first_diff(K, K_o)


STEP_CASES = {
    # (D, n, R, jitter, IC, limiter, strategy, criterion, tableau, steps)
    "1d_none_rk3_gauss": (1, 100, 3.5, 0.0, GAUSS, L.HCC_LIM_NONE, L.HCC_STRAT_NONE, L.HCC_CRIT_NONE, "RK3_SSP", 12),
    "1d_vk_rk2_box": (1, 100, 2.5, 0.2, BOX, L.HCC_LIM_VK, L.HCC_STRAT_NONE, L.HCC_CRIT_NONE, "RK2_SSP", 12),
    "2d_vk_epd1_u1_rk3_gauss": (2, 48, 2.5, 0.0, GAUSS, L.HCC_LIM_VK, L.HCC_STRAT_EPD1, L.HCC_CRIT_U1, "RK3_SSP", 10),
    "2d_none_epd1_u1_rk3_box": (2, 48, 2.5, 0.25, BOX, L.HCC_LIM_NONE, L.HCC_STRAT_EPD1, L.HCC_CRIT_U1, "RK3_SSP", 10),
    "2d_none_sepd0_u1_rk2_box": (2, 48, 2.5, 0.25, BOX, L.HCC_LIM_NONE, L.HCC_STRAT_STRICT_EPD0, L.HCC_CRIT_U1,
                                 "RK2_SSP", 10),
    "2d_none_epd0_u1_rk3_box": (2, 48, 2.2, 0.25, BOX, L.HCC_LIM_NONE, L.HCC_STRAT_EPD0, L.HCC_CRIT_U1, "RK3_SSP", 8),
    "2d_none_epd2_u1_rk3_box": (2, 48, 2.5, 0.25, BOX, L.HCC_LIM_NONE, L.HCC_STRAT_EPD2, L.HCC_CRIT_U1, "RK3_SSP", 8),
    "2d_bj_soft_pad_rk3_box": (2, 48, 2.5, 0.25, BOX, L.HCC_LIM_BJ, L.HCC_STRAT_EPD1, L.HCC_CRIT_PAD, "RK3_SSP", 6),
    # MOODu2 with its smooth-extremum override (gpu_tests.jl:20: MUSCL3 MOODu2 EPD1 RK3 box)
    "2d_none_epd1_u2_rk3_box": (2, 48, 2.5, 0.25, BOX, L.HCC_LIM_NONE, L.HCC_STRAT_EPD1, L.HCC_CRIT_U2, "RK3_SSP", 8),
    "2d_none_epd1_u2plateau_rk3_gauss": (2, 48, 2.5, 0.25, GAUSS, L.HCC_LIM_NONE, L.HCC_STRAT_EPD1, L.HCC_CRIT_U2,
                                         "RK3_SSP", 8),
    "2d_vk_sepd0_u2_rk2_box": (2, 48, 2.5, 0.25, BOX, L.HCC_LIM_VK, L.HCC_STRAT_STRICT_EPD0, L.HCC_CRIT_U2, "RK2_SSP", 6),
}


@pytest.mark.parametrize("case", list(STEP_CASES))
def test_steps_bit_exact(ctx, case):
    """Multi-step runs: dt, states, orders (three levels), MOOD flags and the inadmissible count after every step."""
    D, n, R, jit, f, lim, strat, crit, tab, steps = STEP_CASES[case]
    mode = L.HCC_MODE_SOFT if "soft" in case else L.HCC_MODE_HARD
    plateau = 1e-3 if "plateau" in case else 0.0
    wl = _adv(D, n, R, jit, f, limiter=lim, mode=mode, strategy=strat, criterion=crit, delta=1e-6, tableau_name=tab,
              plateau=plateau)
    g, st, o = build_pair(ctx, wl)
    log = _run_steps(st, o, steps, cfl=0.4)
    if crit in (L.HCC_CRIT_U1, L.HCC_CRIT_U2) and "box" in case:   # the box makes MOOD drop orders
        assert sum(fl for _, fl, _, _ in log) > 0, "MOOD never fired"
        assert max(o1 + o2 for _, _, o1, o2 in log) > 0


def test_only_mood_reaches_order_one(ctx):
    """OnlyMOOD: every checkpoint drops every interior particle one order at a time, 3 -> 2 -> 1."""
    wl = _adv(2, 32, 2.5, 0.25, GAUSS, criterion=L.HCC_CRIT_ONLY)
    g, st, o = build_pair(ctx, wl)
    _run_steps(st, o, 2, cfl=0.4)
    inner = wl["cloud"]["is_boundary"] == 0
    assert np.all(st.orders()[inner] == 1)


def test_euler2d_quadrants_muscl3(ctx):
    """2-D Euler four-quadrant problem with outflow walls, MUSCL3 + VK(hard) + MOOD(EPD1, u1), RK3, CFL 0.4: every
    step bit-exact while MOOD drops particles to orders 2 and 1."""
    wl = W.euler_2d(n=48, seed=3)
    wl["scheme"] = dict(wl["scheme"], muscl_order=3)
    g, st, o = build_pair(ctx, wl)
    log = _run_steps(st, o, 12, cfl=0.4)
    assert sum(fl for _, fl, _, _ in log) > 0
    assert max(o2 for _, _, _, o2 in log) > 0 and max(o1 for _, _, o1, _ in log) > 0


def test_step_cfl_matches_host_loop(ctx):
    """hcc_step_cfl (the CFL estimate queued with the step, dt on the device) equals time_step + step."""
    wl = _adv(2, 40, 2.5, 0.25, BOX, limiter=L.HCC_LIM_VK, strategy=L.HCC_STRAT_EPD1, criterion=L.HCC_CRIT_U1,
              delta=1e-6)
    c = wl["cloud"]
    g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
    a, b = Stepper(ctx, g, wl["scheme"]), Stepper(ctx, g, wl["scheme"])
    a.upload(wl["U0"])
    b.upload(wl["U0"])
    t = 0.0
    for _ in range(6):
        a.step_cfl(0.4)
        dt = 0.4 * b.time_step()
        b.step(t, dt)
        t += dt
    assert np.array_equal(a.download(), b.download())


@pytest.mark.parametrize("n,steps", [(500, 4), (1000, 3)])
def test_step_benchmark_workload_bit_exact(ctx, n, steps):
    """The reference's step benchmark (benchmark/step.jl, stab = both) at 250 000 and at its published 1 000 000
    particles: dt = 0.3 get_time_step, fixed; everything bit-exact after each step."""
    wl = W.step_benchmark(n=n)
    g, st, o = build_pair(ctx, wl, threads=16)
    dt_g, dt_o = st.time_step(), o.time_step()
    assert dt_g == dt_o
    log = _run_steps(st, o, steps, dt=0.3 * dt_g)
    assert len(log) == steps


def test_unsupported_muscl3_setups(ctx):
    wl = _adv(2, 24, 2.5, 0.25, GAUSS)
    c = wl["cloud"]
    g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
    for kw in ({"streamed": True}, {"flux": L.HCC_FLUX_ALIGNED_RUSANOV}, {"mover": L.HCC_MOVER_LAGRANGIAN}):
        sch = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0], [0.5]], order=3, **kw)
        with pytest.raises(L.HccError):
            Stepper(ctx, g, sch)
    sch = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0], [0.5]], order=4)
    with pytest.raises(L.HccError):
        Stepper(ctx, g, sch)


def test_u2_override_accepts_smooth_extrema(ctx):
    """On a smooth Gaussian the DMP undershoots/overshoots at the extremum are accepted by MOODu2's curvature
    override at order 3: fewer MOOD drops than MOODu1 over the same steps (both bit-exact with the oracle)."""
    drops = {}
    for crit in (L.HCC_CRIT_U1, L.HCC_CRIT_U2):
        wl = _adv(2, 48, 2.5, 0.0, lambda x: np.exp(-np.sum((x - 0.5) ** 2, axis=1) / 0.05),
                  limiter=L.HCC_LIM_NONE, strategy=L.HCC_STRAT_EPD1, criterion=crit, delta=1e-12)
        g, st, o = build_pair(ctx, wl)
        log = _run_steps(st, o, 8, cfl=0.4)
        drops[crit] = sum(fl for _, fl, _, _ in log)
    assert drops[L.HCC_CRIT_U1] > 0
    assert drops[L.HCC_CRIT_U2] < drops[L.HCC_CRIT_U1]



def _euler3d_muscl3(n, seed=1):
    """config[2]'s cube at a small size: periodic, R = 2.2 h (a mix of degree-3 and safeguarded degree-2 divergence
    fits), rest state with a Gaussian pressure pulse of width 0.05, MUSCL3 + VK(hard) + MOOD(EPD1, u1), RK3."""
    wl = W.euler_3d(n=n, seed=seed)
    wl["scheme"] = dict(wl["scheme"], muscl_order=3)
    return wl


def test_rhs_3d_advection_and_euler(ctx):
    """MUSCL3 in 3-D: the degree-2 reconstruction with its nine terms (x, y, z, x^2/2, y^2/2, z^2/2, xy, xz, yz),
    the 20-term degree-3 divergence fit where the stencil carries it (R = 2 h: a quarter of the particles have fewer
    than 30 neighbours and take the safeguarded degree 2), for a limited box and the Euler pulse."""
    wl = _adv(3, 12, 2.0, 0.25, BOX, limiter=L.HCC_LIM_VK)
    g, st, o = build_pair(ctx, wl)
    K, G = st.rhs(wl["U0"])
    K_o, G_o = o.rhs(wl["U0"])
    assert np.array_equal(G, G_o), first_diff(G, G_o)
    assert np.array_equal(K, K_o), first_diff(K, K_o)
    counts = np.diff(g.neighbors(weights=False, distances=False)[0])
    assert np.any(counts < 30) and np.any(counts >= 30)   # both divergence degrees occur (2 nb >= 3 x 20 for degree 3)
    e = _euler3d_muscl3(10)
    g, st, o = build_pair(ctx, e)
    K, G = st.rhs(e["U0"])
    K_o, G_o = o.rhs(e["U0"])
    assert np.array_equal(G, G_o), first_diff(G, G_o)
    assert np.array_equal(K, K_o), first_diff(K, K_o)


def test_steps_3d_advection_box_mood(ctx):
    """3-D advection of a box, MUSCL3 + VK(hard) + MOOD(EPD1, u1), RK3: MOOD drops particles through three orders;
    every step bit-exact."""
    wl = _adv(3, 14, 2.2, 0.25, BOX, limiter=L.HCC_LIM_VK, strategy=L.HCC_STRAT_EPD1, criterion=L.HCC_CRIT_U1,
              delta=1e-6)
    g, st, o = build_pair(ctx, wl)
    log = _run_steps(st, o, 4, cfl=0.4)
    assert sum(fl for _, fl, _, _ in log) > 0, "MOOD never fired"


def test_steps_3d_euler_pulse(ctx):
    """The 3-D Euler pulse with MUSCL3, CFL 0.3: dt, states, orders, flags and the inadmissible count bit-exact after
    every step."""
    wl = _euler3d_muscl3(12)
    g, st, o = build_pair(ctx, wl)
    _run_steps(st, o, 4, cfl=wl["cfl"])
