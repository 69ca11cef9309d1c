"""Parity of the HIP path (through the C ABI) with the CPU oracle on identical seeded inputs.

Bar: neighbour lists, bins, orders and MOOD flags bit-exact; states, weights, caches and dt identical to the last
bit as well (the kernels keep the reference's operation order, -ffp-contract=off, correctly rounded FP64 division
and sqrt). The stated floating-point tolerance for states, 1e-12 relative L-infinity, is asserted alongside as the
bound that must never be exceeded.
"""
import numpy as np
import pytest

from hcc_hip import Grid, Stepper
from hcc_hip import _lib as L
from hcc_hip import workloads as W
from helpers import build_pair, first_diff, rel_linf

pytestmark = pytest.mark.gpu

STATE_RTOL = 1e-12   # relative L-infinity bound on states (the stated tolerance)


def _cases():
    return {
        "adv2d_n64": lambda: W.advection_2d(64, 0),
        "euler2d_n64": lambda: W.euler_2d(n=64, seed=0),
        "euler2d_n48_s3": lambda: W.euler_2d(n=48, seed=3),
        "euler3d_n12": lambda: W.euler_3d(n=12),
    }


@pytest.mark.parametrize("case", list(_cases()))
def test_neighbour_search_bit_exact(ctx, case):
    wl = _cases()[case]()
    g, st, o = build_pair(ctx, wl)
    r_g, i_g, w_g, d_g = g.neighbors()
    r_o, i_o, w_o, d_o = o.neighbors()
    assert np.array_equal(r_g, r_o), "slot ranges differ"
    assert np.array_equal(i_g, i_o), f"neighbour indices differ: {first_diff(i_g, i_o)}"
    assert np.array_equal(w_g, w_o), f"weights differ: {first_diff(w_g, w_o)}"
    assert np.array_equal(d_g, d_o), f"distances differ: {first_diff(d_g, d_o)}"
    bo_g, cs_g, cp_g, dims_g = g.bins()
    bo_o, cs_o, cp_o, dims_o = o.bins()
    assert np.array_equal(dims_g, dims_o)
    assert np.array_equal(bo_g, bo_o) and np.array_equal(cs_g, cs_o) and np.array_equal(cp_g, cp_o)
    assert np.array_equal(g.positions(), wl["cloud"]["pos"])   # generated positions are already wrapped
    assert g.S == len(i_o) and g.max_nb == int(np.max(np.diff(r_o)))


@pytest.mark.parametrize("case", list(_cases()))
def test_geometry_caches_bit_exact(ctx, case):
    wl = _cases()[case]()
    g, st, o = build_pair(ctx, wl)
    ig, og, idv, od, dw = st.moments()
    ig_o, og_o, idv_o, od_o, dw_o = o.moments()
    assert np.array_equal(og, og_o) and np.array_equal(od, od_o)
    assert np.array_equal(ig, ig_o), f"reconstruction inverse moments differ: {first_diff(ig, ig_o)}"
    assert np.array_equal(idv[:, 1:, :], idv_o[:, 1:, :]), f"divergence rows differ: {first_diff(idv[:, 1:], idv_o[:, 1:])}"
    assert np.array_equal(dw, dw_o), f"dweights differ: {first_diff(dw, dw_o)}"


@pytest.mark.parametrize("case", list(_cases()))
def test_rhs_and_cfl_bit_exact(ctx, case):
    wl = _cases()[case]()
    g, st, o = build_pair(ctx, wl)
    K, G = st.rhs(wl["U0"])
    K_o, G_o = o.rhs(wl["U0"])
    assert np.array_equal(G, G_o), f"limited gradients differ: {first_diff(G, G_o)}"
    assert np.array_equal(K, K_o), f"K_E differs: {first_diff(K, K_o)}"
    dt, dt_o = st.time_step(), o.time_step()
    assert dt == dt_o, (dt, dt_o)


def _run_steps(st, o, cfl, nsteps, check_each=True):
    t = 0.0
    log = []
    for k in range(nsteps):
        dt_g, dt_o = st.time_step(), o.time_step()
        assert dt_g == dt_o, f"step {k}: CFL dt {dt_g} vs {dt_o}"
        dt = cfl * dt_g
        st.step(t, dt)
        o.step(t, dt)
        t += dt
        if check_each or k == nsteps - 1:
            U, U_o = st.download(), o.state()
            f, f_o = st.mood_flags(), o.mood_flags()
            assert np.array_equal(st.orders(), o.orders()), f"step {k}: particle orders differ"
            assert np.array_equal(f, f_o), f"step {k}: MOOD flags differ ({int(f.sum())} vs {int(f_o.sum())})"
            assert st.inadmissible() == o.inadmissible(), f"step {k}: inadmissible count differs"
            err = rel_linf(U, U_o)
            assert err <= STATE_RTOL, f"step {k}: states differ, rel Linf {err}"
            assert np.array_equal(U, U_o), f"step {k}: states not bit-identical: {first_diff(U, U_o)}"
            log.append((k, int(f.sum()), st.mood_iterations()))
    return log


def test_advection_config0_full_run(ctx):
    """2-D advection on the 64^2 jittered periodic cloud, MUSCL2 + VK + MOOD, RK3, CFL 0.5, to t = 0.5."""
    wl = W.advection_2d(64, 0)
    g, st, o = build_pair(ctx, wl)
    t, k = 0.0, 0
    while t < wl["t_max"]:
        dt_g, dt_o = st.time_step(), o.time_step()
        assert dt_g == dt_o
        dt = min(wl["cfl"] * dt_g, wl["t_max"] - t)
        if dt <= 1e-12:
            break
        st.step(t, dt)
        o.step(t, dt)
        t += dt
        k += 1
        assert np.array_equal(st.mood_flags(), o.mood_flags()), f"step {k}"
        assert np.array_equal(st.orders(), o.orders()), f"step {k}"
    U, U_o = st.download(), o.state()
    assert rel_linf(U, U_o) <= STATE_RTOL
    assert np.array_equal(U, U_o), first_diff(U, U_o)
    assert k > 50


@pytest.mark.parametrize("n,seed,steps", [(64, 0, 20), (48, 3, 20), (96, 7, 6)])
def test_euler2d_quadrants_steps(ctx, n, seed, steps):
    """2-D Euler four-quadrant problem with outflow walls, MOOD firing: every step bit-exact."""
    wl = W.euler_2d(n=n, seed=seed)
    g, st, o = build_pair(ctx, wl)
    log = _run_steps(st, o, wl["cfl"], steps)
    assert sum(f for _, f, _ in log) > 0, "MOOD never fired: the case does not exercise the fixed point"


MOOD_VARIANTS = {
    # strategy, criterion, plateau, Euler substeps checked (else the stored candidate)
    "epd2_u1_sub": (L.HCC_STRAT_EPD2, L.HCC_CRIT_U1, 0.0, True),
    "strict_epd0_u1_sub": (L.HCC_STRAT_STRICT_EPD0, L.HCC_CRIT_U1, 0.0, True),
    "epd1_u2_plateau_sub": (L.HCC_STRAT_EPD1, L.HCC_CRIT_U2, 1e-2, True),
    "epd0_u2_stored": (L.HCC_STRAT_EPD0, L.HCC_CRIT_U2, 0.0, False),
    "strict_epd0_u1_stored": (L.HCC_STRAT_STRICT_EPD0, L.HCC_CRIT_U1, 0.0, False),
    "epd2_u2_plateau_stored": (L.HCC_STRAT_EPD2, L.HCC_CRIT_U2, 1e-3, False),
    "strict_epd0_u2_plateau_sub": (L.HCC_STRAT_STRICT_EPD0, L.HCC_CRIT_U2, 1e-3, True),
}


@pytest.mark.parametrize("variant", list(MOOD_VARIANTS))
def test_mood_variants_euler2d(ctx, variant):
    """MOOD strategies EPD2 / StrictEPD0 and criterion MOODu2 (MOOD.jl:7-64, 140-270) on the four-quadrant Euler
    problem: the unfused fixed point (forcing, two-hop recompute, global range) bit-exact step for step."""
    strat, crit, plateau, sub = MOOD_VARIANTS[variant]
    wl = W.euler_2d(n=48, seed=3)
    sch = wl["scheme"]
    wl["scheme"] = W.scheme(4, L.HCC_PDE_EULER, delta=sch["delta"], bcs=sch["bcs"], strategy=strat, criterion=crit,
                            plateau=plateau, mood_substeps=sub)
    g, st, o = build_pair(ctx, wl)
    log = _run_steps(st, o, wl["cfl"], 8)
    assert sum(f for _, f, _ in log) > 0, "MOOD never fired"


FLUX_VARIANTS = {
    # streamed divergence, numerical flux, MOOD strategy, criterion, Euler substeps checked
    "streamed_rusanov_epd1_u1": (True, L.HCC_FLUX_RUSANOV, L.HCC_STRAT_EPD1, L.HCC_CRIT_U1, True),
    "buffered_aligned_epd1_u1": (False, L.HCC_FLUX_ALIGNED_RUSANOV, L.HCC_STRAT_EPD1, L.HCC_CRIT_U1, True),
    "streamed_aligned_epd0_u1": (True, L.HCC_FLUX_ALIGNED_RUSANOV, L.HCC_STRAT_EPD0, L.HCC_CRIT_U1, True),
    "buffered_aligned_epd2_u1": (False, L.HCC_FLUX_ALIGNED_RUSANOV, L.HCC_STRAT_EPD2, L.HCC_CRIT_U1, True),
    "streamed_aligned_strict_epd0_u2_stored": (True, L.HCC_FLUX_ALIGNED_RUSANOV, L.HCC_STRAT_STRICT_EPD0,
                                               L.HCC_CRIT_U2, False),
}


@pytest.mark.parametrize("variant", list(FLUX_VARIANTS))
def test_muscl_flux_variants_euler2d(ctx, variant):
    """MUSCL(...; streamed = true) and AlignedRusanovFlux (MUSCL.jl:338-339, 407-461) on the four-quadrant Euler
    problem with MOOD dropping orders (the aligned sorting applies on pairs below the full order): bit-exact."""
    streamed, flux, strat, crit, sub = FLUX_VARIANTS[variant]
    wl = W.euler_2d(n=48, seed=3)
    sch = wl["scheme"]
    wl["scheme"] = W.scheme(4, L.HCC_PDE_EULER, delta=sch["delta"], bcs=sch["bcs"], strategy=strat, criterion=crit,
                            plateau=1e-3 if crit == L.HCC_CRIT_U2 else 0.0, mood_substeps=sub, streamed=streamed,
                            flux=flux)
    g, st, o = build_pair(ctx, wl)
    K, _ = st.rhs(wl["U0"])
    K_o, _ = o.rhs(wl["U0"])
    assert np.array_equal(K, K_o), f"K_E differs: {first_diff(K, K_o)}"
    log = _run_steps(st, o, wl["cfl"], 8)
    assert sum(f for _, f, _ in log) > 0, "MOOD never fired"


def test_streamed_aligned_advection_unlimited(ctx):
    """Streamed divergence with AlignedRusanovFlux on the periodic advection cloud, unlimited reconstruction of a
    square pulse (MOOD keeps it bounded): every step bit-exact."""
    wl = W.advection_2d(64, 0)
    sch = wl["scheme"]
    x = wl["cloud"]["pos"]
    wl["U0"] = np.all((x > 0.3) & (x < 0.6), axis=1).astype(float)[:, None]
    wl["scheme"] = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=sch["adv_speed"], delta=sch["delta"], limiter=L.HCC_LIM_NONE,
                            streamed=True, flux=L.HCC_FLUX_ALIGNED_RUSANOV)
    g, st, o = build_pair(ctx, wl)
    log = _run_steps(st, o, wl["cfl"], 10)
    assert sum(f for _, f, _ in log) > 0, "MOOD never fired"


def test_mood_epd2_u2_advection(ctx):
    """EPD2 + MOODu2 with the plateau test on the periodic advection cloud (linear PDE, one component): an unlimited
    MUSCL reconstruction of a square pulse on a sine background, so that MOOD is what keeps the update bounded."""
    wl = W.advection_2d(64, 0)
    sch = wl["scheme"]
    x = wl["cloud"]["pos"]
    wl["U0"] = np.all((x > 0.3) & (x < 0.6), axis=1).astype(float)[:, None] + 0.05 * np.sin(2 * np.pi * x[:, :1])
    wl["scheme"] = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=sch["adv_speed"], delta=sch["delta"], limiter=L.HCC_LIM_NONE,
                            strategy=L.HCC_STRAT_EPD2, criterion=L.HCC_CRIT_U2, plateau=1e-3)
    g, st, o = build_pair(ctx, wl)
    log = _run_steps(st, o, wl["cfl"], 10)
    assert sum(f for _, f, _ in log) > 0, "MOOD never fired"


def test_euler3d_pulse_steps(ctx):
    wl = W.euler_3d(n=12)
    g, st, o = build_pair(ctx, wl)
    _run_steps(st, o, wl["cfl"], 5)


def test_euler2d_256_last_step(ctx):
    """A larger cloud (about 66k nodes): states, orders and flags after 3 steps."""
    wl = W.euler_2d(n=256, seed=1)
    g, st, o = build_pair(ctx, wl)
    _run_steps(st, o, wl["cfl"], 3, check_each=False)


def test_euler2d_full_size_step_bit_exact(ctx):
    """The benchmark workload itself (h = 1/1000, 1 013 475 nodes, LDS-staged passes): three GPU steps evolve the
    four-quadrant state until MOOD fires, then one more step from that state on both sides must agree bit for bit
    (dt, state, orders, flags, inadmissible count). The oracle runs a single step, so the test stays in seconds."""
    wl = W.euler_2d(h=1e-3, seed=0)
    c = wl["cloud"]
    g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
    st = Stepper(ctx, g, wl["scheme"])
    st.upload(wl["U0"])
    t = 0.0
    for _ in range(3):
        dt = wl["cfl"] * st.time_step()
        st.step(t, dt)
        t += dt
    U = st.download()
    from oracle import oracle as O
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], wl["scheme"], threads=16)
    o.set_state(U)
    dt_g, dt_o = st.time_step(), o.time_step()
    assert dt_g == dt_o
    dt = wl["cfl"] * dt_g
    st.step(t, dt)
    o.step(t, dt)
    f, f_o = st.mood_flags(), o.mood_flags()
    assert f.sum() > 0, "MOOD did not fire at full size"
    assert np.array_equal(f, f_o)
    assert np.array_equal(st.orders(), o.orders())
    assert st.inadmissible() == o.inadmissible()
    U1, U1_o = st.download(), o.state()
    assert rel_linf(U1, U1_o) <= STATE_RTOL
    assert np.array_equal(U1, U1_o), first_diff(U1, U1_o)
    rho, p = U1[:, 0], (1.4 - 1.0) * (U1[:, 3] - 0.5 * (U1[:, 1] ** 2 + U1[:, 2] ** 2) / U1[:, 0])
    assert np.all(rho > 0) and np.all(p > 0)


FULL_STEPS = 200       # CFL steps of the 1 013 475-node run compared with the oracle
FULL_CHECK_EVERY = 10  # states, orders, flags and the inadmissible count compared every this many steps


def test_euler2d_full_size_cfl_run_bit_exact(ctx):
    """The benchmark run itself: the 1 013 475-node four-quadrant problem (h = 1/1000, seed 0) advanced by
    hcc_step_cfl -- the device-side CFL estimate fused into the stage-1 gather, the path bench.py times -- for
    FULL_STEPS steps from t = 0 beside the oracle's driver loop (dt = CFL x get_time_step, then the step;
    test/time_integration.jl:107-111). dt is compared every step; states, orders, MOOD flags and the accepted
    inadmissible count every FULL_CHECK_EVERY steps and at the end, bit for bit (the 1e-12 relative L-infinity bound
    asserted too). The reference's own device bar: test/gpu_tests.jl:38-43."""
    import os
    wl = W.euler_2d(h=1e-3, seed=0)
    c = wl["cloud"]
    g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
    st = Stepper(ctx, g, wl["scheme"])
    st.upload(wl["U0"])
    from oracle import oracle as O
    threads = max(1, min(len(os.sched_getaffinity(0)), 16))
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], wl["scheme"], threads=threads)
    o.set_state(wl["U0"])
    cfl, t, fired = wl["cfl"], 0.0, 0
    for k in range(1, FULL_STEPS + 1):
        st.step_cfl(cfl, float("inf"))
        dt = cfl * o.time_step()
        assert st.last_dt() == dt, f"dt differs at step {k}: {st.last_dt()!r} vs {dt!r}"
        o.step(t, dt)
        t += dt
        if k % FULL_CHECK_EVERY == 0 or k == FULL_STEPS:
            f, f_o = st.mood_flags(), o.mood_flags()
            assert np.array_equal(f, f_o), f"MOOD flags differ at step {k}: {first_diff(f, f_o)}"
            assert np.array_equal(st.orders(), o.orders()), f"orders differ at step {k}"
            assert st.inadmissible() == o.inadmissible(), f"inadmissible count differs at step {k}"
            U, U_o = st.download(), o.state()
            assert rel_linf(U, U_o) <= STATE_RTOL, f"step {k}: rel Linf {rel_linf(U, U_o):.3e}"
            assert np.array_equal(U, U_o), f"states differ at step {k}: {first_diff(U, U_o)}"
            fired += int(f.sum())
            print(f"step {k}: t = {t:.6e}, dt = {dt:.6e}, MOOD flags {int(f.sum())}, bit-exact", flush=True)
    assert fired > 0, "MOOD never fired on the checked steps"


def test_euler2d_full_size_dense_state_bit_exact(ctx):
    """The bench's dense state at full size (the four quadrants times 1 + 0.01 sin(6 pi x) sin(10 pi y) on the
    1 013 475-node cloud): no constant neighbourhood, and the first MOOD check of stage 1 flags most particles, so
    from the second step on the first fixed-point iteration runs in its dense form (k_mark_dense, k_flux_marked).
    Five hcc_step_cfl steps beside the oracle: dt every step, states, orders, flags and the inadmissible count bit for
    bit after each."""
    import os
    wl = W.euler_2d(h=1e-3, seed=0)
    c = wl["cloud"]
    x = c["pos"]
    U0 = wl["U0"].copy()
    fac = 1.0 + 0.01 * np.sin(2 * np.pi * 3 * x[:, 0]) * np.sin(2 * np.pi * 5 * x[:, 1])
    U0[:, 0] *= fac
    U0[:, 3] *= fac
    g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
    st = Stepper(ctx, g, wl["scheme"])
    assert st.tuning()["dense"] == 64
    st.upload(U0)
    from oracle import oracle as O
    threads = max(1, min(len(os.sched_getaffinity(0)), 16))
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], wl["scheme"], threads=threads)
    o.set_state(U0)
    cfl, t = wl["cfl"], 0.0
    for k in range(1, 6):
        st.step_cfl(cfl, float("inf"))
        dt = cfl * o.time_step()
        assert st.last_dt() == dt, f"dt differs at step {k}: {st.last_dt()!r} vs {dt!r}"
        o.step(t, dt)
        t += dt
        f, f_o = st.mood_flags(), o.mood_flags()
        assert np.array_equal(f, f_o), f"MOOD flags differ at step {k}: {first_diff(f, f_o)}"
        assert int(f.sum()) * 64 >= g.N, f"step {k}: {int(f.sum())} flags do not select the dense form"
        assert np.array_equal(st.orders(), o.orders()), f"orders differ at step {k}"
        assert st.inadmissible() == o.inadmissible(), f"inadmissible count differs at step {k}"
        U, U_o = st.download(), o.state()
        assert np.array_equal(U, U_o), f"states differ at step {k}: {first_diff(U, U_o)}"
        print(f"dense state step {k}: dt = {dt:.6e}, MOOD flags {int(f.sum())}, bit-exact", flush=True)


def test_upload_breaks_the_step_chain(ctx):
    """Consecutive CFL steps form a chain in which a particle whose K's stay exactly zero keeps the stage records the
    last step wrote (its predictor stores nothing). An upload in the middle of a chain must end it: after five steps
    of the four-quadrant state, the same state with density and energy scaled by 1.1 (the same constant regions, other
    values in every record) is uploaded, and the steps after it match the oracle started from it, bit for bit."""
    from oracle import oracle as O
    wl = W.euler_2d(n=64, seed=0)
    c = wl["cloud"]
    g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
    st = Stepper(ctx, g, wl["scheme"])
    st.upload(wl["U0"])
    for _ in range(5):   # the chain is on from the third step
        st.step_cfl(wl["cfl"])
    U_b = wl["U0"].copy()
    U_b[:, 0] *= 1.1
    U_b[:, 3] *= 1.1
    st.upload(U_b)
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], wl["scheme"], threads=8)
    o.set_state(U_b)
    t = 0.0
    for k in range(1, 7):
        st.step_cfl(wl["cfl"])
        dt = wl["cfl"] * o.time_step()
        assert st.last_dt() == dt, f"step {k} after the upload"
        o.step(t, dt)
        t += dt
        U, U_o = st.download(), o.state()
        assert np.array_equal(U, U_o), f"step {k} after the upload: {first_diff(U, U_o)}"
        assert np.array_equal(st.orders(), o.orders()), f"step {k} after the upload"


def test_step_totals_match_per_step_values(ctx):
    """hcc_step returns once its work is queued; the per-step getters and the summed totals see the same steps."""
    wl = W.euler_2d(n=64, seed=2)
    c = wl["cloud"]
    g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
    st = Stepper(ctx, g, wl["scheme"])
    st.upload(wl["U0"])
    st.totals(reset=True)
    t, iters, ms = 0.0, 0, 0.0
    for _ in range(6):
        dt = wl["cfl"] * st.time_step()
        st.step(t, dt)
        t += dt
        iters += st.mood_iterations()
        ms += st.timing()[0]
    n, it, tot_ms, g_ms, f_ms = st.totals(reset=True)
    assert n == 6 and it == iters
    assert abs(tot_ms - ms) <= 1e-6 * max(ms, 1.0)
    assert 0 < g_ms < tot_ms and 0 < f_ms < tot_ms
    assert st.totals(reset=False)[0] == 0


def test_sampled_pass_timing(ctx):
    """hcc_stepper_set_pass_timing: the pass events of every k-th step (k = 3: steps 0, 3, 6 of 7), none with k = 0;
    the steps that were timed are counted with the totals, and the states do not depend on the timing."""
    wl = W.euler_2d(n=64, seed=2)
    c = wl["cloud"]
    g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
    a, b = Stepper(ctx, g, wl["scheme"]), Stepper(ctx, g, wl["scheme"])
    a.upload(wl["U0"])
    b.upload(wl["U0"])
    a.set_pass_timing(3)
    b.set_pass_timing(0)
    a.totals(reset=True)
    b.totals(reset=True)
    for k in range(7):
        a.step_cfl(wl["cfl"])
        b.step_cfl(wl["cfl"])
        assert np.array_equal(a.download(), b.download()), f"step {k}"
    assert a.timed_passes() == 3 and b.timed_passes() == 0
    n, _, tot_ms, g_ms, f_ms = a.totals(reset=True)
    assert n == 7 and 0 < g_ms < tot_ms and 0 < f_ms < tot_ms
    assert a.timed_passes() == 0, "cleared with the totals"
    n, _, _, g_ms, f_ms = b.totals(reset=True)
    assert n == 7 and g_ms == 0 and f_ms == 0
    for bad in (-1,):
        with pytest.raises(Exception):
            a.set_pass_timing(bad)


def test_step_cfl_matches_host_driver_loop(ctx):
    """hcc_step_cfl (dt = min(CFL * CFL step, cap) on the device) gives the same dt and states, bit for bit, as the
    driver loop with the host-side dt (time_step, then step); the cap binds in the last step."""
    wl = W.euler_2d(n=64, seed=4)
    c = wl["cloud"]
    g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
    a, b = Stepper(ctx, g, wl["scheme"]), Stepper(ctx, g, wl["scheme"])
    a.upload(wl["U0"])
    b.upload(wl["U0"])
    t = 0.0
    for k in range(6):
        cap = 1e-5 if k == 5 else float("inf")
        dt = min(wl["cfl"] * a.time_step(), cap)
        a.step(t, dt)
        b.step_cfl(wl["cfl"], cap)
        assert b.last_dt() == dt
        t += dt
        assert np.array_equal(a.mood_flags(), b.mood_flags())
    Ua, Ub = a.download(), b.download()
    assert np.array_equal(Ua, Ub), first_diff(Ua, Ub)
    assert a.inadmissible() == b.inadmissible()
