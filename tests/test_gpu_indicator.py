"""The learned indicator inside the step (2-D Euler): collecting training data leaves the step unchanged, the
features equal a host restatement, the labels are exactly MOOD's drops, and the two extreme networks reproduce known
schemes bit for bit: one that never fires gives the plain step, one that always fires gives OnlyMOOD's first-order
step (every particle at first order before the flux pass)."""
import numpy as np
import pytest

from hcc_hip import MLP, Stepper
from hcc_hip import _lib as L
from hcc_hip import workloads as W
from helpers import build_pair, first_diff
from oracle import mlp as NN
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _wl(n=48, seed=3, criterion=L.HCC_CRIT_U1):
    wl = W.euler_2d(n=n, seed=seed)
    h = 1.0 / n
    wl["scheme"] = W.scheme(4, L.HCC_PDE_EULER, delta=1e-4 * h * h, criterion=criterion, bcs=wl["scheme"]["bcs"])
    return wl


def _steps(st, o, n, cfl=0.4):
    t = 0.0
    for k in range(n):
        dt = cfl * o.time_step()
        st.step(t, dt)
        o.step(t, dt)
        t += dt
        U, U_o = st.download(), o.state()
        assert np.array_equal(U, U_o), f"step {k}: {first_diff(U, U_o)}"


def test_collection_leaves_the_step_unchanged_and_labels_are_moods_drops(ctx):
    wl = _wl()
    g, st, o = build_pair(ctx, wl)
    st.collect(True)
    t = 0.0
    for k in range(3):
        dt = 0.4 * o.time_step()
        st.step(t, dt)
        o.step(t, dt)
        t += dt
        assert np.array_equal(st.download(), o.state()), f"step {k}"
        drops = np.zeros(g.N, bool)
        for s in range(3):
            f, pre, post = st.samples(s)
            drops |= (pre == 2) & (post == 1)
            assert np.all(post <= pre)
        assert np.array_equal(post, st.orders())
        assert np.array_equal(drops, st.mood_flags().astype(bool)), f"step {k}: labels"
    assert drops.sum() > 0


def test_features_equal_a_host_restatement(ctx):
    """Stage 1 of the first step: the ranges, skews and gradient norms of the four components, the Mach number and
    p / E, from the stage state, the neighbour lists and the limited gradients."""
    wl = _wl()
    g, st, o = build_pair(ctx, wl)
    st.collect(True)
    dt = 0.4 * o.time_step()
    st.step(0.0, dt)
    o.step(0.0, dt)
    f_bits, pre, _ = st.samples(0)
    f = (f_bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    Y = o.stage("Y", 0)
    _, G = st.rhs(Y)                       # the limited gradients of Y_1 (orders reset to 2, as at stage 1)
    r, idx, _, _ = g.neighbors(weights=False, distances=False)
    isb = wl["cloud"]["is_boundary"].astype(bool)
    gam, h = 1.4, 1.0 / 48
    rng = np.random.default_rng(0)
    for i in rng.choice(np.flatnonzero(~isb), 200, replace=False):
        nb = np.concatenate([[i], idx[r[i]:r[i + 1]]])
        u = Y[i]
        mx, mn = Y[nb].max(0), Y[nb].min(0)
        ux, uy = u[1] / u[0], u[2] / u[0]
        p = (gam - 1) * (u[3] - 0.5 * (u[1] * ux + u[2] * uy))
        c = np.sqrt(gam * p / u[0])
        sp = np.hypot(ux, uy)
        sc = np.array([abs(u[0]), abs(u[0]) * (sp + c), abs(u[0]) * (sp + c), abs(u[3])])
        ref = np.concatenate([(mx - mn) / sc, (mx + mn - 2 * u) / sc, h * np.hypot(G[i, 0], G[i, 1]) / sc,
                              [sp / c, p / u[3]]])
        assert np.allclose(f[i, :14], ref, rtol=1e-2, atol=1e-6), (i, f[i, :14], ref)
        assert 0.0 <= f[i, 14] <= f[i, 15] + 1e-6 <= 1.0 + 1e-6


def _const_net(ctx, b3):
    p = np.zeros(NN.P, np.float32)
    p[-1] = b3
    return MLP(ctx, p)


def test_a_network_that_never_fires_gives_the_plain_step(ctx):
    wl = _wl()
    g, st, o = build_pair(ctx, wl)
    net = _const_net(ctx, -100.0)
    st.set_indicator(net, 0.0)
    _steps(st, o, 3)
    assert st.indicator_count() == 0


def test_a_network_that_always_fires_gives_the_first_order_step(ctx):
    """Every candidate (a particle at the full order whose neighbourhood is not constant) takes first order before the
    flux pass: the states of OnlyMOOD, which drops every particle after the high-order pass and recomputes it at first
    order. A particle with a constant neighbourhood has an exactly zero gradient fit (its reconstruction is its value at
    either order) and is never flagged by MOOD's DMP test, so leaving it at the full order changes nothing."""
    wl = _wl()
    g, st, _ = build_pair(ctx, wl)
    wl_o = _wl(criterion=L.HCC_CRIT_ONLY)
    c = wl_o["cloud"]
    o = O.OracleSim(wl_o["domain"], c["pos"], c["is_boundary"], c["tags"], wl_o["scheme"])
    o.set_state(wl_o["U0"])
    net = _const_net(ctx, 100.0)
    st.set_indicator(net, 0.0)
    _steps(st, o, 3)
    assert st.indicator_count() > 0 and int(st.mood_flags().sum()) == 0


def test_the_shipped_network_gains_over_plain_mood(ctx):
    """The shipped network (trained on 32 seeded four-quadrant runs at 250 k nodes) on an unseen seed at 128^2, ten
    CFL steps: fewer MOOD fixed-point iterations than the plain step, final states within 1e-3 relative L1 of plain
    MOOD (its measured gain: 57 -> 22 iterations at 3.7e-4 on seed 777)."""
    from hcc_hip import indicator as IND
    net = MLP(ctx, IND.default_params())
    r = IND.compare(ctx, net, 128, 777, 10, IND.DEFAULT_THRESHOLD)
    assert r["indicator"]["lowered_last_step"] > 0
    assert r["indicator"]["mood_iterations"] < 0.8 * r["plain"]["mood_iterations"], r
    assert r["relative_l1_deviation"] < 1e-3, r
