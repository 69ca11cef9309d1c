"""Ensembles through the C ABI: several point clouds of uneven size in one grid (hcc_grid_concat) advanced by one
stepper. Every instance must equal, bit for bit, a stepper on its own grid: states, orders, MOOD flags and CFL steps
after every step, with one dt for all (hcc_step) or each instance's own CFL step (hcc_step_cfl)."""
import numpy as np
import pytest

from hcc_hip import Ensemble, Grid, Stepper
from hcc_hip import _lib as L
from hcc_hip import workloads as W
from helpers import build_pair, first_diff
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _members(sizes, h=1.0 / 64, seed=0):
    return [W.euler_2d_box(n, h, seed + 1 + k) for k, n in enumerate(sizes)]


def _singles(ctx, wls):
    out = []
    for wl in wls:
        c = wl["cloud"]
        g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
        st = Stepper(ctx, g, wl["scheme"])
        st.upload(wl["U0"])
        out.append((g, st))
    return out


def _check(ens, st_e, singles, k):
    U_e = ens.split(st_e.download())
    o_e, f_e = ens.split(st_e.orders()), ens.split(st_e.mood_flags())
    for m, (g, st) in enumerate(singles):
        U = st.download()
        assert np.array_equal(U_e[m], U), f"step {k}, member {m}: {first_diff(U_e[m], U)}"
        assert np.array_equal(o_e[m], st.orders()), f"step {k}, member {m}: orders"
        assert np.array_equal(f_e[m], st.mood_flags()), f"step {k}, member {m}: MOOD flags"


@pytest.mark.parametrize("mode", ["cfl", "host_dt"])
def test_ensemble_members_equal_their_own_runs(ctx, mode):
    """Four four-quadrant Euler boxes of 20^2 .. 41^2 lattice nodes (one spacing, own seeds) in one grid: six steps,
    each member bit-exact with a stepper on its own grid; per-member CFL steps (mode cfl) or one shared dt."""
    wls = _members([20, 33, 41, 27])
    singles = _singles(ctx, wls)
    ens = Ensemble(ctx, [g for g, _ in singles])
    assert list(ens.sizes) == [wl["cloud"]["pos"].shape[0] for wl in wls]
    assert all(o % 256 == 0 for o in ens.offsets)
    st_e = Stepper(ctx, ens, wls[0]["scheme"])
    st_e.upload(ens.pack([wl["U0"] for wl in wls]))
    t = 0.0
    for k in range(6):
        dts = st_e.time_steps()
        assert np.array_equal(dts, [st.time_step() for _, st in singles]), f"step {k}: CFL steps"
        if mode == "cfl":
            st_e.step_cfl(0.4)
            for _, st in singles:
                st.step_cfl(0.4)
        else:
            dt = 0.4 * float(dts.min())
            st_e.step(t, dt)
            for _, st in singles:
                st.step(t, dt)
            t += dt
        _check(ens, st_e, singles, k)
    assert sum(int(st.mood_flags().sum()) for _, st in singles) > 0   # MOOD fired in the members


def test_ensemble_member_matches_the_oracle(ctx):
    """One member of an ensemble against the oracle directly (its own CFL step, three steps)."""
    wls = _members([24, 30])
    singles = _singles(ctx, wls)
    ens = Ensemble(ctx, [g for g, _ in singles])
    st_e = Stepper(ctx, ens, wls[0]["scheme"])
    st_e.upload(ens.pack([wl["U0"] for wl in wls]))
    wl = wls[1]
    c = wl["cloud"]
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], wl["scheme"])
    o.set_state(wl["U0"])
    t = 0.0
    for k in range(3):
        dt = 0.4 * o.time_step()
        assert st_e.time_steps()[1] == o.time_step()
        st_e.step_cfl(0.4)
        o.step(t, dt)
        t += dt
        U = ens.split(st_e.download())[1]
        assert np.array_equal(U, o.state()), f"step {k}: {first_diff(U, o.state())}"
        assert np.array_equal(ens.split(st_e.mood_flags())[1], o.mood_flags())


def test_ensemble_rejects_mismatched_members(ctx):
    """Members must share D, R, alpha, spacing and wrap lengths: periodic boxes of different sizes (different wrap
    lengths) and different spacings are refused."""
    a = W.advection_2d(16, 0)
    b = W.advection_2d(20, 0)
    gs = []
    for wl in (a, b):
        c = wl["cloud"]
        gs.append(Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"]))
    with pytest.raises(L.HccError, match="differ"):
        Ensemble(ctx, gs)


def test_ensemble_with_the_indicator(ctx):
    """The batched form with the learned indicator (the shipped network) on six uneven four-quadrant members: every
    member bit-exact, step for step, with a stepper on its own grid running the same network (a particle's logit
    depends on its own features only), and the batched-plus-indicator states within 1e-3 relative L1 of the plain
    batched MOOD states after six CFL steps."""
    from hcc_hip import MLP
    from hcc_hip import indicator as IND
    net = MLP(ctx, IND.default_params())
    wls = _members([60, 33, 81, 45, 24, 70], h=1.0 / 200, seed=40)
    singles = _singles(ctx, wls)
    for _, st in singles:
        st.set_indicator(net, IND.DEFAULT_THRESHOLD)
    ens = Ensemble(ctx, [g for g, _ in singles])
    U0 = ens.pack([wl["U0"] for wl in wls])
    st_i, st_p = Stepper(ctx, ens, wls[0]["scheme"]), Stepper(ctx, ens, wls[0]["scheme"])
    st_i.upload(U0)
    st_p.upload(U0)
    st_i.set_indicator(net, IND.DEFAULT_THRESHOLD)
    lowered = 0
    for k in range(6):
        st_i.step_cfl(0.4)
        st_p.step_cfl(0.4)
        for _, st in singles:
            st.step_cfl(0.4)
        _check(ens, st_i, singles, k)
        lowered += st_i.indicator_count()
    assert lowered > 0, "the network never fired"
    U_i, U_p = st_i.download(), st_p.download()
    dev = float(np.abs(U_i - U_p).sum() / np.abs(U_p).sum())
    assert dev <= 1e-3, dev
    assert not np.array_equal(U_i, U_p)   # the indicator changed the step
    net.close()
