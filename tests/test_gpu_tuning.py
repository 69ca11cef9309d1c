"""Launch choices and the autotuner (hcc_stepper_set_tuning, hcc_stepper_autotune): every choice gives the oracle's
step bit for bit (states, dt, orders, MOOD flags), the autotuner leaves the state and the step statistics as they were,
and a second stepper of a tuned shape takes the stored choice without timing."""
import itertools

import numpy as np
import pytest

from hcc_hip import Grid, Stepper
from hcc_hip import workloads as W
from helpers import build_pair, first_diff
from oracle import oracle as O

pytestmark = pytest.mark.gpu

STEPS = 5


def _dense_state(wl):
    """The four quadrants times 1 + 0.01 sin(6 pi x) sin(10 pi y): no constant neighbourhood, and the first MOOD check
    of every stage flags most particles (the dense form of the fixed point's first iteration)."""
    x = wl["cloud"]["pos"]
    U = wl["U0"].copy()
    fac = 1.0 + 0.01 * np.sin(2 * np.pi * 3 * x[:, 0]) * np.sin(2 * np.pi * 5 * x[:, 1])
    U[:, 0] *= fac
    U[:, 3] *= fac
    return U


def _oracle_run(wl):
    c = wl["cloud"]
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], wl["scheme"], threads=8)
    o.set_state(wl["U0"])
    out, t = [], 0.0
    for _ in range(STEPS):
        dt = wl["cfl"] * o.time_step()
        o.step(t, dt)
        t += dt
        out.append((dt, o.state(), o.orders(), o.mood_flags()))
    return out


@pytest.mark.parametrize("case", ["euler2d_n64", "adv2d_n48"])
def test_every_launch_choice_is_bit_exact(ctx, case):
    wl = W.euler_2d(n=64, seed=0) if case == "euler2d_n64" else W.advection_2d(48, 0)
    ref = _oracle_run(wl)
    c = wl["cloud"]
    g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
    st = Stepper(ctx, g, wl["scheme"])
    g0 = 8 if st.M * st.D <= 8 else 16
    choices = [(staged, split, sweep, fuse, dense)
               for staged, split, sweep, fuse in itertools.product([1, 0], [g0, 2 * g0], [65536, 16384], [1, 0])
               for dense in ([64, -1, 0] if staged else [64])]
    for staged, split, sweep, fuse, dense in choices:
        st.set_tuning(staged, split, sweep, fuse, dense)
        assert st.tuning() == {"staged": staged, "split": split, "sweep": sweep, "fuse_cfl": fuse, "dense": dense}
        st.upload(wl["U0"])
        for k, (dt, U_o, ord_o, fl_o) in enumerate(ref):
            st.step_cfl(wl["cfl"])
            tag = f"choice {(staged, split, sweep, fuse, dense)}, step {k}"
            assert st.last_dt() == dt, tag
            U = st.download()
            assert np.array_equal(U, U_o), f"{tag}: {first_diff(U, U_o)}"
            assert np.array_equal(st.orders(), ord_o), tag
            assert np.array_equal(st.mood_flags(), fl_o), tag


def test_autotune_keeps_state_and_statistics_and_caches(ctx):
    wl = W.euler_2d(n=96, seed=2)
    g, st, _ = build_pair(ctx, wl)
    for _ in range(3):
        st.step_cfl(wl["cfl"])
    n0 = st.totals(reset=False)
    U0 = st.download()
    knobs, best, dflt, timed = st.autotune(steps=2, cfl=wl["cfl"])
    print(f"tuned {knobs}: best {best:.4f} ms, default {dflt:.4f} ms", flush=True)
    assert timed and best > 0 and dflt > 0 and best <= dflt
    assert np.array_equal(st.download(), U0), "the autotuner changed the state"
    assert st.totals(reset=False) == n0, "the autotuner changed the statistics"
    # the tuned stepper and an untuned one continue identically
    ref = Stepper(ctx, g, wl["scheme"])
    ref.upload(U0)
    st.upload(U0)
    for k in range(3):
        st.step_cfl(wl["cfl"])
        ref.step_cfl(wl["cfl"])
        assert np.array_equal(st.download(), ref.download()), f"step {k}"
    # a stepper of the same shape takes the stored choice
    other = Stepper(ctx, g, wl["scheme"])
    other.upload(wl["U0"])
    k2, _, _, timed2 = other.autotune(steps=2, cfl=wl["cfl"])
    assert not timed2 and k2 == knobs


@pytest.mark.parametrize("dense", [64, -1, 0])
def test_dense_fixed_point_form_is_bit_exact(ctx, dense):
    """A state whose first MOOD check flags most particles: the first iteration of each fixed point runs in the dense
    form (plain marks, the staged flux pass of the marked particles) from the second step on (64: the last step's count
    decides), in every step (-1) or never (0); states, dt, orders and flags equal the oracle's every step."""
    wl = W.euler_2d(n=64, seed=0)
    wl = dict(wl, U0=_dense_state(wl))
    ref = _oracle_run(wl)
    assert max(int(r[3].sum()) for r in ref) * 64 >= len(wl["U0"]), "the state does not flag enough particles"
    c = wl["cloud"]
    g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
    st = Stepper(ctx, g, wl["scheme"])
    st.set_tuning(dense=dense)
    st.upload(wl["U0"])
    for k, (dt, U_o, ord_o, fl_o) in enumerate(ref):
        st.step_cfl(wl["cfl"])
        tag = f"dense {dense}, step {k}"
        assert st.last_dt() == dt, tag
        U = st.download()
        assert np.array_equal(U, U_o), f"{tag}: {first_diff(U, U_o)}"
        assert np.array_equal(st.orders(), ord_o), tag
        assert np.array_equal(st.mood_flags(), fl_o), tag
