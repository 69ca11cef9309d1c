"""Every compiled (D, M) layout of linear advection through CFL steps with MOOD firing, against the oracle, bit-exact
after every step: the 1-D line with outflow ends and two advected components on the periodic 2-D cloud (both with
unlimited MUSCL, so that MOOD has overshoots to catch), and the periodic 3-D cloud (Venkatakrishnan); the Euler
layouts (2, 4) and (3, 5) are covered by the parity and edge-case files."""
import numpy as np
import pytest

from hcc_hip import _lib as L
from hcc_hip import workloads as W
from helpers import build_pair, first_diff

pytestmark = pytest.mark.gpu


def _adv1d_outflow():
    n = 400
    h = 2.0 / n
    dom = W.domain(1, (-1.0,), (1.0,), (False,), h, 2.5, 0.0, 0)
    cloud = L.cloud_generate(W.domain_desc(dom))
    x = cloud["pos"][:, 0]
    sch = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0]], delta=1e-4 * h, limiter=L.HCC_LIM_NONE,
                   bcs=[(t, L.HCC_BC_OUTFLOW) for t in (1, 2)])
    return {"domain": dom, "scheme": sch, "cloud": cloud, "U0": np.where(np.abs(x) < 0.3, 1.0, 0.2)[:, None]}


def _adv2d_two_components():
    n = 48
    h = 1.0 / n
    dom = W.domain(2, (0.0, 0.0), (1.0, 1.0), (True, True), h, 2.5, 0.25, 2)
    cloud = L.cloud_generate(W.domain_desc(dom))
    x = cloud["pos"]
    sch = W.scheme(2, L.HCC_PDE_LINADV, adv_speed=[[1.0, -0.5], [0.5, 1.0]], delta=1e-4 * h * h, limiter=L.HCC_LIM_NONE)
    box = (np.abs(x[:, 0] - 0.5) < 0.2) & (np.abs(x[:, 1] - 0.5) < 0.2)
    U0 = np.column_stack([np.where(box, 1.0, 0.0), np.sin(2 * np.pi * x[:, 0]) * np.cos(2 * np.pi * x[:, 1])])
    return {"domain": dom, "scheme": sch, "cloud": cloud, "U0": U0}


def _adv3d():
    n = 12
    h = 1.0 / n
    dom = W.domain(3, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (True, True, True), h, 2.2, 0.25, 4)
    cloud = L.cloud_generate(W.domain_desc(dom))
    x = cloud["pos"]
    sch = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0], [0.5], [-0.25]], delta=1e-4 * h ** 3)
    U0 = np.where(np.linalg.norm(x - 0.5, axis=1) < 0.25, 1.0, 0.0)[:, None]
    return {"domain": dom, "scheme": sch, "cloud": cloud, "U0": U0}


CASES = {"adv1d_outflow": _adv1d_outflow, "adv2d_two_components": _adv2d_two_components, "adv3d": _adv3d}


@pytest.mark.parametrize("case", list(CASES))
def test_advection_layouts_cfl_steps(ctx, case):
    wl = CASES[case]()
    g, st, o = build_pair(ctx, wl)
    t, flags = 0.0, 0
    for k in range(6):
        dt_g, dt_o = st.time_step(), o.time_step()
        assert dt_g == dt_o, f"step {k}: CFL step {dt_g} vs {dt_o}"
        dt = 0.5 * dt_g
        print(f"{case} step {k} dt {dt}", flush=True)
        st.step(t, dt)
        o.step(t, dt)
        t += dt
        U, U_o = st.download(), o.state()
        assert np.array_equal(U, U_o), f"step {k}: {first_diff(U, U_o)}"
        assert np.array_equal(st.orders(), o.orders()), f"step {k}: orders"
        assert np.array_equal(st.mood_flags(), o.mood_flags()), f"step {k}: MOOD flags"
        flags += int(np.sum(o.mood_flags()))
    assert flags > 0   # MOOD fired
