"""The committed golden vectors (tests/golden/*.npz, written by tests/golden/make_golden.py) reproduced bit for bit:
by the oracle and the product's host set-up on the CPU, and by the HIP path through the C ABI on the GPU."""
import glob
import os

import numpy as np
import pytest

from hcc_hip import Grid, Stepper
from hcc_hip import _lib as L
from hcc_hip import workloads as W
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
NAMES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, "golden", "*.npz")))


def _case(name):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg.cases()[name]


def test_fixtures_present():
    assert set(NAMES) == {"adv2d_64", "euler2d_n12", "euler2d_n16"}


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    g = np.load(os.path.join(HERE, "golden", name + ".npz"))
    wl, steps = _case(name)
    c = wl["cloud"]                                   # the product's host generator
    assert np.array_equal(c["pos"], g["pos"]) and np.array_equal(c["tags"], g["tags"])
    o = O.OracleSim(wl["domain"], g["pos"], g["is_boundary"], g["tags"], wl["scheme"])
    o.set_state(g["U0"])
    r, idx, w, dist = o.neighbors()
    assert np.array_equal(r, g["ranges"]) and np.array_equal(idx, g["indices"])
    assert np.array_equal(w, g["weights"]) and np.array_equal(dist, g["distances"])
    K, G = o.rhs(g["U0"])
    assert np.array_equal(K, g["K0"]) and np.array_equal(G, g["G0"])
    t = 0.0
    for k in range(int(g["steps"])):
        dt = wl["cfl"] * o.time_step()
        assert dt == float(g[f"dt{k}"])
        o.step(t, dt)
        t += dt
        assert np.array_equal(o.state(), g[f"U{k + 1}"])
        assert np.array_equal(o.orders(), g[f"orders{k + 1}"]) and np.array_equal(o.mood_flags(), g[f"flags{k + 1}"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_reproduces_golden(ctx, name):
    g = np.load(os.path.join(HERE, "golden", name + ".npz"))
    wl, steps = _case(name)
    grid = Grid(ctx, W.domain_desc(wl["domain"]), g["pos"], g["is_boundary"], g["tags"])
    st = Stepper(ctx, grid, wl["scheme"])
    r, idx, w, dist = grid.neighbors()
    assert np.array_equal(r, g["ranges"]) and np.array_equal(idx, g["indices"])
    assert np.array_equal(w, g["weights"]) and np.array_equal(dist, g["distances"])
    ig, og, idv, od, dw = st.moments()
    assert np.array_equal(ig, g["inv_grad"]) and np.array_equal(idv[:, 1:, :], g["inv_div"])
    assert np.array_equal(dw, g["dweights"])
    K, G = st.rhs(g["U0"])
    assert np.array_equal(K, g["K0"]) and np.array_equal(G, g["G0"])
    st.upload(g["U0"])
    t = 0.0
    for k in range(int(g["steps"])):
        dt = wl["cfl"] * st.time_step()
        assert dt == float(g[f"dt{k}"])
        st.step(t, dt)
        t += dt
        assert np.array_equal(st.download(), g[f"U{k + 1}"]), f"state after step {k + 1}"
        assert np.array_equal(st.orders(), g[f"orders{k + 1}"]) and np.array_equal(st.mood_flags(), g[f"flags{k + 1}"])
