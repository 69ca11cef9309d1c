"""Edge cases of the time step through the C ABI against the oracle, bit-exact: empty and one-particle grids, clouds
smaller than a SELL chunk or a staged block, isolated particles (no neighbours) and starved ones (fewer neighbours than
the reconstruction's D unknowns, which drop to first order), neighbourhoods too wide for the LDS staging (the flux pass
then reads neighbours from HBM), and a non-finite state (MOOD's admissibility and the NaN rules of Julia's max/min)."""
import numpy as np
import pytest

from hcc_hip import _lib as L
from hcc_hip import workloads as W
from helpers import build_pair, first_diff

pytestmark = pytest.mark.gpu


def _random_cloud(N, seed, M, h=0.1, R_factor=2.5, periodic=True):
    """N uniform random particles in the unit square (nominal spacing h, R = R_factor h): a sparse cloud has isolated
    and starved particles. Linear advection (M = 1) or Euler (M = 4) with a jump in x."""
    dom = W.domain(2, (0.0, 0.0), (1.0, 1.0), (periodic, periodic), h, R_factor, 0.0, seed)
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, 1.0, (N, 2))
    cloud = {"pos": pos, "is_boundary": np.zeros(N, np.uint8), "tags": np.zeros(N, np.int32)}
    if M == 1:
        sch = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0], [0.5]], delta=1e-4 * h * h)
        U0 = (1.0 + 0.5 * np.sin(2 * np.pi * pos[:, 0]))[:, None]
    else:
        sch = W.scheme(4, L.HCC_PDE_EULER, delta=1e-4 * h * h)
        rho = np.where(pos[:, 0] > 0.5, 1.0, 0.125)
        p = np.where(pos[:, 0] > 0.5, 1.0, 0.1)
        u = np.zeros((N, 2)) + [0.3, -0.2]
        U0 = W.euler_conserved(rho, u, p, 1.4)
    return {"domain": dom, "scheme": sch, "cloud": cloud, "U0": U0}


def _steps(st, o, n, dt):
    t = 0.0
    for k in range(n):
        dt_g, dt_o = st.time_step(), o.time_step()
        assert dt_g == dt_o, f"step {k}: CFL step {dt_g} vs {dt_o}"
        print(f"step {k} dt {dt}", flush=True)
        st.step(t, dt)
        o.step(t, dt)
        t += dt
        U, U_o = st.download(), o.state()
        assert first_diff(U, U_o) is None, f"step {k}: {first_diff(U, U_o)}"
        assert np.array_equal(st.orders(), o.orders()), f"step {k}: orders"
        assert np.array_equal(st.mood_flags(), o.mood_flags()), f"step {k}: MOOD flags"


@pytest.mark.parametrize("M", [1, 4])
def test_empty_grid(ctx, M):
    """A grid without particles: the step and the CFL estimate run (dt = +inf, as the oracle's minimum over no
    particle) and the state stays empty."""
    wl = _random_cloud(0, 1, M)
    g, st, o = build_pair(ctx, wl)
    assert st.time_step() == o.time_step() == np.inf
    st.step(0.0, 0.01)
    assert st.download().shape == (0, M)


@pytest.mark.parametrize("N", [1, 2, 5, 63, 65, 257])
@pytest.mark.parametrize("M", [1, 4])
def test_tiny_and_sparse_clouds(ctx, N, M):
    """Clouds of 1 to 257 particles (partial SELL chunks and staged blocks), with isolated particles (no neighbour:
    K = 0, the CFL estimate skips them) and starved ones (fewer than D neighbours): three steps bit-exact."""
    wl = _random_cloud(N, N + 1, M)
    g, st, o = build_pair(ctx, wl)
    cnt = np.diff(o.neighbors()[0])
    if N in (1, 2):
        assert np.all(cnt == 0)
    if N == 5:
        assert (cnt == 0).any() and (cnt < 2).sum() >= 2
    _steps(st, o, 3, 0.004)


def test_neighbourhoods_too_wide_for_lds(ctx):
    """R = 8 h on a jittered 40^2 Euler cloud (about 200 neighbours a particle): a staged block's halo with its
    gradient records and index table (about 220 KB) exceeds the LDS, so the flux pass reads neighbours from HBM;
    CFL steps with MOOD firing, bit-exact."""
    n, h = 40, 1.0 / 40
    dom = W.domain(2, (0.0, 0.0), (1.0, 1.0), (True, True), h, 8.0, 0.25, 3)
    cloud = L.cloud_generate(W.domain_desc(dom))
    x = cloud["pos"]
    inside = (np.abs(x[:, 0] - 0.5) < 0.2) & (np.abs(x[:, 1] - 0.5) < 0.2)
    U0 = W.euler_conserved(np.where(inside, 4.0, 1.0), np.zeros((len(x), 2)), np.where(inside, 10.0, 0.1), 1.4)
    wl = {"domain": dom, "scheme": W.scheme(4, L.HCC_PDE_EULER, delta=1e-4 * h * h), "cloud": cloud, "U0": U0}
    g, st, o = build_pair(ctx, wl)
    assert np.mean(np.diff(o.neighbors()[0])) > 150
    for k in range(3):
        dt = 0.4 * o.time_step()
        _steps(st, o, 1, dt)
    assert np.sum(o.mood_flags()) > 0


def test_non_finite_state(ctx):
    """One particle's density is NaN and another's pressure negative: the gather's extrema take Julia's NaN rule,
    MOOD finds the candidates inadmissible and the oracle's step is reproduced (NaN where the oracle has NaN)."""
    wl = W.euler_2d(n=24, seed=5)
    U0 = wl["U0"].copy()
    U0[100, 0] = np.nan
    U0[200, 3] = -1.0
    wl["U0"] = U0
    g, st, o = build_pair(ctx, wl)
    st.step(0.0, 1e-4)
    o.step(0.0, 1e-4)
    U, U_o = st.download(), o.state()
    assert np.isnan(U_o).any()
    assert first_diff(U, U_o) is None, first_diff(U, U_o)
    assert np.array_equal(st.orders(), o.orders())
    assert np.array_equal(st.mood_flags(), o.mood_flags())
