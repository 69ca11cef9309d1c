"""Pins of the oracle's MUSCL order 3: exactness on quadratic fields, the neighbour-count safeguard of the divergence
degree with the reference's CFL weight sums, limited reconstructions inside the local bounds, three MOOD levels and
the observed order of accuracy.

MUSCL3 (src/Interpolation/MUSCL.jl:11-32, 351-508): a degree-2 reconstruction (graded Taylor basis with 1/prod(a!)
factors, Interpolators.jl:122-218), a divergence fit of degree 3 with intercept where the stencil carries it, the
direct (Cholesky) solves of the lower orders, and :soft / :hard limiting of the terms of degree 2 (Limiter.jl:135).
"""
import numpy as np
import pytest

from hcc_hip import _lib as L
from hcc_hip import workloads as W
from oracle import oracle as O


def _deep_interior(o, c):
    r, idx, _, _ = o.neighbors()
    isb = c["is_boundary"].astype(bool)
    touch = np.zeros(o.N, bool)
    for i in range(o.N):
        touch[i] = isb[idx[r[i]:r[i + 1]]].any()
    return ~isb & ~touch


def _quadratic(x, D):
    """u = 0.3 + a . x + x^T Q x / 2 with a fixed symmetric Q; returns (u, grad u)."""
    a = np.array([1.7, -0.6, 2.2][:D])
    Q = np.array([[1.3, -0.8, 0.5], [-0.8, 2.1, 0.7], [0.5, 0.7, -1.1]])[:D, :D]
    u = 0.3 + x @ a + 0.5 * np.einsum("ni,ij,nj->n", x, Q, x)
    return u, a[None, :] + x @ Q


def _case(D, n, jitter, order=3, limiter=L.HCC_LIM_NONE, R=2.5, seed=5, M=1, mode=L.HCC_MODE_HARD):
    dom = W.domain(D, (0,) * D, (1,) * D, (False,) * D, 1 / n, R, jitter, seed)
    c = O.cloud_generate(dom)
    speeds = [[0.7, -0.4][:M], [-1.3, 0.9][:M], [0.4, 0.2][:M]][:D]
    sch = W.scheme(M, L.HCC_PDE_LINADV, adv_speed=speeds, limiter=limiter, mode=mode, criterion=L.HCC_CRIT_NONE,
                   strategy=L.HCC_STRAT_NONE, order=order)
    o = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"], sch)
    return o, c, np.array(speeds, float)


@pytest.mark.parametrize("D,n,jitter", [(1, 40, 0.2), (2, 20, 0.25), (3, 8, 0.2)])
def test_quadratic_field_is_exact(D, n, jitter):
    """For a quadratic u the degree-2 reconstruction is exact, both interface states equal u at the midpoint, the
    flux difference c (u(x_i + d/2) - u(x_i)) is quadratic in the offset d, and the divergence fit (degree 3, or 2
    where the safeguard lowers it, both with intercept) reproduces it: K = -c . grad u(x_i) at every interior particle
    whose stencil holds no ghost. A wrong Taylor factor, a basis term out of order between the fit and the interface
    basis, or a wrong scale factor of the degree-2 terms breaks it; so does the degree-1 divergence of MUSCL2 (its
    fit of a quadratic is not exact on a jittered cloud)."""
    o, c, cv = _case(D, n, jitter)
    u, g = _quadratic(c["pos"], D)
    K, G = o.rhs(u[:, None])
    deep = _deep_interior(o, c)
    assert deep.sum() > 10
    exact = -(g @ cv[:, 0])
    assert np.allclose(K[deep, 0], exact[deep], rtol=1e-8, atol=1e-8)
    assert np.allclose(G[deep, :D, 0], g[deep], rtol=1e-8, atol=1e-8)   # the first-derivative terms
    o2 = O.OracleSim(o_dom(D, n, jitter), c["pos"], c["is_boundary"], c["tags"],
                     W.scheme(1, L.HCC_PDE_LINADV, adv_speed=cv.tolist(), limiter=L.HCC_LIM_NONE,
                              criterion=L.HCC_CRIT_NONE, strategy=L.HCC_STRAT_NONE, order=2))
    K2, _ = o2.rhs(u[:, None])
    if jitter > 0:
        assert np.max(np.abs(K2[deep, 0] - exact[deep])) > 1e-4


def o_dom(D, n, jitter, R=2.5, seed=5):
    return W.domain(D, (0,) * D, (1,) * D, (False,) * D, 1 / n, R, jitter, seed)


def test_second_derivative_terms_are_the_hessian():
    """The degree-2 terms of the limited (here unlimited) gradient record are the second derivatives in the order of
    _generate_exponents: x^2, y^2, then xy (2-D); with the 1/2 and 1 Taylor factors they scale back to Q."""
    o, c, _ = _case(2, 20, 0.25)
    u, _ = _quadratic(c["pos"], 2)
    _, G = o.rhs(u[:, None])
    deep = _deep_interior(o, c)
    Q = np.array([[1.3, -0.8], [-0.8, 2.1]])
    assert np.allclose(G[deep, 2, 0], Q[0, 0], rtol=1e-7) and np.allclose(G[deep, 3, 0], Q[1, 1], rtol=1e-7)
    assert np.allclose(G[deep, 4, 0], Q[0, 1], rtol=1e-7)


def test_lower_orders_use_the_direct_solves():
    """Particles below the full order take the direct (Cholesky) solves of degree 1 (MUSCL.jl:371-376, 409-419): on a
    linear field K = -c . grad u at order 2 and the same K as MUSCL2's cached fits to round-off; at order 1 the
    reconstruction is constant (zero gradients)."""
    o, c, cv = _case(2, 20, 0.25)
    a = np.array([1.7, -0.6])
    u = (c["pos"] @ a + 0.3)[:, None]
    deep = _deep_interior(o, c)
    K3 = o.rhs_at(u, np.full(o.N, 2, np.int32))
    assert np.allclose(K3[deep, 0], -(cv[:, 0] @ a), rtol=1e-9, atol=1e-9)
    o2 = O.OracleSim(o_dom(2, 20, 0.25), c["pos"], c["is_boundary"], c["tags"],
                     W.scheme(1, L.HCC_PDE_LINADV, adv_speed=cv.tolist(), limiter=L.HCC_LIM_NONE,
                              criterion=L.HCC_CRIT_NONE, strategy=L.HCC_STRAT_NONE, order=2))
    K2 = o2.rhs_at(u, np.full(o.N, 2, np.int32))
    inner = c["is_boundary"] == 0
    assert np.allclose(K3[inner], K2[inner], rtol=1e-10, atol=1e-10)
    K1 = o.rhs_at(u, np.ones(o.N, np.int32))
    K1b = o2.rhs_at(u, np.ones(o.N, np.int32))
    # order 1: constant reconstruction; MUSCL3 fits the divergence with the degree-1 direct solve where MUSCL2 uses
    # its degree-1 cache: the same fit
    assert np.allclose(K1[inner], K1b[inner], rtol=1e-10, atol=1e-10)


def _sum_abs_c(D, n, R, order=3):
    """Sum over the slots of an interior particle of |c| dx on a uniform periodic lattice (make_pg_1d: [-1, 1],
    ExponentialWeightFunction(1, R dx))."""
    dom = W.domain(D, (-1.0,) * D, (1.0,) * D, (True,) * D, 2.0 / n, R, 0.0, 0)
    c = O.cloud_generate(dom)
    sch = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0]] * D, limiter=L.HCC_LIM_NONE, order=order)
    o = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"], sch)
    r, idx, w, dist = o.neighbors()
    m = o.moments()
    dw = m["dweights"] if isinstance(m, dict) else m[4]
    i = n // 2
    return np.sum(np.abs(dw[r[i]:r[i + 1]])) * (2.0 / n), r[i + 1] - r[i], (o, c, r, w, dist, dw)


def test_divergence_degree_safeguard_and_cfl_weights():
    """unit_schemes.jl:429-441, 495-502: in 1-D with R = 2.5 dx (4 neighbours) the safeguard lowers MUSCL3's
    divergence to degree 2, whose weights sum to |c| dx = 0.644 (degree 3 would give 1.500, the classical fourth-order
    weights 1/12, 2/3); with R = 3.5 dx (6 neighbours) degree 3 is kept."""
    s25, nb25, _ = _sum_abs_c(1, 40, 2.5)
    assert nb25 == 4
    assert s25 == pytest.approx(0.644, abs=6e-4)
    s35, nb35, (o, c, r, w, dist, dw) = _sum_abs_c(1, 40, 3.5)
    assert nb35 == 6
    # degree 3 with intercept on 6 points: an independent weighted least-squares restatement
    i = 20
    d = dist[r[i]:r[i + 1], 0] * (40 / 2.0)
    P = np.stack([np.ones_like(d), d, d * d / 2, d ** 3 / 6], axis=1)
    Wt = w[r[i]:r[i + 1]]
    cfit = np.linalg.solve(P.T @ (Wt[:, None] * P), (P * Wt[:, None]).T)[1] * (40 / 2.0)
    assert np.allclose(dw[r[i]:r[i + 1], 0], cfit, rtol=1e-9, atol=1e-9)


def test_mixed_divergence_degrees_match_least_squares():
    """unit_schemes.jl:444-466: a jittered 2-D periodic cloud holds particles whose stencils carry the degree-3
    divergence (at least 15 neighbours: 2 nb >= 3 (basis_count(2, 3) + 1)) and particles lowered to degree 2; every
    particle's per-slot weights are those of the weighted least-squares fit with intercept of its own degree (an
    independent numpy solve). R = 2.2 dx instead of the reference's 2 dx: with this build's jitter stream no particle
    of a 2 dx cloud reaches 15 neighbours."""
    n = 16
    dom = W.domain(2, (0.0, 0.0), (1.0, 1.0), (True, True), 1.0 / n, 2.2, 0.25, 5)
    c = O.cloud_generate(dom)
    sch = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0], [1.0]], limiter=L.HCC_LIM_NONE, order=3)
    o = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"], sch)
    r, idx, w, dist = o.neighbors()
    m = o.moments()
    dw = m["dweights"] if isinstance(m, dict) else m[4]
    cnt = np.diff(r)
    deg = np.where(cnt >= 15, 3, np.where(cnt >= 6, 2, 1))
    assert (deg == 3).any() and (deg == 2).any()

    def basis(q, d):
        x, y = d[:, 0], d[:, 1]
        cols = [np.ones_like(x), x, y]
        if q >= 2:
            cols += [x * x / 2, y * y / 2, x * y]
        if q >= 3:
            cols += [x ** 3 / 6, y ** 3 / 6, x * x * y / 2, x * y * y / 2]
        return np.stack(cols, axis=1)

    for i in range(o.N):
        dd = dist[r[i]:r[i + 1]] * n
        P = basis(deg[i], dd)
        Wt = w[r[i]:r[i + 1]]
        X = np.linalg.solve(P.T @ (Wt[:, None] * P), (P * Wt[:, None]).T)
        assert np.allclose(dw[r[i]:r[i + 1]], X[1:3].T * n, rtol=1e-7, atol=1e-7 * n), i


ICS_1D = [lambda x: 0.5 * (1 + np.tanh(x / 0.04)), lambda x: (np.abs(x) < 0.3).astype(float), np.abs,
          lambda x: np.exp(-x ** 2 / 0.002), lambda x: np.where(x < 0, 0.0, np.where(x < 0.2, 5 * x, 0.5))]


def _max_overshoot(limiter, mode, ic, N=60):
    dom = W.domain(1, (-1.0,), (1.0,), (True,), 2.0 / N, 2.5, 0.0, 0)
    c = O.cloud_generate(dom)
    sch = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0]], limiter=limiter, mode=mode, order=3,
                   criterion=L.HCC_CRIT_NONE, strategy=L.HCC_STRAT_NONE)
    o = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"], sch)
    u = ic(c["pos"][:, 0])
    _, G = o.rhs(u[:, None])
    r, idx, _, dist = o.neighbors()
    worst = 0.0
    for i in range(o.N):
        nb = idx[r[i]:r[i + 1]]
        lo, hi = min(u[nb].min(), u[i]), max(u[nb].max(), u[i])
        for k in range(r[i], r[i + 1]):
            h = 0.5 * dist[k, 0]
            rec = u[i] + G[i, 0, 0] * h + G[i, 1, 0] * (0.5 * h * h)
            worst = max(worst, rec - hi, lo - rec)
    return worst


@pytest.mark.parametrize("limiter,mode", [(L.HCC_LIM_BJ, L.HCC_MODE_SOFT), (L.HCC_LIM_VK, L.HCC_MODE_HARD)])
def test_limited_reconstructions_stay_in_the_local_bounds(limiter, mode):
    """unit_schemes.jl:611-639: the midpoint reconstructions of a limited MUSCL3 lie within the neighbourhood's
    [min, max] for steps, kinks, narrow peaks and ramps; unlimited they overshoot the box by more than 10 %."""
    for ic in ICS_1D:
        assert _max_overshoot(limiter, mode, ic) <= 1e-12
    assert _max_overshoot(L.HCC_LIM_NONE, L.HCC_MODE_HARD, ICS_1D[1]) > 0.1


def test_soft_mode_leaves_the_second_derivatives_unlimited():
    """Limiter.jl:134-140: :hard scales every term by phi, :soft only the first-derivative ones."""
    o_h, c, _ = _case(2, 20, 0.25, limiter=L.HCC_LIM_VK, mode=L.HCC_MODE_HARD)
    o_s, _, _ = _case(2, 20, 0.25, limiter=L.HCC_LIM_VK, mode=L.HCC_MODE_SOFT)
    o_n, _, _ = _case(2, 20, 0.25, limiter=L.HCC_LIM_NONE)
    x = c["pos"]
    u = np.tanh((x[:, 0] - 0.5) / 0.05)[:, None]
    _, Gh = o_h.rhs(u)
    _, Gs = o_s.rhs(u)
    _, Gn = o_n.rhs(u)
    inner = c["is_boundary"] == 0
    phi_s = np.ones(o_h.N)
    lim = inner & (np.abs(Gn[:, 0, 0]) > 1e-8)
    phi_s[lim] = Gs[lim, 0, 0] / Gn[lim, 0, 0]
    assert (phi_s[lim] < 0.999).sum() > 10
    assert np.array_equal(Gs[inner, 2:, 0], Gn[inner, 2:, 0])          # soft: degree-2 terms untouched
    assert np.allclose(Gh[lim, 2:, 0], Gn[lim, 2:, 0] * phi_s[lim, None], rtol=1e-12, atol=0)   # hard: all scaled
    assert np.allclose(Gh[inner, :2, 0], Gs[inner, :2, 0], rtol=0, atol=0)


def test_only_mood_drops_three_levels():
    """unit_core.jl:928-935 with MUSCL3: OnlyMOOD flags every candidate, each checkpoint drops one order at a time
    until order 1 (3 -> 2 -> 1), so a step ends with every interior particle at order 1 and the first-order step's
    state (the same as a MUSCL2 OnlyMOOD step: both end at order 1 with the degree-1 divergence)."""
    wl = W.advection_2d(n=24, seed=3)
    c = wl["cloud"]
    s3 = dict(wl["scheme"], muscl_order=3, criterion=L.HCC_CRIT_ONLY)
    s2 = dict(wl["scheme"], criterion=L.HCC_CRIT_ONLY)
    o3 = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], s3)
    o2 = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], s2)
    for o in (o3, o2):
        o.set_state(wl["U0"])
    dt = 0.4 * o3.time_step()
    o3.step(0.0, dt)
    o2.step(0.0, dt)
    inner = c["is_boundary"] == 0
    assert np.all(o3.orders()[inner] == 1)
    assert o3.mood_iterations() == 2 * o2.mood_iterations()
    assert np.allclose(o3.state(), o2.state(), rtol=1e-12, atol=1e-14)


def test_interface_orders_with_three_levels():
    """unit_core.jl:688-697 with orders 3 and 1: EPD0 (3, 1), EPD1 the edge minimum, EPD2 the neighbourhood one."""
    dom = W.domain(1, (-1.0,), (1.0,), (True,), 2.0 / 40, 2.5, 0.0, 0)
    c = O.cloud_generate(dom)
    o = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"],
                    W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0]], order=3))
    r, idx, _, _ = o.neighbors()
    i = 10
    nb = idx[r[i]:r[i + 1]]
    x = c["pos"][:, 0]
    j = int(nb[np.argmin(np.abs(x[nb] - (x[i] + 0.1)))])     # i + 2 dx
    k = int(nb[np.argmin(np.abs(x[nb] - (x[i] - 0.1)))])     # i - 2 dx, not a neighbour of j
    orders = np.full(o.N, 3, np.int32)
    orders[j] = 1
    assert o.interface_orders(L.HCC_STRAT_EPD0, orders, i, j) == (3, 1)
    assert o.interface_orders(L.HCC_STRAT_STRICT_EPD0, orders, i, j) == (3, 1)
    assert o.interface_orders(L.HCC_STRAT_EPD1, orders, i, j) == (1, 1)
    assert o.interface_orders(L.HCC_STRAT_EPD1, orders, i, k) == (3, 3)
    assert o.interface_orders(L.HCC_STRAT_EPD2, orders, i, k) == (1, 1)


def _eoc_run(N, order, R, tmax=2.0):
    dom = W.domain(1, (-1.0,), (1.0,), (True,), 2.0 / N, R, 0.0, 0)
    c = O.cloud_generate(dom)
    sch = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0]], limiter=L.HCC_LIM_NONE, criterion=L.HCC_CRIT_NONE,
                   strategy=L.HCC_STRAT_NONE, tableau_name="RK4_Classical", order=order)
    o = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"], sch, threads=4)
    x = c["pos"][:, 0]
    g = lambda y: np.exp(-(y / 0.2) ** 2)
    o.set_state(g(x)[:, None])
    t = 0.0
    while tmax - t > 1e-12:
        dt = min(0.2 * o.time_step(), tmax - t)
        o.step(t, dt)
        t += dt
    u = o.state()[:, 0]
    xe = np.mod(x - tmax + 1.0, 2.0) - 1.0
    ue = g(xe)
    return np.sqrt(np.sum((u - ue) ** 2) / np.sum(ue ** 2))


def test_observed_order_of_accuracy():
    """pipeline_tests.jl:103-112: odd MUSCL orders reach their order once the stencil carries the higher divergence
    degree: MUSCL3 with R = 3.5 dx converges at an observed order in (2.6, 3.6) from N = 200 to 400 (RK4, CFL 0.2,
    one period of a Gaussian); MUSCL2 stays near 2."""
    e200, e400 = _eoc_run(200, 3, 3.5), _eoc_run(400, 3, 3.5)
    eoc = np.log2(e200 / e400)
    assert 2.6 < eoc < 3.6, eoc
    f200, f400 = _eoc_run(200, 2, 3.5), _eoc_run(400, 2, 3.5)
    assert np.log2(f200 / f400) < 2.5
