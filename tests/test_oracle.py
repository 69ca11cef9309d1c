"""Pins of the CPU oracle: closed forms, invariants and the expectations of the reference's own unit tests.

Each check is chosen so that a plausible mistake in the restatement (a dropped term, a wrong sign or index, a
transposed operand) fails it. Reference files are cited next to each check.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from hcc_hip import _lib as L
from hcc_hip import workloads as W
from oracle import oracle as O

EPS = np.finfo(float).eps


# ---- scalar building blocks ----------------------------------------------------------------------------------
def test_weight_function_is_inverse_taylor_quartic():
    """MLSWeightFunctions.jl:68-82: w = 1/P4(alpha r^2 / R^2); 0.4 % off exp at the cutoff for alpha = 1 and 23 %
    for alpha = 3 (the docstring's figures)."""
    R = 0.7
    for alpha in (1.0, 3.0):
        for d2 in (0.0, 0.01, 0.2, 0.49):
            y = Fraction(alpha) * Fraction(d2) / (Fraction(R) * Fraction(R))
            exact = 1 / (1 + y + y ** 2 / 2 + y ** 3 / 6 + y ** 4 / 24)
            assert abs(O.weight(alpha, R, d2) - float(exact)) <= 4 * EPS * float(exact)
    assert O.weight(1.0, R, 0.0) == 1.0
    assert abs(O.weight(1.0, R, R * R) / math.exp(-1) - 1) == pytest.approx(0.0037, abs=5e-4)
    assert abs(O.weight(3.0, R, R * R) / math.exp(-3) - 1) == pytest.approx(0.23, abs=0.01)


def test_venkatakrishnan_matches_textbook():
    """unit_core.jl:102-105: VK(r) = (r^2 + 2r)/(r^2 + r + 2) to 4 eps; VK(-2) = 0; BJ(1e17) = 1."""
    for r in (0.0, 1e-3, 0.5, 1.0, 1.5, 7.0, 1e3, 1e7):
        tb = (r * r + 2 * r) / (r * r + r + 2)
        assert O.limiter_phi(2, r) == pytest.approx(tb, rel=4 * EPS, abs=0)
    assert O.limiter_phi(2, -2.0) == 0.0
    assert O.limiter_phi(2, 1e20) == pytest.approx(1.0, rel=1e-15)
    assert O.limiter_phi(1, 1e17) == 1.0 and O.limiter_phi(1, 0.25) == 0.25


PHI_FLAT, PHI_GAP = 2.000000001, 1e-12   # kPhiFlat, kPhiGap of hcc_math.h


def _phi_scan(rec, lim, dup, ddn, bj):
    """The kernels' phi without a quotient per slot (phi_scan_add / phi_scan_finish in hcc_math.h), restated:
    None where a near tie leaves it to the full fold."""
    p1 = p2 = n1 = n2 = 0.0
    for v, l in zip(rec, lim):
        if not l:
            continue
        if v > 0:
            if v > p1:
                p1, p2 = v, p1
            elif p2 < v < p1:
                p2 = v
        else:
            if v < n1:
                n1, n2 = v, n1
            elif n1 < v < n2:
                n2 = v
    inf = math.inf
    rp = dup / p1 if p1 > 0 else inf
    rn = ddn / n1 if n1 < 0 else inf
    rs = O.lib().orc_jmin(rp, rn)
    if bj:
        return O.lib().orc_jmin(1.0, rs)
    if rs != rs:
        return None
    if not rs < PHI_FLAT:
        return 1.0
    safe = lambda x: x >= rs + PHI_GAP or x >= PHI_FLAT
    rp2 = dup / p2 if p2 > 0 else inf
    rn2 = ddn / n2 if n2 < 0 else inf
    if not ((rp == rs or safe(rp)) and (rn == rs or safe(rn)) and (dup == 0 or safe(rp2)) and (ddn == 0 or safe(rn2))):
        return None
    return O.lib().orc_jmin(1.0, O.limiter_phi(2, rs))


def _phi_fold(rec, lim, dup, ddn, bj):
    """_bounding_phi (Limiter.jl:93-107): phi = min(phi, limiter_phi(r)) over the limited slots, from 1."""
    phi = 1.0
    for v, l in zip(rec, lim):
        if l:
            r = (dup if v > 0 else ddn) / v
            phi = O.lib().orc_jmin(phi, O.limiter_phi(1 if bj else 2, r))
    return phi


def test_limiter_phi_from_extreme_increments():
    """The kernels form phi from the largest increments of each side instead of a quotient per slot; the
    restatement here agrees with the reference's fold on random stencils with near ties, ties, signed zeros, huge
    ratios and limited and unlimited slots, and reports the near ties it leaves to the fold."""
    rng = np.random.default_rng(11)
    undecided = 0
    for trial in range(4000):
        n = int(rng.integers(1, 24))
        rec = rng.standard_normal(n) * 10.0 ** rng.uniform(-3, 3)
        if trial % 5 == 0:   # ties and near ties
            rec[: n // 2] = rec[0] * (1.0 + rng.integers(0, 3, n // 2) * 1e-15)
        lim = rng.random(n) < 0.8
        umax = float(np.max(np.abs(rec))) * rng.choice([0.0, 0.3, 1.9, 2.0, 2.2, 50.0, 1e300])
        dup, ddn = umax, -umax * rng.choice([0.0, 0.5, 1.0, 3.0])
        for bj in (False, True):
            fast, ref = _phi_scan(rec, lim, dup, ddn, bj), _phi_fold(rec, lim, dup, ddn, bj)
            if fast is None:
                undecided += 1
                continue
            assert fast == ref and math.copysign(1, fast) == math.copysign(1, ref), (rec, lim, dup, ddn, bj)
    assert 0 < undecided < 1000   # of the 4000 Venkatakrishnan stencils, a fifth of them built around ties


def test_venkatakrishnan_rounding_facts_behind_the_scan():
    """The two facts phi_scan_finish rests on, checked on the oracle's limiter_phi (Limiter.jl:39-45): the rounded
    phi is >= 1 for every ratio >= kPhiFlat (dense next to it, up to 1e308 and inf), and ratios kPhiGap apart on
    [0, kPhiFlat] never give a smaller rounded phi to the larger ratio (its computed value is within 1e-15 of the
    exact rational). Adjacent ratios can invert by an ulp, which is why a near tie goes to the full fold."""
    rng = np.random.default_rng(7)
    rs = [PHI_FLAT, math.inf]
    x = PHI_FLAT
    for _ in range(3000):
        x = float(np.nextafter(x, np.inf))
        rs.append(x)
    rs += list(rng.uniform(PHI_FLAT, 10.0, 20000)) + list(10.0 ** rng.uniform(1.0, 308.0, 20000))
    rs += list(10.0 ** rng.uniform(13.0, 17.5, 20000))   # s = 1/r where fl(1 + s) and fl(1 + 2s) meet
    for r in rs:
        assert O.limiter_phi(2, r) >= 1.0, r
    for r in rng.uniform(0.0, PHI_FLAT, 3000):
        fr = Fraction(float(r))
        exact = (fr * fr + 2 * fr) / (fr * fr + fr + 2)
        assert abs(Fraction(O.limiter_phi(2, float(r))) - exact) < Fraction(1, 10 ** 15)
    xs = rng.uniform(0.0, PHI_FLAT - PHI_GAP, 40000)
    ys = np.minimum(xs + PHI_GAP * (1.0 + rng.uniform(0.0, 1.0, xs.size)), PHI_FLAT)
    for x, y in zip(xs, ys):
        if y - x >= PHI_GAP:
            assert O.limiter_phi(2, float(y)) >= O.limiter_phi(2, float(x)), (x, y)
    inversions = 0
    for x in rng.uniform(0.5, 2.0, 20000):
        y = float(np.nextafter(x, np.inf))
        inversions += O.limiter_phi(2, y) < O.limiter_phi(2, float(x))
    assert inversions > 0   # without the gap, the smallest ratio would not always give the smallest phi
    assert O.limiter_phi(2, 1.9) < 0.996


def test_julia_min_max_mod_semantics():
    """Utils.jl:84-87 (exact min/max, NaN-propagating, unit_core.jl:90-91) and Base.mod for floats."""
    assert O.lib().orc_jmin(0.1, 0.3) == 0.1 and O.lib().orc_jmax(0.1, 0.3) == 0.3
    assert math.isnan(O.lib().orc_jmin(float("nan"), 1.0)) and math.isnan(O.lib().orc_jmax(1.0, float("nan")))
    assert math.copysign(1, O.lib().orc_jmax(-0.0, 0.0)) == 1 and math.copysign(1, O.lib().orc_jmin(0.0, -0.0)) == -1
    assert O.lib().orc_jmin(1e-17, 1.0) == 1e-17
    assert O.lib().orc_jmod(-0.25, 1.0) == 0.75 and O.lib().orc_jmod(1.25, 1.0) == 0.25
    m = O.lib().orc_jmod(-1.0, 1.0)
    assert m == 0.0 and math.copysign(1, m) == 1


def test_spd_inverse_and_regularisation():
    """_cholesky_moment (Interpolators.jl:46-64): the inverse of a healthy SPD matrix; a singular one is
    regularised by sqrt(eps) tr/B; an indefinite one fails."""
    rng = np.random.default_rng(0)
    for B in (2, 3, 4):
        A = rng.standard_normal((B, B))
        A = A @ A.T + 0.1 * np.eye(B)
        X, ok = O.spd_inverse(A)
        assert ok == 1
        assert np.allclose(X @ A, np.eye(B), atol=1e-12)
    S = np.array([[1.0, 1.0], [1.0, 1.0]])        # rank one: retried with the Tikhonov term
    X, ok = O.spd_inverse(S)
    reg = math.sqrt(EPS) * 2.0 / 2
    assert ok == 1 and np.allclose(X, np.linalg.inv(S + reg * np.eye(2)), rtol=1e-6)
    X, ok = O.spd_inverse(np.array([[-1.0, 0.0], [0.0, 1.0]]))
    assert ok == 0


def test_euler_flux_closed_form():
    """The 2-D/3-D Euler flux, wave speed and admissibility (build-defined PDE), against primitive variables."""
    g = 1.4
    for D in (2, 3):
        rho, p = 1.3, 0.7
        u = np.array([0.3, -1.1, 0.4][:D])
        E = p / (g - 1) + 0.5 * rho * (u @ u)
        U = np.concatenate([[rho], rho * u, [E]])
        F, lam, adm = O.euler(D, g, U)
        c = math.sqrt(g * p / rho)
        for d in range(D):
            ex = np.concatenate([[rho * u[d]], rho * u[d] * u + p * np.eye(D)[d], [(E + p) * u[d]]])
            assert np.allclose(F[d], ex, rtol=1e-14)
            assert lam[d] == pytest.approx(abs(u[d]) + c, rel=1e-14)
        assert adm
        U2 = U.copy(); U2[-1] = 0.5 * rho * (u @ u) - 1e-3      # negative pressure
        assert not O.euler(D, g, U2)[2]
        U3 = U.copy(); U3[0] = float("nan")
        assert not O.euler(D, g, U3)[2]


def test_rk_tableaus_order_conditions_and_ssp():
    """unit_core.jl:595-625 (order conditions to 1e-13), ssp_coefficient (ButcherTableaus.jl:86-118)."""
    a, b, c = (np.array(x, float) for x in W.TABLEAUS["RK3_SSP"])
    assert abs(b.sum() - 1) < 1e-13 and abs(b @ c - 0.5) < 1e-13
    assert abs(b @ c ** 2 - 1 / 3) < 1e-13 and abs(b @ (a @ c) - 1 / 6) < 1e-13
    assert np.allclose(a.sum(1), c, atol=1e-15)
    assert W.tableau("RK3_SSP")["ssp"] == pytest.approx(1.0, abs=1e-9)
    assert W.tableau("RK2_SSP")["ssp"] == pytest.approx(1.0, abs=1e-9)
    assert W.tableau("RK1_Euler")["ssp"] == pytest.approx(1.0, abs=1e-9)
    assert W.tableau("RK2_Ralston")["ssp"] == pytest.approx(0.5, abs=1e-9)
    r = W.tableau("RK3_SSP")["substep_ratio"]
    assert 0.999999999 < r <= 1.0


# ---- neighbour search --------------------------------------------------------------------------------------------
def _brute(pos, R, L):
    N = len(pos)
    d = pos[None, :, :] - pos[:, None, :]
    for a, Lw in enumerate(L):
        if Lw:
            d[..., a] -= Lw * np.round(d[..., a] / Lw)
    d2 = (d * d).sum(-1)
    np.fill_diagonal(d2, np.inf)
    return [set(np.nonzero(d2[i] <= R * R)[0]) for i in range(N)]


@pytest.mark.parametrize("dom", [
    W.domain(2, (0, 0), (1, 1), (True, True), 1 / 6, 2.5, 0.25, 1),
    W.domain(2, (0, 0), (1, 1), (False, False), 1 / 10, 2.5, 0.25, 2),
    W.domain(3, (0, 0, 0), (1, 1, 1), (True, False, True), 1 / 6, 2.2, 0.2, 3),
])
def test_neighbour_lists_equal_brute_force(dom):
    """unit_core.jl:331-341: count = brute force, no duplicates, no self; the slot order is window order then
    descending index within a bin (NeighborLogic.jl:47-67, 153-167)."""
    c = O.cloud_generate(dom)
    o = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"], W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0]] * dom["D"]))
    ranges, idx, w, dist = o.neighbors()
    Lw = [1.0 if p else 0.0 for p in dom["periodic"]]
    expect = _brute(c["pos"], dom["R"], Lw)
    bo, cs, cp, dims = o.bins()
    for i in range(o.N):
        lst = idx[ranges[i]:ranges[i + 1]]
        assert len(set(lst)) == len(lst) and i not in lst
        assert set(lst) == expect[i]
        # within one bin the neighbours come in descending index order
        for a, b in zip(lst[:-1], lst[1:]):
            if bo[a] == bo[b]:
                assert a > b
    # weights are the weight function of |d|^2 and distances the minimum image
    for k in range(0, len(idx), 7):
        d2 = float((dist[k] ** 2).sum())
        assert w[k] == pytest.approx(O.weight(1.0, dom["R"], d2), rel=1e-15)
    # bins: counting sort with descending in-bin order
    for b in range(int(dims[3])):
        seg = cp[cs[b]:cs[b + 1]]
        assert np.all(np.diff(seg) < 0) and np.all(bo[seg] == b)


def test_uniform_line_has_four_neighbours():
    """unit_core.jl:682-686: on a uniform periodic 1-D line, R = 2.5 dx gives 4 neighbours (+-1, +-2)."""
    dom = W.domain(1, (0,), (1,), (True,), 1 / 40, 2.5, 0.0, 0)
    c = O.cloud_generate(dom)
    assert c["N"] == 40
    o = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"], W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0]]))
    r, idx, _, _ = o.neighbors()
    assert np.all(np.diff(r) == 4)


@pytest.mark.parametrize("R", [0.45, 0.6])
def test_small_periodic_domain_no_duplicates(R):
    """unit_core.jl:331-341: with one or two bins on a periodic axis repeated window bins are dropped."""
    dom = {"D": 1, "mins": [0.0], "maxs": [1.0], "periodic": [True], "dx": [0.1], "R": R, "alpha": 1.0,
           "jitter": [0.0], "seed": 0}
    c = O.cloud_generate(dom)
    o = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"], W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0]]))
    r, idx, _, _ = o.neighbors()
    x = c["pos"][:, 0]
    for i in range(o.N):
        lst = idx[r[i]:r[i + 1]]
        assert len(set(lst)) == len(lst) and i not in lst
        dd = x - x[i]
        expect = int(np.sum(np.abs(dd - np.round(dd)) <= R + 1e-12)) - 1
        assert len(lst) == expect


# ---- MLS fits, limiter, flux, divergence -----------------------------------------------------------------------
def _linear_field_case(D=2, jitter=0.25, n=24, limiter=L.HCC_LIM_VK, seed=5):
    dom = W.domain(D, (0,) * D, (1,) * D, (False,) * D, 1 / n, 2.5, jitter, seed)
    c = O.cloud_generate(dom)
    speeds = [[0.7], [-1.3], [0.4]][:D]
    sch = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=speeds, limiter=limiter, criterion=L.HCC_CRIT_NONE,
                   strategy=L.HCC_STRAT_NONE)
    o = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"], sch)
    a = np.array([1.7, -0.6, 2.2][:D])
    U = (c["pos"] @ a + 0.3)[:, None]
    return o, c, U, a, np.array([s[0] for s in speeds])


@pytest.mark.parametrize("limiter", [L.HCC_LIM_NONE, L.HCC_LIM_VK, L.HCC_LIM_BJ])
def test_rhs_exact_for_linear_field(limiter):
    """A linear state is reconstructed exactly, the Rusanov dissipation vanishes, and the divergence fit is exact
    for linear flux differences (unit_core.jl:361-380): K_E = -c . grad(u) at every interior node whose stencil is
    interior (ghost particles keep a zero gradient: the gather pass skips them, TimeSteppers.jl:329-334); the
    limited gradient equals the field gradient at every interior node (the limiters do not clip linear data)."""
    o, c, U, a, cv = _linear_field_case(limiter=limiter)
    K, G = o.rhs(U)
    inner = c["is_boundary"] == 0
    deep = _deep_interior(o, c)
    assert deep.sum() > 100
    assert np.allclose(K[deep, 0], -(cv @ a), rtol=1e-9, atol=1e-9)
    assert np.allclose(G[inner, :, 0], a[None, :], rtol=1e-9)
    assert np.all(K[~inner] == 0)
    assert np.max(np.abs(K[inner & ~deep, 0] + cv @ a)) > 1e-6   # next to the ghosts the zero gradients show


def _deep_interior(o, c):
    r, idx, _, _ = o.neighbors()
    isb = c["is_boundary"].astype(bool)
    touch = np.zeros(o.N, bool)
    for i in range(o.N):
        touch[i] = isb[idx[r[i]:r[i + 1]]].any()
    return ~isb & ~touch


def test_rhs_exact_for_linear_field_3d():
    o, c, U, a, cv = _linear_field_case(D=3, n=10, jitter=0.2)
    K, G = o.rhs(U)
    deep = _deep_interior(o, c)
    assert deep.sum() > 20
    assert np.allclose(K[deep, 0], -(cv @ a), rtol=1e-9, atol=1e-9)


def test_constant_state_has_zero_rhs():
    """unit_core.jl:401-407: a constant flux has zero divergence (the divergence weights sum to zero)."""
    wl = W.euler_2d(n=32, seed=4)
    c = wl["cloud"]
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], wl["scheme"])
    U = np.tile(wl["U0"][0], (o.N, 1))
    K, G = o.rhs(U)
    assert np.max(np.abs(K)) < 1e-12 and np.max(np.abs(G)) == 0.0
    ig, og, idv, od, dw = o.moments()
    r, idx, w, dist = o.neighbors()
    sums = np.add.reduceat(dw, r[:-1], axis=0)
    assert np.max(np.abs(sums[np.diff(r) > 0])) < 1e-9 * np.max(np.abs(dw))


def test_limited_interfaces_stay_in_bounds():
    """unit_schemes.jl:611-639: the limited midpoint reconstruction of every interface stays within the
    neighbourhood bounds (to 1e-12)."""
    wl = W.euler_2d(n=40, seed=2)
    c = wl["cloud"]
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], wl["scheme"])
    U = wl["U0"]
    K, G = o.rhs(U)
    r, idx, w, dist = o.neighbors()
    for i in np.nonzero(c["is_boundary"] == 0)[0][::7]:
        nb = idx[r[i]:r[i + 1]]
        lo = np.minimum(U[i], U[nb].min(0))
        hi = np.maximum(U[i], U[nb].max(0))
        rec = U[i][None, :] + 0.5 * dist[r[i]:r[i + 1]] @ G[i]
        assert np.all(rec <= hi + 1e-12 * (1 + np.abs(hi))) and np.all(rec >= lo - 1e-12 * (1 + np.abs(lo)))


# ---- boundary conditions, conservation, MOOD ------------------------------------------------------------------
def test_outflow_copies_outermost_interior_value():
    """unit_core.jl:535-554: on a 1-D line with outflow on both walls every ghost holds the state of the
    outermost interior particle on its side (zero-gradient)."""
    dx = 2 / 40
    dom = W.domain(1, (-1,), (1,), (False,), dx, 2.5, 0.0, 0)
    c = O.cloud_generate(dom)
    sch = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[0.0]], criterion=L.HCC_CRIT_NONE, strategy=L.HCC_STRAT_NONE,
                   bcs=[(1, L.HCC_BC_OUTFLOW), (2, L.HCC_BC_OUTFLOW)])
    o = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"], sch)
    x = c["pos"][:, 0]
    o.set_state(x[:, None].copy())
    o.step(0.0, 1e-3)
    U = o.state()[:, 0]
    inner = c["is_boundary"] == 0
    xlo, xhi = x[inner].min(), x[inner].max()
    for i in np.nonzero(~inner)[0]:
        assert U[i] == (xlo if x[i] < 0 else xhi)
    assert np.array_equal(U[inner], x[inner])


def test_mass_conserved_on_uniform_periodic_lattice():
    """pipeline_tests.jl:43-64: on a uniform periodic lattice sum(U) is conserved to rtol 1e-12."""
    h = 1 / 32
    dom = W.domain(2, (0, 0), (1, 1), (True, True), h, 2.5, 0.0, 0)
    c = O.cloud_generate(dom)
    sch = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0], [1.0]], delta=1e-4 * h * h)
    o = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"], sch)
    x = c["pos"]
    U0 = np.exp(-((x[:, 0] - 0.5) ** 2 + (x[:, 1] - 0.5) ** 2) / 0.01)[:, None]
    o.set_state(U0)
    for k in range(10):
        o.step(0, 0.5 * o.time_step())
    assert abs(o.state().sum() / U0.sum() - 1) < 1e-12


def test_only_mood_drives_all_orders_to_one():
    """unit_core.jl:928-935 and :748-766: OnlyMOOD flags every interior particle; one step ends at order 1
    everywhere and every interior particle carries the step flag; the fixed point terminates."""
    wl = W.advection_2d(16, 1)
    sch = dict(wl["scheme"], criterion=L.HCC_CRIT_ONLY)
    c = wl["cloud"]
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], sch)
    o.set_state(wl["U0"])
    o.step(0, 0.5 * o.time_step())
    assert np.all(o.orders() == 1) and np.all(o.mood_flags() == 1)
    # order-1 everywhere equals the first-order scheme: gradients zero, so the result equals a run without
    # MOOD at order 1 -- checked through boundedness of the update
    U = o.state()
    assert U.min() >= wl["U0"].min() - 1e-14 and U.max() <= wl["U0"].max() + 1e-14


def test_substep_is_what_mood_checks():
    """unit_schemes.jl:240-273: MOOD checks W = Y + (dt/C) K; the ratio is 1/C of the SSP tableau."""
    tb = W.tableau("RK3_SSP")
    assert tb["substep_ratio"] == pytest.approx(1.0, abs=1e-9)
    assert W.tableau("RK2_Ralston")["substep_ratio"] == pytest.approx(2.0, abs=1e-8)


def test_euler_quadrants_mood_keeps_positivity():
    """The four-quadrant Euler problem: MOOD fires, the PAD keeps density and pressure positive."""
    wl = W.euler_2d(n=48, seed=0)
    c = wl["cloud"]
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], wl["scheme"])
    o.set_state(wl["U0"])
    fired = 0
    for k in range(10):
        o.step(0, wl["cfl"] * o.time_step())
        fired += int(o.mood_flags().sum())
    U = o.state()
    rho = U[:, 0]
    p = 0.4 * (U[:, 3] - 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2) / rho)
    assert fired > 0 and np.all(rho > 0) and np.all(p > 0)
    assert o.inadmissible() == 0


def test_outflow_keeps_constant_state_constant():
    """A uniform Euler state with outflow walls is a steady solution: it stays constant to round-off."""
    wl = W.euler_2d(n=32, seed=1)
    c = wl["cloud"]
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], wl["scheme"])
    U = np.tile(wl["U0"][0], (o.N, 1))
    o.set_state(U)
    for k in range(3):
        o.step(0, wl["cfl"] * o.time_step())
    assert np.max(np.abs(o.state() - U)) < 1e-13
