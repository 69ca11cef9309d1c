"""Pins of the oracle's MOOD strategies and criteria: EPD0 / EPD1 / EPD2 / StrictEPD0 interface orders, the order
reach, StrictEPD0's forcing from the snapshot, the dirty-mask expansion and MOODu2's plateau test, each against the
expectations of the reference's own unit tests (test/unit_core.jl), plus step-level invariants that tell the
strategies apart.
"""
import numpy as np
import pytest

from hcc_hip import _lib as L
from hcc_hip import workloads as W
from oracle import oracle as O


def _line40():
    """make_pg_1d(40) of the reference tests: a uniform periodic line, R = 2.5 dx (4 neighbours each)."""
    dom = W.domain(1, (0,), (1,), (True,), 1 / 40, 2.5, 0.0, 0)
    c = O.cloud_generate(dom)
    o = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"], W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0]]))
    r, idx, _, _ = o.neighbors()
    return o, c["pos"][:, 0], r, idx


def _at(x, v):
    """index of the particle at position v (mod 1) on the line"""
    d = np.abs(x - v % 1.0)
    return int(np.argmin(np.minimum(d, 1 - d)))


def test_interface_orders():
    """unit_core.jl:688-697: EPD0 and StrictEPD0 keep (o_i, o_j); EPD1 takes the edge minimum; EPD2 the minimum over
    N(i) u N(j), so a low-order particle next to i lowers the edge (i, k) on the other side."""
    o, x, r, idx = _line40()
    i = _at(x, x[10])
    j, k = _at(x, x[i] + 2 / 40), _at(x, x[i] - 2 / 40)   # neighbours of i on either side; j is not adjacent to k
    assert j in idx[r[i]:r[i + 1]] and k in idx[r[i]:r[i + 1]] and k not in idx[r[j]:r[j + 1]]
    orders = np.full(o.N, 3, np.int32)
    orders[j] = 1
    assert o.interface_orders(L.HCC_STRAT_EPD0, orders, i, j) == (3, 1)
    assert o.interface_orders(L.HCC_STRAT_STRICT_EPD0, orders, i, j) == (3, 1)
    assert o.interface_orders(L.HCC_STRAT_EPD1, orders, i, j) == (1, 1)
    assert o.interface_orders(L.HCC_STRAT_EPD1, orders, i, k) == (3, 3)
    assert o.interface_orders(L.HCC_STRAT_EPD2, orders, i, k) == (1, 1)
    # EPD2 on an edge whose two neighbourhoods miss j: three spacings from j on the far side of k
    far = _at(x, x[k] - 1 / 40)
    assert o.interface_orders(L.HCC_STRAT_EPD2, orders, k, far) == (3, 3)


def test_order_reach():
    """unit_core.jl:699-706: a MUSCL flux depends on orders one hop away (two with EPD2)."""
    o, _, _, _ = _line40()
    assert o.order_reach(L.HCC_STRAT_EPD0) == 1 and o.order_reach(L.HCC_STRAT_EPD1) == 1
    assert o.order_reach(L.HCC_STRAT_STRICT_EPD0) == 1 and o.order_reach(L.HCC_STRAT_EPD2) == 2


def test_mask_expansion():
    """unit_core.jl:708-715: each hop adds the n_nb particles next to the marked set on the line."""
    o, x, r, idx = _line40()
    mask = np.zeros(o.N, bool)
    mask[7] = True
    m1 = o.expand(mask, 1)
    assert m1.sum() == 1 + 4 and set(np.flatnonzero(m1)) == {7, *idx[r[7]:r[8]]}
    assert o.expand(m1, 1).sum() == 1 + 8
    assert np.array_equal(o.expand(mask, 2), o.expand(m1, 1))


def test_forcing_uses_the_snapshot():
    """unit_core.jl:717-746: the neighbours of a dropped particle are pulled down to its order (one hop, no cascade,
    the particle itself not marked); a neighbour that dropped itself is forced from the snapshot."""
    o, x, r, idx = _line40()
    p = 20
    nbs = list(idx[r[p]:r[p + 1]])
    orders = np.full(o.N, 3, np.int32)
    orders[p] = 1
    trig = np.zeros(o.N, bool)
    trig[p] = True
    new, ch = o.force(trig, orders.copy(), orders, np.zeros(o.N, bool))
    assert all(new[j] == 1 for j in nbs) and all(ch[j] for j in nbs)
    p3 = _at(x, x[p] + 3 / 40)
    assert new[p3] == 3 and not ch[p3]
    assert ch.sum() == len(nbs) and not ch[p]
    q = nbs[0]
    orders = np.full(o.N, 3, np.int32)
    orders[p], orders[q] = 1, 2
    trig[q] = True
    new, ch = o.force(trig, orders.copy(), orders, np.zeros(o.N, bool))
    assert new[q] == 1 and ch[q]
    assert all(new[j] == 1 for j in nbs)


# ---- MOODu2 (unit_core.jl:781-880): a 2-D periodic lattice with hand-set states ---------------------------------
def _lattice20():
    n = 20
    dom = W.domain(2, (0, 0), (1, 1), (True, True), 1 / n, 2.5, 0.0, 0)
    c = O.cloud_generate(dom)
    o = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"],
                    W.scheme(2, L.HCC_PDE_LINADV, adv_speed=[[1.0, 1.0], [1.0, 1.0]]))
    r, idx, _, _ = o.neighbors()
    return o, c["pos"], r, idx, 1 / n


def _nearest(x, c):
    return int(np.argmin(np.sum((x - np.asarray(c)) ** 2, axis=1)))


def _local_extrema(U, r, idx, p, m):
    v = np.concatenate([[U[p, m]], U[idx[r[p]:r[p + 1]], m]])
    return v.min(), v.max()


DELTA = 1e-8


def test_u2_below_order_three_is_the_dmp():
    """unit_core.jl:830-841: at MUSCL order 2 there are no curvature coefficients: MOODu2 is the plain DMP on the
    components that violate it (an undershoot of the Gaussian's discrete minimum is flagged)."""
    o, x, r, idx, dx = _lattice20()
    g = np.exp(-((x[:, 0] - 0.5) ** 2 + (x[:, 1] - 0.5) ** 2) / 0.1)
    U = np.column_stack([1 - 0.5 * g, np.sin(2 * np.pi * x[:, 0])])
    p = _nearest(x, (0.5, 0.5))
    lo1, _ = _local_extrema(U, r, idx, p, 0)
    assert o.u2_flags(U, p, [lo1 - 10 * DELTA, U[p, 1]], DELTA)
    assert not o.u2_flags(U, p, [lo1 - 0.5 * DELTA, U[p, 1]], DELTA)   # inside the padding d
    lo2, _ = _local_extrema(U, r, idx, p, 1)
    assert o.u2_flags(U, p, [U[p, 0], lo2 - 10 * DELTA], DELTA)
    assert not o.u2_flags(U, p, U[p], DELTA)


def test_u2_plateau():
    """unit_core.jl:843-864: a ripple of amplitude 1e-6 far from a bump of height 0.5: with the plateau test on
    (1e-3 of the global range 0.5) an overshoot of the ripple by 5 d is accepted, without it flagged."""
    o, x, r, idx, dx = _lattice20()
    narrow = np.exp(-((x[:, 0] - 0.5) ** 2 + (x[:, 1] - 0.5) ** 2) / 0.01)
    ripple = 1e-6 * np.sin(20 * np.pi * x[:, 0] + 0.7)
    U = np.column_stack([1 - 0.5 * narrow + ripple, np.zeros(len(x))])
    p = _nearest(x, (0.1, 0.1))
    lo, hi = _local_extrema(U, r, idx, p, 0)
    assert hi - lo < 1e-5
    cand = [hi + 5 * DELTA, 0.0]
    assert o.u2_flags(U, p, cand, DELTA, plateau=0.0)
    assert not o.u2_flags(U, p, cand, DELTA, plateau=1e-3)
    # the threshold is the global range times plateau: just above the local range it exempts, just below it does not
    rng = U[:, 0].max() - U[:, 0].min()
    assert not o.u2_flags(U, p, cand, DELTA, plateau=1.01 * (hi - lo) / rng)
    assert o.u2_flags(U, p, cand, DELTA, plateau=0.99 * (hi - lo) / rng)


def test_u2_jump_is_flagged_with_the_plateau_test_on():
    """unit_core.jl:866-874: an overshoot next to a unit jump is flagged with or without the plateau test."""
    o, x, r, idx, dx = _lattice20()
    U = np.column_stack([(x[:, 0] >= 0.5).astype(float), np.zeros(len(x))])
    p = _nearest(x, (0.5 - dx / 2, 0.5))
    _, hi = _local_extrema(U, r, idx, p, 0)
    for plateau in (0.0, 1e-3, 1e-2):
        assert o.u2_flags(U, p, [hi + 1e-3, 0.0], DELTA, plateau=plateau)


# ---- step-level invariants ----------------------------------------------------------------------------------------
def _euler(strategy, criterion, plateau=0.0, n=40, seed=2, substeps=True):
    wl = W.euler_2d(n=n, seed=seed)
    wl["scheme"] = W.scheme(4, L.HCC_PDE_EULER, delta=wl["scheme"]["delta"], bcs=wl["scheme"]["bcs"],
                            strategy=strategy, criterion=criterion, plateau=plateau, mood_substeps=substeps)
    c = wl["cloud"]
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], wl["scheme"], threads=4)
    o.set_state(wl["U0"])
    return wl, o


def test_strict_epd0_forces_every_neighbour_of_a_flagged_particle():
    """StrictEPD0 (MOOD.jl:22-64): after a step every interior neighbour of a flagged particle (order 1: it dropped)
    sits at order 1 too; EPD0 (no forcing) leaves some at order 2 on the same problem."""
    seen = {}
    for strat in (L.HCC_STRAT_STRICT_EPD0, L.HCC_STRAT_EPD0):
        wl, o = _euler(strat, L.HCC_CRIT_U1)
        isb = wl["cloud"]["is_boundary"].astype(bool)
        r, idx, _, _ = o.neighbors()
        o.step(0.0, wl["cfl"] * o.time_step())
        f, ords = o.mood_flags().astype(bool), o.orders()
        assert f.sum() > 0
        assert np.all(ords[f] == 1)
        nb_of_flagged = np.unique(np.concatenate([idx[r[p]:r[p + 1]] for p in np.flatnonzero(f)]))
        nb_of_flagged = nb_of_flagged[~isb[nb_of_flagged]]
        seen[strat] = int(np.sum(ords[nb_of_flagged] == 2))
    assert seen[L.HCC_STRAT_STRICT_EPD0] == 0 and seen[L.HCC_STRAT_EPD0] > 0


def test_u2_without_plateau_equals_u1_at_order_two():
    """_u2_fails (MOOD.jl:215-270): at MUSCL order 2 with the plateau test off MOODu2 is MOODu1, step for step."""
    wl, o1 = _euler(L.HCC_STRAT_EPD1, L.HCC_CRIT_U1)
    _, o2 = _euler(L.HCC_STRAT_EPD1, L.HCC_CRIT_U2, plateau=0.0)
    for k in range(3):
        dt = wl["cfl"] * o1.time_step()
        o1.step(0.0, dt)
        o2.step(0.0, dt)
        assert np.array_equal(o1.state(), o2.state()) and np.array_equal(o1.mood_flags(), o2.mood_flags())
    # the plateau test only exempts: on the same state it never flags more than MOODu1 on the first check
    _, o3 = _euler(L.HCC_STRAT_EPD1, L.HCC_CRIT_U2, plateau=1e-2)
    _, o4 = _euler(L.HCC_STRAT_EPD1, L.HCC_CRIT_U1)
    dt = wl["cfl"] * o3.time_step()
    o3.step(0.0, dt)
    o4.step(0.0, dt)
    assert o3.mood_flags().sum() <= o4.mood_flags().sum()


def test_epd2_drops_more_interfaces_than_epd1():
    """EPD2's interface orders are the minimum over both neighbourhoods (never above EPD1's edge minimum), so with
    the same orders the EPD2 right-hand side differs from EPD1's exactly on the particles within two hops of an
    order-1 particle."""
    o, x, r, idx = _line40()
    orders = np.full(o.N, 2, np.int32)
    orders[5] = 1
    for i in range(o.N):
        for j in idx[r[i]:r[i + 1]]:
            e1 = o.interface_orders(L.HCC_STRAT_EPD1, orders, i, j)
            e2 = o.interface_orders(L.HCC_STRAT_EPD2, orders, i, j)
            assert e2[0] <= e1[0]
            near = {5, *idx[r[5]:r[6]]}
            assert (e2[0] == 1) == (i in near or j in near)


# ---- MUSCL variants: streamed divergence, AlignedRusanovFlux (unit_schemes.jl:6-68, 397-427) --------------------
def _periodic_cloud(n, R_factor, jitter, seed, M, speeds, **kw):
    dom = W.domain(2, (0, 0), (1, 1), (True, True), 1 / n, R_factor, jitter, seed)
    c = O.cloud_generate(dom)
    sch = W.scheme(M, L.HCC_PDE_LINADV, adv_speed=speeds, limiter=L.HCC_LIM_NONE, strategy=L.HCC_STRAT_NONE,
                   criterion=L.HCC_CRIT_NONE, **kw)
    return O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"], sch, threads=4), c["pos"]


def _first_order_coefficients(flux):
    """K_i = sum_j a_ij (u_j - u_i) at order 1 everywhere (constant reconstructions): column j from a unit state at j."""
    o, _ = _periodic_cloud(24, 2.5, 0.2, 9, 1, [[1.0], [0.7]], flux=flux)
    r, idx, _, _ = o.neighbors()
    ones = np.ones(o.N, np.int32)
    a = []
    for j in range(o.N):
        U = np.zeros((o.N, 1))
        U[j] = 1.0
        K = o.rhs_at(U, ones)[:, 0]
        a.append(K)
    A = np.array(a).T                     # A[i, j] = a_ij
    return np.concatenate([A[i, idx[r[i]:r[i + 1]]] for i in range(o.N)])


def test_aligned_rusanov_first_order_coefficients_are_non_negative():
    """unit_schemes.jl:40-68: sorted by the signs of the divergence weights every first-order coefficient a_ij is
    non-negative (the update is local-extremum-diminishing); sorted by the distance some are negative on a jittered
    cloud with R = 2.5 dx."""
    a_al = _first_order_coefficients(O.FLUX_ALIGNED_RUSANOV)
    a_geo = _first_order_coefficients(O.FLUX_RUSANOV)
    assert len(a_al) > 1000 and a_al.min() >= -1e-12
    assert np.count_nonzero(a_geo < -1e-12) > 0


def test_streamed_divergence_matches_the_buffered_fit():
    """unit_schemes.jl:397-427: the streamed flux pass gives the divergence of the buffered fit up to round-off, with
    both fluxes, at full order and with a strip of particles dropped to order 1; at full order the aligned flux is the
    plain one bit for bit (no pair is below the full order)."""
    speeds = [[1.0, 0.5], [0.7, -0.3]]
    res = {}
    for flux in (O.FLUX_RUSANOV, O.FLUX_ALIGNED_RUSANOV):
        for streamed in (False, True):
            o, x = _periodic_cloud(24, 3.5, 0.2, 21, 2, speeds, flux=flux, streamed=streamed)
            U = np.column_stack([np.exp(-np.sum((x - 0.5) ** 2, axis=1) / 0.05),
                                 np.sin(2 * np.pi * x[:, 0]) * np.cos(2 * np.pi * x[:, 1])])
            for drop in (False, True):
                orders = np.full(o.N, 2, np.int32)
                if drop:
                    orders[np.abs(x[:, 0] - 0.5) < 0.1] = 1
                res[flux, streamed, drop] = o.rhs_at(U, orders)
    for flux in (O.FLUX_RUSANOV, O.FLUX_ALIGNED_RUSANOV):
        for drop in (False, True):
            Kb, Ks = res[flux, False, drop], res[flux, True, drop]
            scale = np.abs(Kb).max()
            assert np.abs(Kb - Ks).max() < 1e-11 * scale
            assert not np.array_equal(Kb, Ks)          # a different summation order: not the same bits
    for streamed in (False, True):
        assert np.array_equal(res[O.FLUX_RUSANOV, streamed, False], res[O.FLUX_ALIGNED_RUSANOV, streamed, False])
        assert not np.array_equal(res[O.FLUX_RUSANOV, streamed, True], res[O.FLUX_ALIGNED_RUSANOV, streamed, True])
