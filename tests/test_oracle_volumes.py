"""Pins of the oracle's particle volumes (ShepardVolumes), conservation diagnostics (total_mass / interior_mass) and
the driver loop (solve_equation) against the reference's unit-test expectations, an independent numpy restatement
and closed forms."""
import numpy as np
import pytest

from hcc_hip import _lib as L
from hcc_hip import workloads as W
from oracle import oracle as O


def _line(n, periodic=True, jitter=0.0, seed=0):
    dom = W.domain(1, (-1,), (1,), (periodic,), 2 / n, 2.5, jitter, seed)
    c = O.cloud_generate(dom)
    o = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"], W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0]]))
    return o, c


def test_uniform_periodic_line_volumes_equal_dx():
    """unit_schemes.jl:914-923: on a uniform periodic cloud every Shepard volume is dx and they sum to |Omega|."""
    o, c = _line(50)
    v = o.volumes()
    assert np.allclose(v, 2 / 50, rtol=1e-12, atol=0)
    assert v.sum() == pytest.approx(2.0, rel=1e-14)
    v3 = o.volumes(3)
    assert np.allclose(v3, 2 / 50, rtol=1e-12, atol=0) and v3.sum() == pytest.approx(2.0, rel=1e-14)


def test_total_mass_of_a_constant():
    """unit_schemes.jl:932-935: total_mass of U = 2 on the 40-particle line [-1, 1] is 4."""
    o, c = _line(40)
    o.set_scheme(W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0]]))
    o.set_state(np.full((o.N, 1), 2.0))
    assert o.mass()[0] == pytest.approx(4.0, rel=1e-14)
    assert np.array_equal(o.mass(True), o.mass())   # periodic: no ghosts


def _numpy_shepard(o, sweeps, dom_volume):
    """Volumes.jl:183-231 restated with numpy from the neighbour lists and stored weights (independent of the C++)."""
    r, idx, w, _ = o.neighbors()
    N = o.N
    owner = np.repeat(np.arange(N), np.diff(r))
    S = 1.0 + np.bincount(owner, weights=w, minlength=N)   # w(0) = 1 for the exponential kernel
    V = 1.0 / S
    for _ in range(sweeps):
        V = (1.0 / S) * V + np.bincount(owner, weights=w / S[idx] * V[idx], minlength=N)
    V = np.maximum(V, 0.0)
    return V * (dom_volume / V.sum())


@pytest.mark.parametrize("sweeps", [0, 1, 3])
def test_shepard_volumes_match_an_independent_restatement(sweeps):
    """A jittered bounded 2-D cloud with its ghost band: the oracle's volumes equal a numpy restatement of the
    Shepard partition to rounding (the sums differ in order only); Sigma V = |Omega| and V >= 0."""
    wl = W.euler_2d(n=24, seed=3)
    c = wl["cloud"]
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], wl["scheme"])
    v = o.volumes(sweeps)
    ref = _numpy_shepard(o, sweeps, 1.0)
    assert np.max(np.abs(v - ref) / ref) < 1e-13
    assert v.sum() == pytest.approx(1.0, rel=1e-13) and np.all(v >= 0)


def test_shepard_quadrature_on_a_jittered_line():
    """Volumes.jl:36-50 and unit_schemes.jl:880-890: on a jittered bounded line the ghosts get volumes too, so
    sum V = |Omega| = 2 exactly up to rounding while sum V x^2 converges to 2/3 at first order only."""
    errs = []
    for n in (80, 160, 320, 640):
        o, c = _line(n, periodic=False, jitter=0.3, seed=1)
        v = o.volumes()
        assert v.sum() == pytest.approx(2.0, rel=1e-13)
        errs.append(abs(np.sum(v * c["pos"][:, 0] ** 2) - 2 / 3))
    rate = np.log2(errs[0] / errs[-1]) / 3
    assert 0.6 < rate < 1.6 and errs[-1] < 0.02


def test_interior_mass_excludes_ghosts():
    """Volumes.jl:103-117: total_mass sums V U over every particle, interior_mass skips the ghosts."""
    wl = W.euler_2d(n=24, seed=1)
    c = wl["cloud"]
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], wl["scheme"])
    o.set_state(wl["U0"])
    v = o.volumes()
    isb = c["is_boundary"].astype(bool)
    tot, inner = o.mass(), o.mass(True)
    assert np.allclose(tot, (v[:, None] * wl["U0"]).sum(0), rtol=1e-13)
    assert np.allclose(inner, (v[~isb, None] * wl["U0"][~isb]).sum(0), rtol=1e-13)
    assert inner[0] < tot[0] and inner[3] < tot[3]   # density and energy: the ghosts hold positive content


def test_snapshot_times_are_correctly_rounded():
    """range(0, 0.5, length = 11) holds the doubles nearest k / 20 (TwicePrecision), not k * (0.5 / 10): the third
    element is 0.15, where 3 * 0.05 rounds to 0.15000000000000002."""
    t = O.snapshot_times(0.5, 10)
    assert t == [float(f"0.{5 * k:02d}") if k < 10 else 0.5 for k in range(11)][:10] + [0.5]
    assert 3 * (0.5 / 10) != t[3] and t[3] == 0.15


def test_solve_equation_driver_semantics():
    """test/time_integration.jl:72-132 on the periodic advection cloud: the first snapshot is the initial state, each
    later one the first step reaching its time, the last the end state at t = tmax; dt is the CFL step capped at
    tmax - t."""
    wl = W.advection_2d(24, 0)
    c = wl["cloud"]
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], wl["scheme"], threads=4)
    o.set_state(wl["U0"])
    xs, us, ts, k, ms = O.solve_equation(o, 0.2, wl["cfl"], True, 4)
    t_snap = O.snapshot_times(0.2, 4)
    assert ts[0] == 0.0 and ts[-1] == 0.2 and np.array_equal(us[0], wl["U0"])
    assert np.array_equal(us[-1], o.state())
    for c_, t in enumerate(ts[1:-1], start=1):
        assert t >= t_snap[c_] and t - t_snap[c_] < 0.2 / 4
    assert k > 4 and all(np.isfinite(m[0]).all() for m in ms)


def test_solve_equation_conserves_mass_on_a_uniform_periodic_lattice():
    """pipeline_tests.jl:43-64: on a uniform periodic lattice (equal volumes, symmetric weights) the discrete content
    sum V U is conserved by the flux-form step, snapshot after snapshot."""
    h = 1 / 24
    dom = W.domain(2, (0, 0), (1, 1), (True, True), h, 2.5, 0.0, 0)
    c = O.cloud_generate(dom)
    sch = W.scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0], [1.0]], delta=1e-4 * h * h)
    o = O.OracleSim(dom, c["pos"], c["is_boundary"], c["tags"], sch, threads=4)
    x = c["pos"]
    o.set_state(np.exp(-((x[:, 0] - 0.5) ** 2 + (x[:, 1] - 0.5) ** 2) / 0.01)[:, None])
    xs, us, ts, k, ms = O.solve_equation(o, 0.1, 0.5, True, 3)
    m0 = ms[0][0][0]
    assert all(abs(m[0][0] / m0 - 1) < 1e-12 for m in ms)
