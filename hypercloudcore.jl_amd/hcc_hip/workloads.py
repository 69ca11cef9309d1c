"""Synthetic workloads of the explicit-step path: point clouds, initial conditions and scheme settings.

Each workload returns plain data (a domain dict, node arrays, an initial state, a scheme dict) so that the GPU
path and the CPU checker receive identical inputs.

* ``advection_2d``: 2-D linear advection, velocity (1, 1), periodic unit square, jittered h = 1/n lattice
  (jitter 0.25 h), Gaussian u = exp(-|x - (0.5, 0.5)|^2 / 0.01); MUSCL order 2, Venkatakrishnan(:hard), Rusanov,
  MOOD(EPD1, MOODu1(1e-4 h^2)), SSP-RK3, CFL 0.5. The reference constructs it as in benchmark/step.jl:56-75 and
  test/time_integration.jl:107-111.
* ``euler_2d``: 2-D compressible Euler (gamma = 1.4) on the unit square with outflow on all four faces
  (tags 1-4), jittered h lattice, four constant quadrants split at (0.5, 0.5) whose rho ~ U[0.1, 1.5],
  p ~ U[0.03, 1.5], u, v ~ U[-1, 1] are drawn from the seed; same scheme, CFL 0.4.
* ``euler_3d``: 3-D compressible Euler on the periodic unit cube, rest state rho = 1, p = 1 with a Gaussian pressure
  pulse of amplitude 0.5 and width 0.05 at the centre, R = 2.2 h.
"""
import math

import numpy as np

from . import _lib as L

# ---- explicit tableaus (src/TimeStepping/ButcherTableaus.jl:124-147) ----------------------------------------
TABLEAUS = {
    "RK1_Euler": ([[0.0]], [1.0], [0.0]),
    "RK2_SSP": ([[0.0, 0.0], [1.0, 0.0]], [1 / 2, 1 / 2], [0.0, 1.0]),
    "RK2_Ralston": ([[0.0, 0.0], [2 / 3, 0.0]], [1 / 4, 3 / 4], [0.0, 2 / 3]),
    "RK3_SSP": ([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.25, 0.25, 0.0]], [1 / 6, 1 / 6, 2 / 3], [0.0, 1.0, 0.5]),
    "RK4_Classical": ([[0.0, 0.0, 0.0, 0.0], [0.5, 0.0, 0.0, 0.0], [0.0, 0.5, 0.0, 0.0], [0.0, 0.0, 1.0, 0.0]],
                      [1 / 6, 1 / 3, 1 / 3, 1 / 6], [0.0, 0.5, 0.5, 1.0]),
}


def _imex_tableaus():
    """IMEX tableaus (ButcherTableaus.jl:183-264): explicit part (a~, b~, c~) and implicit part (a, b, c), entries
    computed with the reference's Float64 expressions."""
    g2 = 1 - 1 / math.sqrt(2)                      # ARS222, PRSSP3: 1 - 1/sqrt(2)
    d2 = 1 - 1 / (2 * g2)
    g3 = (3.0 + math.sqrt(3.0)) / 6.0             # ARS233
    return {
        "IMEX_Euler": dict(a_t=[[0.0, 0.0], [1.0, 0.0]], b_t=[1.0, 0.0], c_t=[0.0, 1.0],
                           a_im=[[0.0, 0.0], [0.0, 1.0]], b_im=[0.0, 1.0], c_im=[0.0, 1.0]),
        "IMEX_ARS233": dict(a_t=[[0.0, 0.0, 0.0], [g3, 0.0, 0.0], [g3 - 1.0, 2.0 * (1 - g3), 0.0]],
                            b_t=[0.0, 0.5, 0.5], c_t=[0.0, g3, 1.0 - g3],
                            a_im=[[0.0, 0.0, 0.0], [0.0, g3, 0.0], [0.0, 1.0 - 2 * g3, g3]],
                            b_im=[0.0, 0.5, 0.5], c_im=[0.0, g3, 1.0 - g3]),
        "IMEX_ARS222": dict(a_t=[[0.0, 0.0, 0.0], [g2, 0.0, 0.0], [d2, 1 - d2, 0.0]], b_t=[d2, 1 - d2, 0.0],
                            c_t=[0.0, g2, 1.0], a_im=[[0.0, 0.0, 0.0], [0.0, g2, 0.0], [0.0, 1 - g2, g2]],
                            b_im=[0.0, 1 - g2, g2], c_im=[0.0, g2, 1.0]),
        "IMEX_PRSSP3": dict(a_t=[[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.25, 0.25, 0.0]], b_t=[1 / 6, 1 / 6, 2 / 3],
                            c_t=[0.0, 1.0, 0.5], a_im=[[g2, 0.0, 0.0], [1 - 2 * g2, g2, 0.0], [0.5 - g2, 0.0, g2]],
                            b_im=[1 / 6, 1 / 6, 2 / 3], c_im=[g2, 1 - g2, 0.5]),
        "IMEX_SSP2332": dict(a_t=[[0.0, 0.0, 0.0], [0.5, 0.0, 0.0], [0.5, 0.5, 0.0]], b_t=[1 / 3, 1 / 3, 1 / 3],
                             c_t=[0.0, 0.5, 1.0], a_im=[[0.25, 0.0, 0.0], [0.0, 0.25, 0.0], [1 / 3, 1 / 3, 1 / 3]],
                             b_im=[1 / 3, 1 / 3, 1 / 3], c_im=[0.25, 0.25, 1.0]),
    }


IMEX_TABLEAUS = _imex_tableaus()


def ssp_coefficient(a_t, b_t):
    """SSP coefficient by bisection on the canonical Shu-Osher form (ButcherTableaus.jl:101-118).

    alpha_r = r K (I + r K)^-1 with K the Butcher matrix with b as its last row; the largest r for which
    alpha_r and (I - alpha_r) e have no entry below -1e-13. The LU rounding of Julia's right division is not
    reproduced bit for bit; the time stepper takes the resulting ratio 1/C as an input.
    """
    a_t = np.asarray(a_t, float)
    s = len(b_t)
    K = np.zeros((s + 1, s + 1))
    K[:s, :s] = a_t
    K[s, :s] = b_t
    I = np.eye(s + 1)

    def ok(r):
        A = r * K
        alpha = np.linalg.solve((I + A).T, A.T).T
        return alpha.min() >= -1e-13 and ((I - alpha) @ np.ones(s + 1)).min() >= -1e-13

    r_lo = r_hi = 1e-3
    if not ok(r_lo):
        return 0.0
    while ok(r_hi) and r_hi < 1e3:
        r_hi *= 2
    r_lo = r_hi / 2
    for _ in range(60):
        mid = (r_lo + r_hi) / 2
        if ok(mid):
            r_lo = mid
        else:
            r_hi = mid
    return r_lo


def tableau(name):
    if name in IMEX_TABLEAUS:   # an implicit part: SSP coefficient 0 (ButcherTableaus.jl:20-21), no substep check
        return dict(IMEX_TABLEAUS[name], ssp=0.0, substep_ratio=0.0)
    a_t, b_t, c_t = TABLEAUS[name]
    C = ssp_coefficient(a_t, b_t)
    return {"a_t": a_t, "b_t": b_t, "c_t": c_t, "ssp": C, "substep_ratio": (1.0 / C) if C > 0 else 0.0}


def scheme(M, pde, *, gamma=1.4, adv_speed=None, limiter=L.HCC_LIM_VK, mode=L.HCC_MODE_HARD, strategy=L.HCC_STRAT_EPD1,
           criterion=L.HCC_CRIT_U1, delta=0.0, plateau=0.0, tableau_name="RK3_SSP", bcs=(), mood_substeps=True,
           streamed=False, flux=L.HCC_FLUX_RUSANOV, mover=L.HCC_MOVER_NONE, mover_alpha=1.0, order=2, ratio=0.5):
    tb = tableau(tableau_name)
    return {"M": M, "pde": pde, "gamma": gamma, "adv_speed": adv_speed or [], "muscl_order": order, "limiter": limiter,
            "limiter_mode": mode, "strategy": strategy, "criterion": criterion, "delta": delta, "plateau": plateau,
            "ratio": ratio,
            "streamed": streamed, "flux": flux, "a_t": tb["a_t"], "a_im": tb.get("a_im"), "b_im": tb.get("b_im"),
            "c_im": tb.get("c_im"), "mover": mover, "mover_alpha": mover_alpha,
            "b_t": tb["b_t"], "c_t": tb["c_t"], "substep_ratio": tb["substep_ratio"] if mood_substeps else 0.0,
            "bcs": list(bcs)}


def domain(D, mins, maxs, periodic, h, R_factor, jitter_factor, seed, alpha=1.0):
    dx = [h] * D
    return {"D": D, "mins": list(mins), "maxs": list(maxs), "periodic": list(periodic), "dx": dx, "R": R_factor * h,
            "alpha": alpha, "jitter": [jitter_factor * h] * D, "seed": seed}


def domain_desc(dom):
    return L.domain_desc(dom["D"], dom["mins"], dom["maxs"], dom["periodic"], dom["dx"], dom["R"], dom["alpha"],
                         dom["jitter"], dom["seed"])


def euler_conserved(rho, u, p, gamma):
    """(rho, rho u_1..u_D, E) with E = p / (gamma - 1) + 0.5 rho |u|^2 (u: (N, D))."""
    u = np.atleast_2d(u)
    ke = 0.5 * rho * np.sum(u * u, axis=1)
    return np.column_stack([rho, rho[:, None] * u, p / (gamma - 1.0) + ke])


def advection_2d(n=64, seed=0):
    h = 1.0 / n
    dom = domain(2, (0.0, 0.0), (1.0, 1.0), (True, True), h, 2.5, 0.25, seed)
    sch = scheme(1, L.HCC_PDE_LINADV, adv_speed=[[1.0], [1.0]], delta=1e-4 * h * h)
    cloud = L.cloud_generate(domain_desc(dom))
    x = cloud["pos"]
    U0 = np.exp(-((x[:, 0] - 0.5) ** 2 + (x[:, 1] - 0.5) ** 2) / 0.01)[:, None]
    return {"name": f"advection2d_n{n}", "domain": dom, "scheme": sch, "cloud": cloud, "U0": U0, "cfl": 0.5,
            "t_max": 0.5}


def quadrant_states(seed, gamma=1.4):
    """rho, p, u, v of the four quadrants (index q = [x >= 0.5] + 2 [y >= 0.5]) from the seed."""
    r = L.rng_uniform(seed + 0x51, 16).reshape(4, 4)
    rho = 0.1 + 1.4 * r[:, 0]
    p = 0.03 + 1.47 * r[:, 1]
    u = -1.0 + 2.0 * r[:, 2]
    v = -1.0 + 2.0 * r[:, 3]
    return rho, p, u, v


def euler_2d(h=1e-3, seed=0, gamma=1.4, n=None):
    if n is not None:
        h = 1.0 / n
    dom = domain(2, (0.0, 0.0), (1.0, 1.0), (False, False), h, 2.5, 0.25, seed)
    bcs = [(t, L.HCC_BC_OUTFLOW) for t in (1, 2, 3, 4)]
    sch = scheme(4, L.HCC_PDE_EULER, gamma=gamma, delta=1e-4 * h * h, bcs=bcs)
    cloud = L.cloud_generate(domain_desc(dom))
    x = cloud["pos"]
    rho_q, p_q, u_q, v_q = quadrant_states(seed, gamma)
    q = (x[:, 0] >= 0.5).astype(int) + 2 * (x[:, 1] >= 0.5).astype(int)
    U0 = euler_conserved(rho_q[q], np.column_stack([u_q[q], v_q[q]]), p_q[q], gamma)
    return {"name": f"euler2d_h{h:g}", "domain": dom, "scheme": sch, "cloud": cloud, "U0": U0, "cfl": 0.4, "steps": 200}


def euler_2d_box(n, h, seed=0, gamma=1.4):
    """The four-quadrant Euler problem on the box [0, n h]^2 with outflow ghosts: the config of euler_2d with the
    spacing h shared by every box and the split at the box centre. An ensemble member (boxes of one h and different n
    share every kernel constant)."""
    L_box = n * h
    dom = domain(2, (0.0, 0.0), (L_box, L_box), (False, False), h, 2.5, 0.25, seed)
    bcs = [(t, L.HCC_BC_OUTFLOW) for t in (1, 2, 3, 4)]
    sch = scheme(4, L.HCC_PDE_EULER, gamma=gamma, delta=1e-4 * h * h, bcs=bcs)
    cloud = L.cloud_generate(domain_desc(dom))
    x = cloud["pos"]
    rho_q, p_q, u_q, v_q = quadrant_states(seed, gamma)
    q = (x[:, 0] >= 0.5 * L_box).astype(int) + 2 * (x[:, 1] >= 0.5 * L_box).astype(int)
    U0 = euler_conserved(rho_q[q], np.column_stack([u_q[q], v_q[q]]), p_q[q], gamma)
    return {"name": f"euler2d_box_n{n}", "domain": dom, "scheme": sch, "cloud": cloud, "U0": U0, "cfl": 0.4}


def ensemble_sizes(count, n_min, n_max, seed=0):
    """Lattice sides of `count` ensemble members with node counts log-uniform in [n_min, n_max] (seeded)."""
    r = L.rng_uniform(seed + 0xE5, count)
    nodes = np.exp(np.log(n_min) + (np.log(n_max) - np.log(n_min)) * r)
    return np.maximum(np.rint(np.sqrt(nodes)).astype(int), 4)


def euler_2d_ensemble(count, n_min, n_max, h=1e-3, seed=0):
    """`count` four-quadrant Euler boxes of spacing h with node counts log-uniform in [n_min, n_max]; member k has
    seed seed + 1 + k (its own quadrant states and jitter)."""
    return [euler_2d_box(int(n), h, seed + 1 + k) for k, n in enumerate(ensemble_sizes(count, n_min, n_max, seed))]


def euler_3d(n=160, seed=0, gamma=1.4):
    h = 1.0 / n
    dom = domain(3, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (True, True, True), h, 2.2, 0.25, seed)
    sch = scheme(5, L.HCC_PDE_EULER, gamma=gamma, delta=1e-4 * h ** 3)
    cloud = L.cloud_generate(domain_desc(dom))
    x = cloud["pos"]
    r2 = np.sum((x - 0.5) ** 2, axis=1)
    p = 1.0 + 0.5 * np.exp(-r2 / (0.05 ** 2))
    U0 = euler_conserved(np.ones(len(x)), np.zeros((len(x), 3)), p, gamma)
    return {"name": f"euler3d_n{n}", "domain": dom, "scheme": sch, "cloud": cloud, "U0": U0, "cfl": 0.3, "steps": 100}


def step_benchmark(n=1000, M=1, stab="both", D=2):
    """The reference's own step benchmark (benchmark/step.jl:15-80): linear advection of a diagonal system of M
    equations (speed (1 - 0.5 (d - 1)) (1 + 0.2 (m - 1)) along axis d for equation m) with MUSCL3 + RusanovFlux,
    SSP-RK3, on the uniform periodic n^D lattice of the unit box, ExponentialWeightFunction(1, 2.5 dx); the state
    exp(-|x - 0.5|^2 / 0.02) (1 + 0.1 (m - 1)); stab "both": VenkatakrishnanLimiter(:hard) + MOOD(EPD1, MOODu1(1e-6)),
    "limiter", "mood" or "none"; dt = 0.3 get_time_step at t = 0, fixed."""
    h = 1.0 / n
    dom = domain(D, (0.0,) * D, (1.0,) * D, (True,) * D, h, 2.5, 0.0, 0)
    speeds = [[(1 - 0.5 * d) * (1 + 0.2 * m) for m in range(M)] for d in range(D)]
    lim = L.HCC_LIM_VK if stab in ("limiter", "both") else L.HCC_LIM_NONE
    mood = stab in ("mood", "both")
    sch = scheme(M, L.HCC_PDE_LINADV, adv_speed=speeds, limiter=lim, mode=L.HCC_MODE_HARD,
                 strategy=L.HCC_STRAT_EPD1 if mood else L.HCC_STRAT_NONE,
                 criterion=L.HCC_CRIT_U1 if mood else L.HCC_CRIT_NONE, delta=1e-6, order=3)
    cloud = L.cloud_generate(domain_desc(dom))
    x = cloud["pos"]
    g = np.exp(-np.sum((x - 0.5) ** 2, axis=1) / 0.02)
    U0 = np.column_stack([g * (1 + 0.1 * m) for m in range(M)])
    return {"name": f"step_benchmark_{D}d_n{n}_m{M}_{stab}", "domain": dom, "scheme": sch, "cloud": cloud, "U0": U0,
            "dt_factor": 0.3, "steps": 30, "warmup": 3}


def mean_neighbours(grid):
    return grid.S / max(grid.N, 1)


def bytes_per_node_stage(D, M, nnb, B_g=None, B_d=None):
    """Algorithmic HBM bytes per node-stage of the reference dataflow (int32 indices, perfect neighbour reuse):
    gather + flux passes, RK3 stage combine, MOOD check and the CFL estimate (a third per stage)."""
    B_g = D if B_g is None else B_g
    B_d = D + 1 if B_d is None else B_d
    gather = 8 * M + nnb * (4 + 8 * D + 8) + 8 * B_g * B_g + 8 * B_g * M
    flux = 8 * M + 8 * B_g * M + nnb * (4 + 8 * D + 8) + 8 * B_d * B_d + 8 * M
    combine = 8 * M * 14 / 3
    mood = 8 * M * 5 + 4 * nnb
    cfl = (8 * M + nnb * (4 + 8 * D)) / 3
    return gather + flux + combine + mood + cfl


def isclose_rel_linf(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


__all__ = ["TABLEAUS", "ssp_coefficient", "tableau", "scheme", "domain", "domain_desc", "advection_2d", "euler_2d",
           "euler_3d", "quadrant_states", "euler_conserved", "bytes_per_node_stage", "isclose_rel_linf", "math"]
