"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it as the checker
of the HIP path (and as the timed CPU baseline). Nothing in the product imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built: run oracle/build_oracle.py (or __graft_entry__.build())")
        _lib = C.CDLL(LIB)
        vp, i64, i32, d = C.c_void_p, C.c_int64, C.c_int32, C.c_double
        P = C.POINTER
        sig = {
            "orc_cloud_generate": (i64, [C.c_int, vp, vp, vp, vp, d, vp, C.c_uint64, i64, vp, vp, vp, P(i64), vp, P(i64)]),
            "orc_sim_new": (vp, [C.c_int, C.c_int, i64, vp, vp, vp, vp, vp, vp, vp, d, d, C.c_int]),
            "orc_sim_free": (None, [vp]),
            "orc_sim_set_threads": (None, [vp, C.c_int]),
            "orc_sim_set_scheme": (None, [vp, C.c_int, d, vp, C.c_int, C.c_int, C.c_int, C.c_int, d, C.c_int, vp, vp, vp,
                                          d, C.c_int, vp, vp, d, C.c_int, C.c_int, vp, vp, vp, C.c_int, d, d]),
            "orc_sim_set_state": (None, [vp, vp]),
            "orc_sim_get_state": (None, [vp, vp]),
            "orc_sim_time_step": (d, [vp]),
            "orc_sim_step": (None, [vp, d, d]),
            "orc_sim_rhs": (None, [vp, vp, vp, vp]),
            "orc_sim_rhs_orders": (None, [vp, vp, vp, vp]),
            "orc_sim_num_slots": (i64, [vp]),
            "orc_sim_get_neighbors": (None, [vp, vp, vp, vp, vp]),
            "orc_sim_get_moments": (None, [vp, vp, vp, vp, vp, vp]),
            "orc_sim_get_orders": (None, [vp, vp]),
            "orc_sim_get_flags": (None, [vp, vp]),
            "orc_sim_inadmissible": (i64, [vp]),
            "orc_sim_mood_iterations": (i64, [vp]),
            "orc_sim_get_bins": (None, [vp, vp, vp, vp, vp]),
            "orc_sim_interface_orders": (None, [vp, C.c_int, vp, i64, i64, vp]),
            "orc_sim_order_reach": (C.c_int, [vp, C.c_int]),
            "orc_sim_force": (None, [vp, vp, vp, vp, vp]),
            "orc_sim_expand": (None, [vp, vp, C.c_int]),
            "orc_sim_u2_flags": (C.c_int, [vp, vp, i64, vp, d, d]),
            "orc_sim_u2_flags3": (C.c_int, [vp, vp, i64, vp, d, d, d, vp, vp]),
            "orc_sim_volumes": (None, [vp, C.c_int, vp]),
            "orc_sim_mass": (None, [vp, C.c_int, vp]),
            "orc_sim_set_source": (None, [vp, C.c_int, vp]),
            "orc_sim_get_stage": (None, [vp, C.c_int, C.c_int, vp]),
            "orc_sim_get_positions": (None, [vp, vp]),
            "orc_sim_set_positions": (None, [vp, vp, C.c_int]),
            "orc_weight": (d, [d, d, d]),
            "orc_limiter_phi": (d, [C.c_int, d]),
            "orc_jmax": (d, [d, d]),
            "orc_jmin": (d, [d, d]),
            "orc_jmod": (d, [d, d]),
            "orc_spd_inverse": (C.c_int, [C.c_int, vp, vp]),
            "orc_euler": (None, [C.c_int, d, vp, vp, vp, vp]),
            "orc_rng_uniform": (None, [C.c_uint64, i64, vp]),
            "orc_num_threads_max": (C.c_int, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(_lib, name)
            f.restype, f.argtypes = res, args
    return _lib


FLUX_RUSANOV, FLUX_ALIGNED_RUSANOV = 0, 1   # scheme["flux"]: RusanovFlux, AlignedRusanovFlux


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _arr(v, dt, n=3):
    a = np.zeros(n, dt)
    a[:len(v)] = v
    return a


def cloud_generate(dom):
    D = dom["D"]
    mins, maxs = _arr(dom["mins"], np.float64), _arr(dom["maxs"], np.float64)
    per = _arr([int(bool(x)) for x in dom["periodic"]], np.int32)
    dx, jit = _arr(dom["dx"], np.float64), _arr(dom["jitter"], np.float64)
    n = C.c_int64(0)
    ties = C.c_int64(0)
    canvas = np.zeros(2 * D)
    L = lib()
    L.orc_cloud_generate(D, _p(mins), _p(maxs), _p(per), _p(dx), dom["R"], _p(jit), dom["seed"], 0, None, None, None,
                         C.byref(n), _p(canvas), C.byref(ties))
    N = n.value
    pos = np.zeros((N, D))
    isb = np.zeros(N, np.uint8)
    tags = np.zeros(N, np.int32)
    L.orc_cloud_generate(D, _p(mins), _p(maxs), _p(per), _p(dx), dom["R"], _p(jit), dom["seed"], N, _p(pos), _p(isb),
                         _p(tags), C.byref(n), _p(canvas), C.byref(ties))
    return {"pos": pos, "is_boundary": isb, "tags": tags, "canvas": canvas, "morton_ties": ties.value, "N": N}


def basis_count(D, q):
    """Terms of degree 1..q in D variables, binomial(D + q, q) - 1 (Interpolators.jl:170-177)."""
    if q <= 0:
        return 0
    n = 1
    for k in range(1, q + 1):
        n = (n * (D + k)) // k
    return n - 1


class OracleSim:
    """Grid + MUSCL (order 2 or 3) + GeneralRKTimeStepper restated on the CPU."""

    def __init__(self, dom, pos, isb, tags, scheme=None, threads=1):
        D = dom["D"]
        self.D = D
        self.N = pos.shape[0]
        self.M = scheme["M"] if scheme else 1
        self.order = int(scheme.get("muscl_order", 2)) if scheme else 2
        if self.order == 3 and scheme and (scheme.get("streamed") or scheme.get("flux", FLUX_RUSANOV) != FLUX_RUSANOV
                                           or scheme.get("mover", 0)):
            raise ValueError("the MUSCL3 restatement covers the buffered RusanovFlux on static grids")
        # basis sizes: reconstruction of degree order - 1, divergence (with intercept) of degree 1 or 3
        self.BG = basis_count(D, self.order - 1)
        self.BD = basis_count(D, 1 if self.order <= 2 else 3) + 1
        self._keep = [np.ascontiguousarray(pos, np.float64), np.ascontiguousarray(isb, np.uint8),
                      np.ascontiguousarray(tags, np.int32)]
        mins, maxs = _arr(dom["mins"], np.float64), _arr(dom["maxs"], np.float64)
        per = _arr([int(bool(x)) for x in dom["periodic"]], np.int32)
        dx = _arr(dom["dx"], np.float64)
        self.h = lib().orc_sim_new(D, self.M, self.N, _p(self._keep[0]), _p(self._keep[1]), _p(self._keep[2]), _p(mins),
                                   _p(maxs), _p(per), _p(dx), dom["R"], dom["alpha"], self.order)
        if not self.h:
            raise ValueError(f"oracle has no (D={D}, M={self.M}, order {self.order}) variant")
        lib().orc_sim_set_threads(self.h, int(threads))
        if scheme:
            self.set_scheme(scheme)

    def set_scheme(self, sch):
        s = len(sch["b_t"])
        a_t = np.zeros((s, s))
        for i in range(s):
            for j in range(s):
                a_t[i, j] = sch["a_t"][i][j]
        b_t, c_t = np.array(sch["b_t"], float), np.array(sch["c_t"], float)
        a_im = np.array(sch.get("a_im") or np.zeros((s, s)), float).reshape(s, s).copy()
        b_im = np.array(sch.get("b_im") or np.zeros(s), float)
        c_im = np.array(sch.get("c_im") or np.zeros(s), float)
        c = np.zeros((3, 8))
        for d, row in enumerate(sch.get("adv_speed", [])):
            c[d, :len(row)] = row
        bcs = sch.get("bcs", [])
        tags = np.array([t for t, _ in bcs] or [0], np.int32)
        kinds = np.array([k for _, k in bcs] or [0], np.int32)
        self._keep += [a_t, b_t, c_t, c, tags, kinds, a_im, b_im, c_im]
        lib().orc_sim_set_scheme(self.h, sch["pde"], sch.get("gamma", 1.4), _p(c), sch["limiter"],
                                 sch.get("limiter_mode", 1), sch["strategy"], sch["criterion"], sch.get("delta", 0.0), s,
                                 _p(a_t), _p(b_t), _p(c_t), sch.get("substep_ratio", 0.0), len(bcs), _p(tags), _p(kinds),
                                 sch.get("plateau", 0.0), int(bool(sch.get("streamed", False))),
                                 int(sch.get("flux", FLUX_RUSANOV) == FLUX_ALIGNED_RUSANOV), _p(a_im), _p(b_im),
                                 _p(c_im), int(sch.get("mover", 0)), float(sch.get("mover_alpha", 1.0)),
                                 float(sch.get("ratio", 0.5)))

    def set_threads(self, n):
        lib().orc_sim_set_threads(self.h, int(n))

    def set_state(self, U):
        U = np.ascontiguousarray(U, np.float64).reshape(self.N, self.M)
        lib().orc_sim_set_state(self.h, _p(U))

    def state(self):
        U = np.zeros((self.N, self.M))
        lib().orc_sim_get_state(self.h, _p(U))
        return U

    def time_step(self):
        return lib().orc_sim_time_step(self.h)

    def step(self, t, dt):
        lib().orc_sim_step(self.h, float(t), float(dt))

    def rhs(self, Y):
        Y = np.ascontiguousarray(Y, np.float64).reshape(self.N, self.M)
        K = np.zeros((self.N, self.M))
        G = np.zeros((self.N, self.BG, self.M))
        lib().orc_sim_rhs(self.h, _p(Y), _p(K), _p(G))
        return K, G

    def rhs_at(self, Y, orders):
        """K_E of a full gather + flux pass with the given particle orders."""
        Y = np.ascontiguousarray(Y, np.float64).reshape(self.N, self.M)
        o = np.ascontiguousarray(orders, np.int32)
        K = np.zeros((self.N, self.M))
        lib().orc_sim_rhs_orders(self.h, _p(Y), _p(o), _p(K))
        return K

    def neighbors(self):
        S = lib().orc_sim_num_slots(self.h)
        ranges = np.zeros(self.N + 1, np.int64)
        idx = np.zeros(max(S, 1), np.int32)
        w = np.zeros(max(S, 1))
        dist = np.zeros((max(S, 1), self.D))
        lib().orc_sim_get_neighbors(self.h, _p(ranges), _p(idx), _p(w), _p(dist))
        return ranges, idx[:S], w[:S], dist[:S]

    def moments(self):
        D, N = self.D, self.N
        S = lib().orc_sim_num_slots(self.h)
        ig = np.zeros((N, self.BG, self.BG))
        og = np.zeros(N, np.uint8)
        idv = np.zeros((N, self.BD, self.BD))
        od = np.zeros(N, np.uint8)
        dw = np.zeros((max(S, 1), D))
        lib().orc_sim_get_moments(self.h, _p(ig), _p(og), _p(idv), _p(od), _p(dw))
        return ig, og, idv, od, dw[:S]

    def bins(self):
        dims = np.zeros(4, np.int64)
        lib().orc_sim_get_bins(self.h, None, None, None, _p(dims))
        nb = int(dims[3])
        bo = np.zeros(self.N, np.int32)
        cs = np.zeros(nb + 2, np.int64)
        cp = np.zeros(self.N, np.int32)
        lib().orc_sim_get_bins(self.h, _p(bo), _p(cs), _p(cp), _p(dims))
        return bo, cs, cp, dims

    def orders(self):
        o = np.zeros(self.N, np.int32)
        lib().orc_sim_get_orders(self.h, _p(o))
        return o

    def mood_flags(self):
        f = np.zeros(self.N, np.uint8)
        lib().orc_sim_get_flags(self.h, _p(f))
        return f

    def inadmissible(self):
        return lib().orc_sim_inadmissible(self.h)

    # MOOD building blocks on their own (the reference's unit tests call the same functions)
    def interface_orders(self, strategy, orders, i, j):
        o = np.ascontiguousarray(orders, np.int32)
        out = np.zeros(2, np.int32)
        lib().orc_sim_interface_orders(self.h, strategy, _p(o), i, j, _p(out))
        return int(out[0]), int(out[1])

    def order_reach(self, strategy):
        return lib().orc_sim_order_reach(self.h, strategy)

    def force(self, triggered, snapshot, orders, changed):
        t = np.ascontiguousarray(triggered, np.uint8)
        sn = np.ascontiguousarray(snapshot, np.int32)
        o = np.array(orders, np.int32)
        c = np.array(changed, np.uint8)
        lib().orc_sim_force(self.h, _p(t), _p(sn), _p(o), _p(c))
        return o, c.astype(bool)

    def expand(self, mask, hops):
        m = np.array(mask, np.uint8)
        lib().orc_sim_expand(self.h, _p(m), hops)
        return m.astype(bool)

    def u2_flags(self, U, p, cand, delta, plateau=0.0):
        U = np.ascontiguousarray(U, np.float64).reshape(self.N, self.M)
        c = np.ascontiguousarray(cand, np.float64)
        return bool(lib().orc_sim_u2_flags(self.h, _p(U), p, _p(c), delta, plateau))

    def u2_flags3(self, U, p, cand, delta, plateau, ratio, grads, orders):
        """The MOODu2 check of particle p with the given gradient records (N x BG x M) and particle orders."""
        U = np.ascontiguousarray(U, np.float64)
        cand = np.ascontiguousarray(cand, np.float64)
        g = np.ascontiguousarray(grads, np.float64)
        o = np.ascontiguousarray(orders, np.int32)
        return bool(lib().orc_sim_u2_flags3(self.h, _p(U), int(p), _p(cand), float(delta), float(plateau),
                                            float(ratio), _p(g), _p(o)))

    def set_source(self, mode, rates):
        """The decay source S_i(U) = -rates_i U: mode 1 explicit, 2 implicit, 0 none."""
        r = np.ascontiguousarray(rates, np.float64)
        self._keep.append(r)
        lib().orc_sim_set_source(self.h, int(mode), _p(r))

    def stage(self, which, i):
        """Stage buffer of the last step: which 'Y', 'K_E' or 'K_I', stage i (0-based)."""
        out = np.zeros((self.N, self.M))
        lib().orc_sim_get_stage(self.h, {"Y": 0, "K_E": 1, "K_I": 2}[which], int(i), _p(out))
        return out

    def positions(self):
        out = np.zeros((self.N, self.D))
        lib().orc_sim_get_positions(self.h, _p(out))
        return out

    def set_positions(self, x, full_search=False):
        """set_positions!(pg, X, full_search): frozen-topology refresh of distances and weights, or a full search."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.N, self.D)
        lib().orc_sim_set_positions(self.h, _p(x), int(bool(full_search)))

    def volumes(self, sweeps=-1):
        """The particle volumes (ShepardVolumes); sweeps >= 0 recomputes them with that many sweeps first."""
        out = np.zeros(self.N)
        lib().orc_sim_volumes(self.h, int(sweeps), _p(out))
        return out

    def mass(self, interior_only=False):
        """total_mass / interior_mass: sum_i V_i U_i (left fold)."""
        out = np.zeros(self.M)
        lib().orc_sim_mass(self.h, int(bool(interior_only)), _p(out))
        return out

    def mood_iterations(self):
        return lib().orc_sim_mood_iterations(self.h)

    def __del__(self):
        try:
            if self.h:
                lib().orc_sim_free(self.h)
                self.h = None
        except Exception:
            pass


def weight(alpha, R, d2):
    return lib().orc_weight(alpha, R, d2)


def limiter_phi(kind, r):
    return lib().orc_limiter_phi(kind, r)


def spd_inverse(A):
    A = np.ascontiguousarray(A, np.float64)
    X = np.zeros_like(A)
    ok = lib().orc_spd_inverse(A.shape[0], _p(A), _p(X))
    return X, ok


def euler(D, gamma, U):
    U = np.ascontiguousarray(U, np.float64)
    F = np.zeros((D, D + 2))
    lam = np.zeros(D)
    adm = C.c_int32()
    lib().orc_euler(D, gamma, _p(U), _p(F), _p(lam), C.byref(adm))
    return F, lam, bool(adm.value)


def rng_uniform(seed, n):
    out = np.zeros(n)
    lib().orc_rng_uniform(seed, n, _p(out))
    return out


def snapshot_times(tmax, snapshots):
    """range(0, tmax, length = snapshots + 1) (test/time_integration.jl:92): Julia's TwicePrecision range, whose
    elements are the correctly rounded k tmax / snapshots."""
    from fractions import Fraction
    n = int(snapshots)
    return [0.0 if k == 0 else (float(tmax) if k == n else float(Fraction(k) * Fraction(float(tmax)) / n))
            for k in range(n + 1)]


def solve_equation(o, tmax, dt, is_cfl=True, snapshots=10, remove_ghosts=False, isb=None, pos=None):
    """solve_equation (test/time_integration.jl:72-132) on the oracle: the driver loop with dt = dt * get_time_step
    (is_cfl) capped to tmax - t, a stop below 1e-12, and snapshots of the state at the first step reaching each of
    the snapshots + 1 evenly spaced times (the last one after the loop). Returns (xs, us, ts, k_step, masses) with
    masses[c] = (total_mass, interior_mass) of snapshot c."""
    t_snap = snapshot_times(tmax, snapshots)
    xs, us, ts, ms = [None] * (snapshots + 1), [None] * (snapshots + 1), [0.0] * (snapshots + 1), [None] * (snapshots + 1)
    keep = np.ones(o.N, bool) if not remove_ghosts else ~np.asarray(isb, bool)

    def save(c, t):
        ts[c] = t
        us[c] = o.state()[keep]
        xs[c] = (o.positions() if pos is None else np.asarray(pos))[keep]   # a moving grid: the current positions
        ms[c] = (o.mass(False), o.mass(True))

    t, k_step, counter = 0.0, 0, 0
    save(counter, t)
    counter += 1
    while t < tmax:
        dt_val = dt * o.time_step() if is_cfl else dt
        dt_val = min(dt_val, tmax - t)
        if dt_val <= 1e-12:
            break
        o.step(t, dt_val)
        t += dt_val
        k_step += 1
        while counter < snapshots and t >= t_snap[counter]:   # 1-based: snap_counter <= snapshots
            save(counter, t)
            counter += 1
    if counter <= snapshots:
        save(snapshots, t)
    return xs, us, ts, k_step, ms
