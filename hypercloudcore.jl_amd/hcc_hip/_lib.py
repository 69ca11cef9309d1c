"""ctypes binding of libhcc.so (the C ABI declared in include/hcc.h).

The library is the only compute path: if it is missing, or no GPU is present when a device call is made, the
call raises. There is no CPU fallback.
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libhcc.so")

HCC_PDE_LINADV, HCC_PDE_EULER = 1, 2
HCC_LIM_NONE, HCC_LIM_BJ, HCC_LIM_VK = 0, 1, 2
HCC_MODE_SOFT, HCC_MODE_HARD = 0, 1
HCC_STRAT_NONE, HCC_STRAT_EPD0, HCC_STRAT_EPD1, HCC_STRAT_EPD2, HCC_STRAT_STRICT_EPD0 = 0, 1, 2, 3, 4
HCC_CRIT_NONE, HCC_CRIT_U1, HCC_CRIT_PAD, HCC_CRIT_ONLY, HCC_CRIT_U2 = 0, 1, 2, 3, 4
HCC_FLUX_RUSANOV, HCC_FLUX_ALIGNED_RUSANOV = 0, 1
HCC_SRC_NONE, HCC_SRC_DECAY_EXPLICIT, HCC_SRC_DECAY_IMPLICIT = 0, 1, 2
HCC_MOVER_NONE, HCC_MOVER_LAGRANGIAN, HCC_MOVER_SCALED = 0, 1, 2   # HCC_MOVER_SCALED_LAGRANGIAN in hcc.h
HCC_BC_NONE, HCC_BC_OUTFLOW, HCC_BC_DIRICHLET = 0, 1, 2
HCC_E_CAPACITY = -4

# every symbol include/hcc.h declares
EXPORTS = [
    "hcc_ctx_create", "hcc_ctx_destroy", "hcc_last_error", "hcc_ctx_synchronize", "hcc_cloud_generate",
    "hcc_rng_uniform", "hcc_grid_create", "hcc_grid_destroy", "hcc_grid_num_slots", "hcc_grid_get_neighbors",
    "hcc_grid_get_bins", "hcc_grid_get_positions", "hcc_stepper_create", "hcc_stepper_destroy", "hcc_state_upload",
    "hcc_state_download", "hcc_time_step", "hcc_step", "hcc_get_orders", "hcc_get_mood_flags", "hcc_get_inadmissible",
    "hcc_get_mood_iterations", "hcc_get_moments", "hcc_rhs", "hcc_last_step_timing", "hcc_step_totals",
    "hcc_step_cfl", "hcc_last_dt", "hcc_grid_volumes", "hcc_total_mass", "hcc_solve", "hcc_set_source",
    "hcc_grid_concat", "hcc_time_steps", "hcc_mlp_create", "hcc_mlp_destroy", "hcc_mlp_set_params",
    "hcc_mlp_get_params", "hcc_mlp_forward", "hcc_mlp_gradients", "hcc_mlp_gradients_noalloc", "hcc_mlp_train",
    "hcc_set_indicator", "hcc_indicator_collect", "hcc_get_indicator_samples", "hcc_get_indicator_count",
    "hcc_stepper_set_graph", "hcc_stepper_graph_stats", "hcc_stepper_set_tuning", "hcc_stepper_get_tuning",
    "hcc_stepper_autotune", "hcc_stepper_set_pass_timing", "hcc_step_timed_passes",
]
MLP_PARAMS = 5313


class DomainDesc(C.Structure):
    _fields_ = [("D", C.c_int32), ("mins", C.c_double * 3), ("maxs", C.c_double * 3), ("periodic", C.c_int32 * 3),
                ("dx", C.c_double * 3), ("R", C.c_double), ("alpha", C.c_double), ("jitter", C.c_double * 3),
                ("seed", C.c_uint64)]


class StepperDesc(C.Structure):
    _fields_ = [("M", C.c_int32), ("pde", C.c_int32), ("gamma", C.c_double), ("adv_speed", (C.c_double * 8) * 3),
                ("muscl_order", C.c_int32), ("limiter", C.c_int32), ("limiter_mode", C.c_int32),
                ("mood_strategy", C.c_int32), ("mood_criterion", C.c_int32), ("mood_delta", C.c_double),
                ("stages", C.c_int32), ("a_t", (C.c_double * 5) * 5), ("b_t", C.c_double * 5), ("c_t", C.c_double * 5),
                ("substep_ratio", C.c_double), ("n_bc", C.c_int32), ("bc_tag", C.c_int32 * 16),
                ("bc_kind", C.c_int32 * 16), ("mood_plateau", C.c_double), ("muscl_streamed", C.c_int32),
                ("flux", C.c_int32), ("a_im", (C.c_double * 5) * 5), ("b_im", C.c_double * 5),
                ("c_im", C.c_double * 5), ("mover", C.c_int32), ("mover_alpha", C.c_double),
                ("mood_ratio", C.c_double)]


_lib = None


def lib():
    """Load libhcc.so (raises if it has not been built: there is no other implementation)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libhcc.so not found at {LIB_PATH}: build it with __graft_entry__.build() "
                           "(the HIP library is the only implementation; there is no fallback)")
    L = C.CDLL(LIB_PATH)
    vp, i64, i32, d = C.c_void_p, C.c_int64, C.c_int32, C.c_double
    P = C.POINTER
    sig = {
        "hcc_ctx_create": (C.c_int, [C.c_int, P(vp)]),
        "hcc_ctx_destroy": (C.c_int, [vp]),
        "hcc_last_error": (C.c_char_p, [vp]),
        "hcc_ctx_synchronize": (C.c_int, [vp]),
        "hcc_cloud_generate": (C.c_int, [P(DomainDesc), i64, vp, vp, vp, P(i64), vp, P(i64)]),
        "hcc_rng_uniform": (C.c_int, [C.c_uint64, i64, vp]),
        "hcc_grid_create": (C.c_int, [vp, P(DomainDesc), i64, vp, vp, vp, P(vp)]),
        "hcc_grid_destroy": (C.c_int, [vp]),
        "hcc_grid_num_slots": (C.c_int, [vp, P(i64), P(i64)]),
        "hcc_grid_get_neighbors": (C.c_int, [vp, vp, vp, vp, vp]),
        "hcc_grid_get_bins": (C.c_int, [vp, vp, vp, vp, vp]),
        "hcc_grid_get_positions": (C.c_int, [vp, vp]),
        "hcc_stepper_create": (C.c_int, [vp, vp, P(StepperDesc), P(vp)]),
        "hcc_stepper_destroy": (C.c_int, [vp]),
        "hcc_state_upload": (C.c_int, [vp, vp]),
        "hcc_state_download": (C.c_int, [vp, vp]),
        "hcc_time_step": (C.c_int, [vp, P(d)]),
        "hcc_step": (C.c_int, [vp, d, d]),
        "hcc_get_orders": (C.c_int, [vp, vp]),
        "hcc_get_mood_flags": (C.c_int, [vp, vp]),
        "hcc_get_inadmissible": (C.c_int, [vp, P(i64)]),
        "hcc_get_mood_iterations": (C.c_int, [vp, P(i64)]),
        "hcc_get_moments": (C.c_int, [vp, vp, vp, vp, vp, vp]),
        "hcc_rhs": (C.c_int, [vp, vp, vp, vp]),
        "hcc_last_step_timing": (C.c_int, [vp, P(d), P(d), P(d)]),
        "hcc_step_totals": (C.c_int, [vp, C.c_int, P(i64), P(i64), P(d), P(d), P(d)]),
        "hcc_stepper_set_pass_timing": (C.c_int, [vp, C.c_int32]),
        "hcc_step_timed_passes": (C.c_int, [vp, P(i64)]),
        "hcc_step_cfl": (C.c_int, [vp, C.c_double, C.c_double]),
        "hcc_last_dt": (C.c_int, [vp, P(d)]),
        "hcc_grid_volumes": (C.c_int, [vp, C.c_int32, vp]),
        "hcc_total_mass": (C.c_int, [vp, C.c_int32, vp]),
        "hcc_solve": (C.c_int, [vp, d, d, C.c_int32, C.c_int32, C.c_int32, vp, vp, vp, vp, P(i64), P(d)]),
        "hcc_set_source": (C.c_int, [vp, C.c_int32, vp]),
        "hcc_grid_concat": (C.c_int, [vp, C.c_int32, vp, P(vp), vp]),
        "hcc_time_steps": (C.c_int, [vp, vp]),
        "hcc_mlp_create": (C.c_int, [vp, P(vp)]),
        "hcc_mlp_destroy": (C.c_int, [vp]),
        "hcc_mlp_set_params": (C.c_int, [vp, vp]),
        "hcc_mlp_get_params": (C.c_int, [vp, vp]),
        "hcc_mlp_forward": (C.c_int, [vp, vp, i64, vp]),
        "hcc_mlp_gradients": (C.c_int, [vp, vp, vp, i64, C.c_float, vp, P(d)]),
        "hcc_mlp_gradients_noalloc": (C.c_int, [vp, vp, vp, i64, C.c_float, vp]),
        "hcc_mlp_train": (C.c_int, [vp, vp, vp, i64, C.c_int32, C.c_int32, C.c_float, C.c_float, vp]),
        "hcc_stepper_set_graph": (C.c_int, [vp, C.c_int32]),
        "hcc_stepper_set_tuning": (C.c_int, [vp, vp]),
        "hcc_stepper_get_tuning": (C.c_int, [vp, vp]),
        "hcc_stepper_autotune": (C.c_int, [vp, C.c_int32, d, vp, P(d), P(d), P(i32)]),
        "hcc_stepper_graph_stats": (C.c_int, [vp, P(i64), P(i64), P(i64), vp]),
        "hcc_set_indicator": (C.c_int, [vp, vp, C.c_float]),
        "hcc_indicator_collect": (C.c_int, [vp, C.c_int32]),
        "hcc_get_indicator_samples": (C.c_int, [vp, C.c_int32, vp, vp, vp]),
        "hcc_get_indicator_count": (C.c_int, [vp, P(i64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    _lib = L
    return L


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class HccError(RuntimeError):
    pass


def check(rc, ctx=None, what=""):
    if rc != 0:
        msg = lib().hcc_last_error(ctx)
        raise HccError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def domain_desc(D, mins, maxs, periodic, dx, R, alpha=1.0, jitter=None, seed=0):
    dd = DomainDesc()
    dd.D = D
    for a in range(D):
        dd.mins[a], dd.maxs[a] = float(mins[a]), float(maxs[a])
        dd.periodic[a] = int(bool(periodic[a]))
        dd.dx[a] = float(dx[a])
        dd.jitter[a] = float(jitter[a]) if jitter is not None else 0.0
    dd.R, dd.alpha, dd.seed = float(R), float(alpha), int(seed)
    return dd


def cloud_generate(dd):
    """Host point cloud (lattice + jitter + ghost band + tags, Morton-ordered)."""
    L = lib()
    n = C.c_int64(0)
    ties = C.c_int64(0)
    canvas = np.zeros(2 * dd.D)
    rc = L.hcc_cloud_generate(C.byref(dd), 0, None, None, None, C.byref(n), ptr(canvas), C.byref(ties))
    if rc not in (0, HCC_E_CAPACITY):
        check(rc, None, "hcc_cloud_generate")
    N = n.value
    pos = np.zeros((N, dd.D))
    isb = np.zeros(N, np.uint8)
    tags = np.zeros(N, np.int32)
    check(L.hcc_cloud_generate(C.byref(dd), N, ptr(pos), ptr(isb), ptr(tags), C.byref(n), ptr(canvas), C.byref(ties)),
          None, "hcc_cloud_generate")
    return {"pos": pos, "is_boundary": isb, "tags": tags, "canvas": canvas, "morton_ties": ties.value, "N": N}


def rng_uniform(seed, n):
    out = np.zeros(int(n))
    check(lib().hcc_rng_uniform(int(seed), int(n), ptr(out)), None, "hcc_rng_uniform")
    return out


_live_contexts = weakref.WeakSet()


def _close_live_contexts():
    """At interpreter exit, before modules are torn down: release every context still open (its steppers, grids and
    networks first), while the library is still reachable. Objects left to module teardown could otherwise not call
    into the library and would leave streams, events and pinned memory to the HIP runtime's own exit."""
    for c in list(_live_contexts):
        try:
            c.close()
        except Exception:
            pass


atexit.register(_close_live_contexts)


class Context:
    """One GPU context (device + stream). Grids and steppers hold its handle: closing the context closes the ones
    still alive first (steppers, then grids), so that none is destroyed later through a freed context. Contexts still
    open at interpreter exit are closed by an exit handler."""

    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        check(L.hcc_ctx_create(int(device), C.byref(h)), None, "hcc_ctx_create")
        self.h = h
        self._steppers = weakref.WeakSet()
        self._grids = weakref.WeakSet()
        _live_contexts.add(self)

    def close(self):
        if getattr(self, "h", None):
            for obj in list(self._steppers) + list(self._grids):
                obj.close()
            lib().hcc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Grid:
    """Device particle grid: nodes uploaded and the neighbour search run on the GPU."""

    def __init__(self, ctx, dd, pos, is_boundary, tags):
        L = lib()
        self.ctx, self.dd = ctx, dd
        self.D = dd.D
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        self.N = pos.shape[0]
        self._keep = (pos, np.ascontiguousarray(is_boundary, np.uint8), np.ascontiguousarray(tags, np.int32))
        h = C.c_void_p()
        check(L.hcc_grid_create(ctx.h, C.byref(dd), self.N, ptr(self._keep[0]), ptr(self._keep[1]), ptr(self._keep[2]),
                                C.byref(h)), ctx.h, "hcc_grid_create")
        self.h = h
        ctx._grids.add(self)

    def _slots(self):
        S, mx = C.c_int64(), C.c_int64()
        check(lib().hcc_grid_num_slots(self.h, C.byref(S), C.byref(mx)), self.ctx.h, "hcc_grid_num_slots")
        return S.value, mx.value

    @property
    def S(self):
        """Neighbour slots of the current lists (a moving grid's search after each step changes them)."""
        return self._slots()[0]

    @property
    def max_nb(self):
        return self._slots()[1]

    def neighbors(self, weights=True, distances=True):
        S = self.S
        ranges = np.zeros(self.N + 1, np.int64)
        idx = np.zeros(max(S, 1), np.int32)
        w = np.zeros(max(S, 1)) if weights else None
        dist = np.zeros((max(S, 1), self.D)) if distances else None
        check(lib().hcc_grid_get_neighbors(self.h, ptr(ranges), ptr(idx), ptr(w), ptr(dist)), self.ctx.h,
              "hcc_grid_get_neighbors")
        return ranges, idx[:S], (w[:S] if w is not None else None), (dist[:S] if dist is not None else None)

    def bins(self):
        dims = np.zeros(4, np.int64)
        check(lib().hcc_grid_get_bins(self.h, None, None, None, ptr(dims)), self.ctx.h, "hcc_grid_get_bins")
        nb = int(dims[3])
        bin_of = np.zeros(self.N, np.int32)
        cs = np.zeros(nb + 2, np.int64)
        cp = np.zeros(self.N, np.int32)
        check(lib().hcc_grid_get_bins(self.h, ptr(bin_of), ptr(cs), ptr(cp), ptr(dims)), self.ctx.h, "hcc_grid_get_bins")
        return bin_of, cs, cp, dims

    def positions(self):
        out = np.zeros((self.N, self.D))
        check(lib().hcc_grid_get_positions(self.h, ptr(out)), self.ctx.h, "hcc_grid_get_positions")
        return out

    def volumes(self, sweeps=-1):
        """Particle volumes (ShepardVolumes); sweeps >= 0 recomputes them with that many sweeps first."""
        out = np.zeros(self.N)
        check(lib().hcc_grid_volumes(self.h, int(sweeps), ptr(out)), self.ctx.h, "hcc_grid_volumes")
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().hcc_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bf16_rows(X):
    """Feature rows as bfloat16 bit patterns (round to nearest even), 16 per row."""
    f = np.ascontiguousarray(X, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return np.ascontiguousarray((((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16))


class MLP:
    """The troubled-node indicator network on the device (hcc_mlp_*): 16 -> 64 -> 64 -> 1, bf16 operands on the
    matrix cores, fp32 parameters and accumulation."""

    def __init__(self, ctx, params=None):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().hcc_mlp_create(ctx.h, C.byref(h)), ctx.h, "hcc_mlp_create")
        self.h = h
        ctx._grids.add(self)
        if params is not None:
            self.set_params(params)

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == MLP_PARAMS
        check(lib().hcc_mlp_set_params(self.h, ptr(p)), self.ctx.h, "hcc_mlp_set_params")

    def params(self):
        p = np.zeros(MLP_PARAMS, np.float32)
        check(lib().hcc_mlp_get_params(self.h, ptr(p)), self.ctx.h, "hcc_mlp_get_params")
        return p

    def forward(self, X):
        Xb = bf16_rows(X) if X.dtype != np.uint16 else np.ascontiguousarray(X)
        z = np.zeros(Xb.shape[0], np.float32)
        check(lib().hcc_mlp_forward(self.h, ptr(Xb), Xb.shape[0], ptr(z)), self.ctx.h, "hcc_mlp_forward")
        return z

    def gradients(self, X, y, pos_weight=1.0):
        Xb = bf16_rows(X) if X.dtype != np.uint16 else np.ascontiguousarray(X)
        yb = np.ascontiguousarray(y, np.uint8)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_double()
        check(lib().hcc_mlp_gradients(self.h, ptr(Xb), ptr(yb), Xb.shape[0], float(pos_weight), ptr(g), C.byref(loss)),
              self.ctx.h, "hcc_mlp_gradients")
        return loss.value, g

    def train(self, X, y, epochs, batch, lr, pos_weight=1.0):
        Xb = bf16_rows(X) if X.dtype != np.uint16 else np.ascontiguousarray(X)
        yb = np.ascontiguousarray(y, np.uint8)
        losses = np.zeros(epochs)
        check(lib().hcc_mlp_train(self.h, ptr(Xb), ptr(yb), Xb.shape[0], int(epochs), int(batch), float(lr),
                                  float(pos_weight), ptr(losses)), self.ctx.h, "hcc_mlp_train")
        return losses

    def close(self):
        if getattr(self, "h", None):
            lib().hcc_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble(Grid):
    """One device grid holding copies of several instance grids (hcc_grid_concat): instance k's nodes at
    [offsets[k], offsets[k] + sizes[k]), each instance padded to whole 256-node blocks with inert boundary nodes. A
    Stepper on it advances every instance in one step; pack / split convert between per-instance arrays and the
    ensemble's."""

    def __init__(self, ctx, grids):
        self.ctx, self.dd = ctx, grids[0].dd
        self.D = grids[0].D
        n = len(grids)
        handles = (C.c_void_p * n)(*[g.h for g in grids])
        offs = np.zeros(n + 1, np.int64)
        h = C.c_void_p()
        check(lib().hcc_grid_concat(ctx.h, n, handles, C.byref(h), ptr(offs)), ctx.h, "hcc_grid_concat")
        self.h = h
        ctx._grids.add(self)
        self.offsets = offs
        self.sizes = np.array([g.N for g in grids], np.int64)
        self.N = int(offs[-1])

    def pack(self, parts, fill=0.0):
        """Per-instance arrays (N_k x ...) into one ensemble array (padding rows = fill)."""
        first = np.asarray(parts[0])
        out = np.full((self.N,) + first.shape[1:], fill, first.dtype)
        for k, a in enumerate(parts):
            out[self.offsets[k]:self.offsets[k] + self.sizes[k]] = a
        return out

    def split(self, a):
        return [a[self.offsets[k]:self.offsets[k] + self.sizes[k]] for k in range(len(self.sizes))]


def stepper_desc(scheme):
    """hcc_stepper_desc from a plain scheme dict (see hcc_hip.workloads)."""
    sd = StepperDesc()
    sd.M = scheme["M"]
    sd.pde = scheme["pde"]
    sd.gamma = scheme.get("gamma", 1.4)
    for d, row in enumerate(scheme.get("adv_speed", [])):
        for m, v in enumerate(row):
            sd.adv_speed[d][m] = v
    sd.muscl_order = scheme.get("muscl_order", 2)
    sd.limiter, sd.limiter_mode = scheme["limiter"], scheme.get("limiter_mode", HCC_MODE_HARD)
    sd.mood_strategy, sd.mood_criterion = scheme["strategy"], scheme["criterion"]
    sd.mood_delta = scheme.get("delta", 0.0)
    sd.mood_plateau = scheme.get("plateau", 0.0)
    sd.muscl_streamed = int(bool(scheme.get("streamed", False)))
    sd.flux = scheme.get("flux", HCC_FLUX_RUSANOV)
    sd.mover = scheme.get("mover", HCC_MOVER_NONE)
    sd.mover_alpha = scheme.get("mover_alpha", 1.0)
    sd.mood_ratio = scheme.get("ratio", 0.5)
    a_t, b_t, c_t = scheme["a_t"], scheme["b_t"], scheme["c_t"]
    s = len(b_t)
    sd.stages = s
    for i in range(s):
        sd.b_t[i], sd.c_t[i] = b_t[i], c_t[i]
        for j in range(s):
            sd.a_t[i][j] = a_t[i][j]
    if scheme.get("a_im") is not None:   # implicit part of an IMEX tableau
        for i in range(s):
            sd.b_im[i], sd.c_im[i] = scheme["b_im"][i], scheme["c_im"][i]
            for j in range(s):
                sd.a_im[i][j] = scheme["a_im"][i][j]
    sd.substep_ratio = scheme.get("substep_ratio", 0.0)
    bcs = scheme.get("bcs", [])
    sd.n_bc = len(bcs)
    for k, (tag, kind) in enumerate(bcs):
        sd.bc_tag[k], sd.bc_kind[k] = tag, kind
    return sd


def basis_count(D, q):
    """Terms of degree 1..q in D variables, binomial(D + q, q) - 1 (Interpolators.jl:170-177)."""
    if q <= 0:
        return 0
    n = 1
    for k in range(1, q + 1):
        n = (n * (D + k)) // k
    return n - 1


class Stepper:
    """Device time stepper (GeneralRKTimeStepper with MUSCL order 2 or 3 + MOOD) bound to one grid."""

    def __init__(self, ctx, grid, scheme):
        self.ctx, self.grid, self.scheme = ctx, grid, scheme
        self.N, self.M, self.D = grid.N, scheme["M"], grid.D
        self.order = int(scheme.get("muscl_order", 2))
        # reconstruction terms (degree order - 1) and divergence terms with intercept (degree 1, or 3 for MUSCL3)
        self.BG = basis_count(self.D, self.order - 1)
        self.BD = basis_count(self.D, 1 if self.order <= 2 else 3) + 1
        self.sd = stepper_desc(scheme)
        h = C.c_void_p()
        check(lib().hcc_stepper_create(ctx.h, grid.h, C.byref(self.sd), C.byref(h)), ctx.h, "hcc_stepper_create")
        self.h = h
        ctx._steppers.add(self)

    def upload(self, U):
        U = np.ascontiguousarray(U, np.float64).reshape(self.N, self.M)
        check(lib().hcc_state_upload(self.h, ptr(U)), self.ctx.h, "hcc_state_upload")

    def download(self):
        U = np.zeros((self.N, self.M))
        check(lib().hcc_state_download(self.h, ptr(U)), self.ctx.h, "hcc_state_download")
        return U

    def time_step(self):
        v = C.c_double()
        check(lib().hcc_time_step(self.h, C.byref(v)), self.ctx.h, "hcc_time_step")
        return v.value

    def time_steps(self):
        """The CFL step of every instance of an ensemble grid (one value for a plain grid)."""
        n = len(getattr(self.grid, "sizes", [0]))
        out = np.zeros(n)
        check(lib().hcc_time_steps(self.h, ptr(out)), self.ctx.h, "hcc_time_steps")
        return out

    def set_graph(self, on=True):
        """hcc_step_cfl as one hipGraph launch per step (hcc_stepper_set_graph)."""
        check(lib().hcc_stepper_set_graph(self.h, 1 if on else 0), self.ctx.h, "hcc_stepper_set_graph")

    def graph_stats(self):
        """(launches, captures, re-runs, iterations captured per checkpoint)."""
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        bounds = np.zeros(5, np.int32)
        check(lib().hcc_stepper_graph_stats(self.h, C.byref(a), C.byref(b), C.byref(c), ptr(bounds)), self.ctx.h,
              "hcc_stepper_graph_stats")
        return a.value, b.value, c.value, bounds[:len(self.scheme["b_t"])].copy()

    def set_tuning(self, staged=1, split=0, sweep=65536, fuse_cfl=1, dense=64):
        """Launch choices (hcc_stepper_set_tuning); every choice gives the same results."""
        k = np.array([staged, split, sweep, fuse_cfl, dense], np.int32)
        check(lib().hcc_stepper_set_tuning(self.h, ptr(k)), self.ctx.h, "hcc_stepper_set_tuning")

    def tuning(self):
        k = np.zeros(5, np.int32)
        check(lib().hcc_stepper_get_tuning(self.h, ptr(k)), self.ctx.h, "hcc_stepper_get_tuning")
        return {"staged": int(k[0]), "split": int(k[1]), "sweep": int(k[2]), "fuse_cfl": int(k[3]), "dense": int(k[4])}

    def autotune(self, steps=3, cfl=0.4):
        """hcc_stepper_autotune: (chosen knobs, best ms/step, default ms/step, timed)."""
        k = np.zeros(5, np.int32)
        best, dflt, timed = C.c_double(), C.c_double(), C.c_int32()
        check(lib().hcc_stepper_autotune(self.h, int(steps), float(cfl), ptr(k), C.byref(best), C.byref(dflt),
                                         C.byref(timed)), self.ctx.h, "hcc_stepper_autotune")
        return ({"staged": int(k[0]), "split": int(k[1]), "sweep": int(k[2]), "fuse_cfl": int(k[3]), "dense": int(k[4])},
                best.value,
                dflt.value, bool(timed.value))

    def set_indicator(self, mlp, threshold=0.0):
        """The learned indicator before each flux pass (mlp None: off)."""
        self._ind = mlp
        check(lib().hcc_set_indicator(self.h, mlp.h if mlp is not None else None, float(threshold)), self.ctx.h,
              "hcc_set_indicator")

    def collect(self, on=True):
        check(lib().hcc_indicator_collect(self.h, int(bool(on))), self.ctx.h, "hcc_indicator_collect")

    def samples(self, stage):
        """The last step's stage: (features N x 16 as bf16 bits, orders before the flux pass, after the checkpoint)."""
        f = np.zeros((self.N, 16), np.uint16)
        pre, post = np.zeros(self.N, np.int8), np.zeros(self.N, np.int8)
        check(lib().hcc_get_indicator_samples(self.h, int(stage), ptr(f), ptr(pre), ptr(post)), self.ctx.h,
              "hcc_get_indicator_samples")
        return f, pre, post

    def indicator_count(self):
        c = C.c_int64()
        check(lib().hcc_get_indicator_count(self.h, C.byref(c)), self.ctx.h, "hcc_get_indicator_count")
        return c.value

    def step(self, t, dt):
        check(lib().hcc_step(self.h, float(t), float(dt)), self.ctx.h, "hcc_step")

    def orders(self):
        o = np.zeros(self.N, np.int32)
        check(lib().hcc_get_orders(self.h, ptr(o)), self.ctx.h, "hcc_get_orders")
        return o

    def mood_flags(self):
        f = np.zeros(self.N, np.uint8)
        check(lib().hcc_get_mood_flags(self.h, ptr(f)), self.ctx.h, "hcc_get_mood_flags")
        return f

    def inadmissible(self):
        v = C.c_int64()
        check(lib().hcc_get_inadmissible(self.h, C.byref(v)), self.ctx.h, "hcc_get_inadmissible")
        return v.value

    def mood_iterations(self):
        v = C.c_int64()
        check(lib().hcc_get_mood_iterations(self.h, C.byref(v)), self.ctx.h, "hcc_get_mood_iterations")
        return v.value

    def moments(self):
        D, N, S = self.D, self.N, self.grid.S
        ig = np.zeros((N, self.BG, self.BG))
        og = np.zeros(N, np.uint8)
        idv = np.zeros((N, self.BD, self.BD))
        od = np.zeros(N, np.uint8)
        dw = np.zeros((max(S, 1), D))
        check(lib().hcc_get_moments(self.h, ptr(ig), ptr(og), ptr(idv), ptr(od), ptr(dw)), self.ctx.h, "hcc_get_moments")
        return ig, og, idv, od, dw[:S]

    def rhs(self, Y):
        Y = np.ascontiguousarray(Y, np.float64).reshape(self.N, self.M)
        K = np.zeros((self.N, self.M))
        G = np.zeros((self.N, self.BG, self.M))
        check(lib().hcc_rhs(self.h, ptr(Y), ptr(K), ptr(G)), self.ctx.h, "hcc_rhs")
        return K, G

    def timing(self):
        """(step ms, summed full gather-pass ms, summed full flux-pass ms) of the last step"""
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        check(lib().hcc_last_step_timing(self.h, C.byref(a), C.byref(b), C.byref(c)), self.ctx.h, "hcc_last_step_timing")
        return a.value, b.value, c.value

    def step_cfl(self, cfl, dt_cap=float("inf")):
        """One step with dt = min(cfl * CFL step, dt_cap) computed on the device; returns once queued"""
        check(lib().hcc_step_cfl(self.h, float(cfl), float(dt_cap)), self.ctx.h, "hcc_step_cfl")

    def last_dt(self):
        v = C.c_double()
        check(lib().hcc_last_dt(self.h, C.byref(v)), self.ctx.h, "hcc_last_dt")
        return v.value

    def set_source(self, mode, rates):
        """The decay source S_i(U) = -rates_i U: mode HCC_SRC_DECAY_EXPLICIT / _IMPLICIT, or HCC_SRC_NONE."""
        r = np.ascontiguousarray(rates, np.float64) if rates is not None else None
        check(lib().hcc_set_source(self.h, int(mode), ptr(r)), self.ctx.h, "hcc_set_source")

    def mass(self, interior_only=False):
        """total_mass (or interior_mass) of the device state: sum_i V_i U_i, M values."""
        out = np.zeros(self.M)
        check(lib().hcc_total_mass(self.h, int(bool(interior_only)), ptr(out)), self.ctx.h, "hcc_total_mass")
        return out

    def solve(self, tmax, dt, is_cfl=True, snapshots=10, remove_ghosts=False, masses=True):
        """solve_equation on the device: (xs, us, ts, steps, elapsed, masses) with xs / us of shape
        (snapshots + 1, Ns, D / M), ts (snapshots + 1,), masses (snapshots + 1, 2, M) = (total, interior) or None."""
        ns = self.N if not remove_ghosts else int(np.count_nonzero(~self.grid._keep[1].astype(bool)))
        ts = np.zeros(snapshots + 1)
        xs = np.zeros((snapshots + 1, ns, self.D))
        us = np.zeros((snapshots + 1, ns, self.M))
        ms = np.zeros((snapshots + 1, 2, self.M)) if masses else None
        steps, el = C.c_int64(), C.c_double()
        check(lib().hcc_solve(self.h, float(tmax), float(dt), int(bool(is_cfl)), int(snapshots), int(bool(remove_ghosts)),
                              ptr(ts), ptr(xs), ptr(us), ptr(ms), C.byref(steps), C.byref(el)), self.ctx.h, "hcc_solve")
        return xs, us, ts, steps.value, el.value, ms

    def totals(self, reset=True):
        """(steps, MOOD iterations, step ms, full gather-pass ms, full flux-pass ms) summed since the last reset"""
        n, it = C.c_int64(), C.c_int64()
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        check(lib().hcc_step_totals(self.h, int(reset), C.byref(n), C.byref(it), C.byref(a), C.byref(b), C.byref(c)),
              self.ctx.h, "hcc_step_totals")
        return n.value, it.value, a.value, b.value, c.value

    def set_pass_timing(self, every=1):
        """Record the gather / flux pass events every `every`-th step (0: never; hcc_stepper_set_pass_timing)."""
        check(lib().hcc_stepper_set_pass_timing(self.h, int(every)), self.ctx.h, "hcc_stepper_set_pass_timing")

    def timed_passes(self):
        """The steps of the current totals whose passes were timed (hcc_step_timed_passes); read before a reset."""
        n = C.c_int64()
        check(lib().hcc_step_timed_passes(self.h, C.byref(n)), self.ctx.h, "hcc_step_timed_passes")
        return n.value

    def close(self):
        if getattr(self, "h", None):
            lib().hcc_stepper_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
