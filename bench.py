"""Throughput of the explicit time step: 2-D compressible Euler, MUSCL(2) + Venkatakrishnan + Rusanov + MOOD
(EPD1, u1), SSP-RK3, CFL 0.4, outflow walls, on the jittered h = 1e-3 cloud of the unit square (about 1.01 M nodes
including the ghost band), four seeded constant quadrants.

One step = get_time_step (CFL reduction on the device) + one RK3 step with its MOOD fixed points, exactly as the
reference's driver loop (test/time_integration.jl:107-111). The timed region holds K steps with the state resident
in HBM; the metric is node-stage updates per second = N * 3 * K / seconds.

Prints one JSON line (rank 0). --gpus N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "hypercloudcore.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "node-stage updates/s, 2D Euler MUSCL+MOOD @1M nodes; % of HBM roofline"


def flux_pass_bytes(N, N_int, S_int, D, M, k_terms=1.0):
    """Algorithmic HBM bytes of one full flux pass of this implementation (neighbour data read once): per interior
    node its SELL indices (4 B per slot), count, order, state, gradients, position, the derivative rows of the
    divergence inverse, the output K and the MOODu1 bounds stored by the gather pass (2M doubles); boundary nodes
    write a zero K. The fused stage predictor adds per interior node the U^n state, k_terms earlier K rows (1 on
    average over the three RK3 passes) and the next stage state."""
    per_node = 4 + 1 + 1 + 8 * M + 8 * D * M + 8 * D + 8 * D * (D + 1) + 8 * M + 1 + 16 * M
    predictor = 8 * M + k_terms * 8 * M + 8 * M
    return N_int * (per_node + predictor) + 4 * S_int + (N - N_int) * (8 * M + 1)


def gather_pass_bytes(N_int, S_int, D, M):
    """Algorithmic HBM bytes of one full gather pass: per interior node its indices, count, order, state, position,
    the reconstruction inverse, the gradient record and the MOODu1 bounds it stores (2M doubles)."""
    per_node = 4 + 1 + 1 + 8 * M + 8 * D + 8 * D * D + 8 * D * M + 1 + 16 * M
    return N_int * per_node + 4 * S_int



def cpu_baseline(wl, steps, threads):
    """The oracle (C++ restatement of the reference CPU path, OpenMP over host cores) on a bounded sample: a few
    steps of the same cloud and state."""
    from oracle import oracle as O
    c = wl["cloud"]
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], wl["scheme"], threads=threads)
    o.set_state(wl["U0"])
    t0 = time.perf_counter()
    for k in range(steps):
        dt = wl["cfl"] * o.time_step()
        o.step(0.0, dt)
    t = time.perf_counter() - t0
    return {"value": c["N"] * 3 * steps / t, "unit": "node-stage updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps (CFL + RK3 + MOOD) of the same {c['N']}-node cloud and initial state, "
                      f"{t:.1f} s on {threads} host threads"}


def copy_calibration(nbytes=256 << 20, reps=3):
    """A plain device-to-device copy of a known size through the HIP runtime: the reference dispatch for scaling
    rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0          # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}), flush=True)


def ensemble_run(args, ctx=None):
    """The batched form: `members` four-quadrant Euler boxes (one spacing h = 1e-3, node counts log-uniform in
    [n_min, n_max], seeded) in one ensemble grid advanced by one stepper, each member with its own CFL step, with and
    without the learned indicator, against the plain form (a stepper per member, stepped in turn). All forms are timed
    over the same number of steps from the same initial states; members are compared bit for bit afterwards."""
    import hcc_hip
    from hcc_hip import Context, Ensemble, Grid, Stepper
    from hcc_hip import workloads as W

    t_setup = time.perf_counter()
    wls = W.euler_2d_ensemble(args.members, args.n_min, args.n_max, h=1e-3, seed=args.seed)
    own_ctx = ctx is None
    ctx = ctx or Context(0)
    grids, singles = [], []
    for wl in wls:
        c = wl["cloud"]
        g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
        st = Stepper(ctx, g, wl["scheme"])
        st.upload(wl["U0"])
        grids.append(g)
        singles.append(st)
    ens = Ensemble(ctx, grids)
    st_e = Stepper(ctx, ens, wls[0]["scheme"])
    U0 = ens.pack([wl["U0"] for wl in wls])
    st_e.upload(U0)
    tuning = None
    if not getattr(args, "no_autotune", False):   # the batched stepper's launch choices (its state restored)
        st_e.step_cfl(wls[0]["cfl"])
        knobs, ms_best, ms_def, timed = st_e.autotune(steps=2, cfl=wls[0]["cfl"])
        tuning = dict(knobs, ms_per_step_best=round(ms_best, 3), ms_per_step_default=round(ms_def, 3), timed=timed)
        st_e.upload(U0)
    st_i, net = None, None
    if args.indicator:   # the batched form with the learned indicator (the shipped network)
        from hcc_hip import MLP
        from hcc_hip import indicator as IND
        net = MLP(ctx, IND.default_params())
        st_i = Stepper(ctx, ens, wls[0]["scheme"])
        st_i.upload(U0)
        st_i.set_indicator(net, IND.DEFAULT_THRESHOLD)
        if tuning:
            st_i.set_tuning(tuning["staged"], tuning["split"], tuning["sweep"], tuning["fuse_cfl"], tuning["dense"])
    nodes = int(ens.sizes.sum())
    t_setup = time.perf_counter() - t_setup
    print(f"# ensemble setup {t_setup:.1f} s: {len(wls)} members, {nodes} nodes (ensemble grid {ens.N} with padding)",
          file=sys.stderr, flush=True)
    cfl = wls[0]["cfl"]
    steps, warmup = args.ensemble_steps, args.ensemble_warmup

    def run(steppers, k):
        for _ in range(k):
            for st in steppers:
                st.step_cfl(cfl)
        hcc_hip.lib().hcc_ctx_synchronize(ctx.h)

    def timed(steppers):
        run(steppers, warmup)
        t0 = time.perf_counter()
        run(steppers, steps)
        return time.perf_counter() - t0

    wall_e = timed([st_e])
    print(f"# ensemble batched: {steps} steps in {wall_e:.3f} s", file=sys.stderr, flush=True)
    wall_i = timed([st_i]) if st_i is not None else None
    if wall_i is not None:
        print(f"# ensemble batched with the indicator: {steps} steps in {wall_i:.3f} s", file=sys.stderr, flush=True)
    wall_g = None
    if getattr(args, "graph_form", False):   # the batched form as hipGraph-captured steps, from the same state
        st_g = Stepper(ctx, ens, wls[0]["scheme"])
        st_g.upload(U0)
        if tuning:
            st_g.set_tuning(tuning["staged"], tuning["split"], tuning["sweep"], tuning["fuse_cfl"], tuning["dense"])
        st_g.set_graph(True)
        wall_g = timed([st_g])
        assert np.array_equal(st_g.download(), st_e.download()), "captured batched steps differ"
        g_stats = st_g.graph_stats()
        st_g.close()
        print(f"# ensemble batched, captured: {steps} steps in {wall_g:.3f} s", file=sys.stderr, flush=True)
    wall_p = timed(singles)
    print(f"# ensemble plain: {steps} steps in {wall_p:.3f} s", file=sys.stderr, flush=True)
    U_full = st_e.download()
    U_e = ens.split(U_full)
    for m in sorted({0, len(wls) // 2, len(wls) - 1}):
        assert np.array_equal(U_e[m], singles[m].download()), f"member {m} differs from its own stepper"
    stages = 3
    value = nodes * stages * steps / wall_e
    plain = nodes * stages * steps / wall_p
    out = {
        "workload": f"{len(wls)} four-quadrant 2-D Euler boxes, h=1e-3, node counts log-uniform in "
                    f"[{args.n_min}, {args.n_max}] (seed {args.seed}), each its own seeded IC, jitter and CFL step; "
                    f"MUSCL2+VK(hard)+Rusanov+MOOD(EPD1,u1), SSP-RK3, CFL 0.4; {steps} timed steps after {warmup} "
                    f"warm-up steps, every form from the same initial states",
        "members": len(wls), "nodes": nodes, "ensemble_nodes_with_padding": int(ens.N),
        "smallest": int(ens.sizes.min()), "largest": int(ens.sizes.max()), "setup_s": round(t_setup, 1),
        "tuning": tuning,
        "batched": {"value": value, "ms_per_step": 1e3 * wall_e / steps,
                    "form": "one ensemble grid, one stepper (hcc_grid_concat + hcc_step_cfl)"},
        "plain": {"value": plain, "ms_per_step": 1e3 * wall_p / steps,
                  "form": "one stepper per member, stepped in turn (hcc_step_cfl)"},
        "batched_vs_plain": value / plain,
    }
    if wall_g is not None:
        out["batched_graph"] = {"value": nodes * stages * steps / wall_g, "ms_per_step": 1e3 * wall_g / steps,
                                "launches": g_stats[0], "captures": g_stats[1], "reruns": g_stats[2],
                                "form": "the batched form, each step one hipGraph launch (bit-identical states)"}
    if wall_i is not None:
        U_i = st_i.download()
        dev = float(np.abs(U_i - U_full).sum() / np.abs(U_full).sum())
        v_i = nodes * stages * steps / wall_i
        out["batched_indicator"] = {"value": v_i, "ms_per_step": 1e3 * wall_i / steps, "vs_plain": v_i / plain,
                                    "vs_batched": v_i / value, "relative_l1_deviation_from_plain_mood": dev,
                                    "form": "the batched form with the shipped indicator network"}
    for st in singles + [st_e] + ([st_i] if st_i is not None else []):
        st.close()
    if net is not None:
        net.close()
    for g in grids:
        g.close()
    ens.close()
    if own_ctx:
        ctx.close()
    return out


# The reference's published step benchmark (benchmark/step.jl; docs/src/gpu_roadmap.md:202-216, Float64, one RK3
# step of linear advection with MUSCL3 on a uniform periodic cloud): milliseconds per step on an RX 9070 XT (gfx1201)
# and on an i5-13600K limited to 8 cores.
REFERENCE_STEP = {   # (n per axis, equations, stab): (reference GPU ms, reference CPU ms)
    (1000, 1, "both"): (29.0, 387.0),
    (500, 1, "both"): (8.7, 98.0),
    (500, 1, "none"): (5.3, 71.0),
    (500, 1, "limiter"): (7.1, 87.0),
    (500, 1, "mood"): (7.3, 82.0),
}


def reference_step_run(ctx, steps=30, warmup=3, only=None):
    """benchmark/step.jl on this engine: for each case the grid, the stepper, dt = 0.3 get_time_step at t = 0 and
    then warm-up and timed steps ts(eq, pg, k dt, dt) with that fixed dt (hcc_step), wall clock over the timed steps;
    the checksum is the reference's sum of rho_1 over the particles."""
    import hcc_hip
    from hcc_hip import Grid, Stepper
    from hcc_hip import workloads as W
    sync = lambda: hcc_hip.lib().hcc_ctx_synchronize(ctx.h)   # noqa: E731
    cases = []
    for (n, M, stab), (ref_gpu, ref_cpu) in REFERENCE_STEP.items():
        if only is not None and (n, M, stab) != only:
            continue
        wl = W.step_benchmark(n=n, M=M, stab=stab)
        c = wl["cloud"]
        g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
        st = Stepper(ctx, g, wl["scheme"])
        st.upload(wl["U0"])
        dt = wl["dt_factor"] * st.time_step()
        for k in range(warmup):
            st.step((k + 1) * dt, dt)
        sync()
        st.totals(reset=True)
        t0 = time.perf_counter()
        for k in range(steps):
            st.step((warmup + k + 1) * dt, dt)
        sync()
        wall = time.perf_counter() - t0
        n_steps, iters, tot_ms, tot_g, tot_f = st.totals(reset=True)
        U = st.download()
        ms = 1e3 * wall / steps
        cases.append({"case": f"2-D, {stab}, {M} equation{'s' if M > 1 else ''}", "particles": g.N, "stab": stab,
                      "ms_per_step": round(ms, 4), "reference_gpu_ms": ref_gpu, "reference_cpu_ms": ref_cpu,
                      "speedup_vs_reference_gpu": round(ref_gpu / ms, 2),
                      "node_stage_updates_per_s": g.N * 3 * steps / wall,
                      "gather_avg_ms": round(tot_g / (3 * n_steps), 4), "flux_avg_ms": round(tot_f / (3 * n_steps), 4),
                      "mood_iterations": iters, "checksum": float(np.sum(U[:, 0]))})
        st.close()
        g.close()
    return {"workload": "the reference's own benchmark (benchmark/step.jl): linear advection, MUSCL3 + RusanovFlux, "
                        "SSP-RK3, uniform periodic n^2 lattice of the unit square, R = 2.5 dx, Gaussian state; stab "
                        "both = VenkatakrishnanLimiter(:hard) + MOOD(EPD1, MOODu1(1e-6)); dt = 0.3 get_time_step, "
                        f"fixed; {warmup} warm-up and {steps} timed steps",
            "reference_hardware": "RX 9070 XT (gfx1201, ROCm 7.2) Float64 GPU, i5-13600K limited to 8 cores "
                                  "(docs/src/gpu_roadmap.md, 2026-10-15)",
            "cases": cases}


def cube_run(ctx, steps=5, warmup=1):
    """The 3-D Euler cube with MLS order 3 (the pressure pulse on the jittered periodic 160^3 cube, R = 2.2 h,
    MUSCL3 + VK(hard) + Rusanov + MOOD(EPD1, u1), SSP-RK3, CFL 0.3): warm-up and timed CFL steps from t = 0."""
    import hcc_hip
    from hcc_hip import Grid, Stepper
    from hcc_hip import workloads as W
    sync = lambda: hcc_hip.lib().hcc_ctx_synchronize(ctx.h)   # noqa: E731
    t_setup = time.perf_counter()
    wl = W.euler_3d(n=160, seed=0)
    wl["scheme"] = dict(wl["scheme"], muscl_order=3)
    c = wl["cloud"]
    g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
    st = Stepper(ctx, g, wl["scheme"])
    st.upload(wl["U0"])
    sync()
    setup_s = time.perf_counter() - t_setup
    for _ in range(warmup):
        st.step_cfl(wl["cfl"])
    sync()
    st.totals(reset=True)
    wall = timed_steps(st, steps, wl["cfl"], sync)
    n_steps, iters, tot_ms, tot_g, tot_f = st.totals(reset=True)
    U = st.download()
    out = {"workload": "3-D Euler pressure pulse (rest state, amplitude 0.5, width 0.05) on the jittered periodic "
                       "160^3 cube, R = 2.2 h; MUSCL3 (degree-2 reconstruction, degree-3 divergence) + VK(hard) + "
                       f"Rusanov + MOOD(EPD1, u1), SSP-RK3, CFL 0.3; {warmup} warm-up and {steps} timed CFL steps "
                       "from t = 0 (MLS order 3 of config[2] with MUSCL in place of meshless WENO)",
           "nodes": g.N, "mean_neighbours": round(g.S / g.N, 1), "setup_s": round(setup_s, 1),
           "value": g.N * 3 * steps / wall, "ms_per_step": 1e3 * wall / steps,
           "flux_avg_ms": tot_f / (3 * n_steps), "gather_avg_ms": tot_g / (3 * n_steps), "mood_iterations": iters,
           "finite": bool(np.all(np.isfinite(U)))}
    st.close()
    g.close()
    return out


def graph_compare(args):
    """hcc_step_cfl uncaptured against hipGraph-captured steps (hcc_stepper_set_graph) on the four-quadrant problem at
    each h = 1/n: the same 5 warm-up and `steps` timed CFL steps from t = 0, and a bit-for-bit comparison of the final
    states. One JSON line per size."""
    import hcc_hip
    from hcc_hip import Context, Grid, Stepper
    from hcc_hip import workloads as W
    ctx = Context(0)
    sync = lambda: hcc_hip.lib().hcc_ctx_synchronize(ctx.h)   # noqa: E731
    for n in args.graph_compare:
        wl = W.euler_2d(n=n, seed=args.seed)
        c = wl["cloud"]
        g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
        res = {}
        for mode in ("uncaptured", "graph"):
            st = Stepper(ctx, g, wl["scheme"])
            st.upload(wl["U0"])
            st.set_graph(mode == "graph")
            for _ in range(5):
                st.step_cfl(wl["cfl"])
            sync()
            wall = timed_steps(st, args.steps, wl["cfl"], sync)
            res[mode] = {"ms_per_step": 1e3 * wall / args.steps, "value": g.N * 3 * args.steps / wall,
                         "U": st.download()}
            if mode == "graph":
                la, ca, re, bo = st.graph_stats()
                res[mode].update({"launches": la, "captures": ca, "reruns": re, "iterations_captured": bo.tolist()})
            st.close()
        same = bool(np.array_equal(res["uncaptured"].pop("U"), res["graph"].pop("U")))
        print(json.dumps({"nodes_per_axis": n, "nodes": g.N, "steps": args.steps, "bit_identical": same,
                          "graph_vs_uncaptured": res["uncaptured"]["ms_per_step"] / res["graph"]["ms_per_step"],
                          **res}), flush=True)
        g.close()
    ctx.close()


def ensemble_main(args):
    out = ensemble_run(args)
    b = out["batched_indicator"] if "batched_indicator" in out else out["batched"]
    line = {"metric": "node-stage updates/s, batched ensemble of 2D Euler MUSCL+MOOD instances", "value": b["value"],
            "unit": "node-stage updates/s", "n_gpus": 1, "steps": args.ensemble_steps,
            "warmup": args.ensemble_warmup, "ms_per_step": b["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": out["workload"], "members": out["members"], "nodes": out["nodes"]},
            "ensemble": out}
    print(json.dumps(line), flush=True)


VARIANTS = ("moving", "imex", "epd2_u2", "strict_epd0", "streamed", "aligned", "solve", "euler3d", "euler3d_muscl3",
            "muscl3")


def variant_workload(W, L, name, n, seed):
    """The 1 M-node four-quadrant problem with one of the scheme variants built beside the headline path."""
    import numpy as np
    h = 1.0 / n
    wl = W.euler_2d(h=h, seed=seed)
    bcs = wl["scheme"]["bcs"]
    kw = dict(delta=1e-4 * h * h, bcs=bcs)
    if name == "moving":   # Lagrangian particles on the periodic unit square (movers take no outflow conditions)
        dom = W.domain(2, (0.0, 0.0), (1.0, 1.0), (True, True), h, 2.5, 0.25, seed)
        cloud = L.cloud_generate(W.domain_desc(dom))
        x = cloud["pos"]
        rho_q, p_q, u_q, v_q = W.quadrant_states(seed, 1.4)
        q = (x[:, 0] >= 0.5).astype(int) + 2 * (x[:, 1] >= 0.5).astype(int)
        wl = {"domain": dom, "cloud": cloud, "cfl": 0.4,
              "U0": W.euler_conserved(rho_q[q], np.column_stack([u_q[q], v_q[q]]), p_q[q], 1.4),
              "scheme": W.scheme(4, L.HCC_PDE_EULER, delta=1e-4 * h * h, mover=L.HCC_MOVER_LAGRANGIAN)}
        return wl, "LagrangianMover on the periodic unit square (same four quadrants)"
    if name == "euler3d":   # the 3-D cube: 160^3 jittered periodic nodes (4.1 M), a pressure pulse, MUSCL2
        wl = W.euler_3d(n=160, seed=seed)
        return wl, ("3-D Euler pressure pulse on the jittered periodic 160^3 cube (4.1 M nodes, R = 2.2 h), "
                    "MUSCL2+VK(hard)+Rusanov+MOOD(EPD1,u1), SSP-RK3, CFL 0.3 (neighbours through the caches)")
    if name == "euler3d_muscl3":   # the same cube with MLS order 3: degree-2 reconstruction, degree-3 divergence
        wl = W.euler_3d(n=160, seed=seed)
        wl["scheme"] = dict(wl["scheme"], muscl_order=3)
        return wl, ("3-D Euler pressure pulse on the jittered periodic 160^3 cube (4.1 M nodes, R = 2.2 h), "
                    "MUSCL3+VK(hard)+Rusanov+MOOD(EPD1,u1) (degree-2 reconstruction, degree-3 divergence where 2 nb "
                    ">= 60, else degree 2), SSP-RK3, CFL 0.3 (neighbours through the caches)")
    if name == "muscl3":   # the headline problem with MUSCL order 3
        wl["scheme"] = dict(wl["scheme"], muscl_order=3)
        return wl, "MUSCL3+VK(hard)+Rusanov+MOOD(EPD1,u1) (degree-2 reconstruction, degree-3 divergence)"
    if name == "imex":
        kw.update(tableau_name="IMEX_ARS233")
        desc = "IMEX ARS(2,3,3) with the implicit decay source, rate 2 (1 + 0.5 sin 2 pi x)"
    elif name == "epd2_u2":
        kw.update(strategy=L.HCC_STRAT_EPD2, criterion=L.HCC_CRIT_U2, plateau=0.05)
        desc = "MOOD(EPD2, MOODu2 with plateau 0.05)"
    elif name == "strict_epd0":
        kw.update(strategy=L.HCC_STRAT_STRICT_EPD0)
        desc = "MOOD(StrictEPD0, MOODu1)"
    elif name == "streamed":
        kw.update(streamed=True)
        desc = "MUSCL streamed divergence"
    elif name == "aligned":
        kw.update(flux=L.HCC_FLUX_ALIGNED_RUSANOV)
        desc = "AlignedRusanovFlux"
    else:
        desc = "the headline scheme through hcc_solve (the device driver loop, snapshots at the ends only)"
    wl["scheme"] = W.scheme(4, L.HCC_PDE_EULER, **kw)
    return wl, desc


def variant_main(args):
    """One variant of the step on the 1 M-node problem: node-stage updates/s over the timed CFL steps."""
    import hcc_hip
    from hcc_hip import Context, Grid, Stepper
    from hcc_hip import _lib as L
    from hcc_hip import workloads as W
    wl, desc = variant_workload(W, L, args.variant, args.n, args.seed)
    c = wl["cloud"]
    ctx = Context(0)
    g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
    st = Stepper(ctx, g, wl["scheme"])
    st.upload(wl["U0"])
    if args.variant == "imex":
        st.set_source(L.HCC_SRC_DECAY_IMPLICIT, 2.0 * (1 + 0.5 * np.sin(2 * np.pi * c["pos"][:, 0])))
    for _ in range(args.warmup):
        st.step_cfl(wl["cfl"])
    hcc_hip.lib().hcc_ctx_synchronize(ctx.h)
    st.totals(reset=True)
    if args.variant == "solve":   # the driver loop to t = 0.005 (about 20 CFL steps), timed by hcc_solve itself
        _, _, _, n_run, wall, _ = st.solve(0.005, wl["cfl"], True, 1)
        args.steps = int(n_run)
    else:
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st.step_cfl(wl["cfl"])
        hcc_hip.lib().hcc_ctx_synchronize(ctx.h)
        wall = time.perf_counter() - t0
    n_steps, iters, tot_ms, _, _ = st.totals(reset=True)
    U = st.download()
    assert np.all(np.isfinite(U)), "non-finite state"
    stages = len(wl["scheme"]["b_t"])
    out = {"metric": f"node-stage updates/s, Euler MUSCL+MOOD, variant {args.variant}",
           "value": g.N * stages * args.steps / wall, "unit": "node-stage updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": desc if args.variant.startswith("euler3d") else
                      f"2-D Euler four-quadrant, h=1/{args.n}, " + desc,
                      "nodes": g.N, "stages": stages,
                      "variant": args.variant},
           "detail": {"step_event_ms": tot_ms / max(n_steps, 1), "mood_iterations": iters}}
    print(json.dumps(out), flush=True)


def host_cpu():
    """The host's CPU model and core counts (the CPU baseline's context)."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def profile_value(name, key):
    """A figure recorded in a committed rocprofv3 summary under profiles/ (a line `key = value`), or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + " ="):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


def timed_steps(st, k, cfl, sync):
    """k driver-loop steps (CFL estimate + step, dt kept on the device), wall seconds."""
    t0 = time.perf_counter()
    for _ in range(k):
        st.step_cfl(cfl)
    sync()
    return time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--pass-every", type=int, default=4,
                    help="record the gather / flux pass events every k-th timed step (1: every step)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1000, help="nodes per axis (h = 1/n)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--run-steps", type=int, default=200, help="steps of the configured run (config: 200)")
    ap.add_argument("--window", type=int, default=100,
                    help="the timed steps start after this many steps of the run (set-up steps + warm-up steps)")
    ap.add_argument("--cpu-steps", type=int, default=25)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true", help="leave out the batched-ensemble block")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch choices")
    ap.add_argument("--timed-only", action="store_true",
                    help="only the set-up, warm-up and timed steps (no full run, dense state, ensemble or CPU "
                         "baseline): the short run the rocprofv3 summaries under profiles/ profile")
    ap.add_argument("--host-dt", action="store_true",
                    help="return dt to the host between the CFL estimate and the step (hcc_time_step + hcc_step)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device-to-device copy (counter calibration for profiles/)")
    ap.add_argument("--ensemble", action="store_true",
                    help="only the batched form: many Euler boxes of uneven size in one ensemble grid, against a "
                         "stepper per box")
    ap.add_argument("--indicator", action="store_true",
                    help="the headline run with the learned troubled-node indicator (the shipped network)")
    ap.add_argument("--no-ensemble-indicator", dest="ens_indicator", action="store_false",
                    help="the ensemble block without its learned-indicator form")
    ap.add_argument("--variant", choices=VARIANTS, default=None,
                    help="a scheme variant built beside the headline path, on the same 1 M-node problem")
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--n-min", type=int, default=20000)
    ap.add_argument("--n-max", type=int, default=400000)
    ap.add_argument("--ensemble-steps", type=int, default=10)
    ap.add_argument("--ensemble-warmup", type=int, default=2)
    ap.add_argument("--graph-form", action="store_true",
                    help="the ensemble block also times the batched form as hipGraph-captured steps")
    ap.add_argument("--reference-step", action="store_true",
                    help="only the reference's own step benchmark (benchmark/step.jl, MUSCL3 advection)")
    ap.add_argument("--no-reference-step", action="store_true", help="leave out the reference step-benchmark block")
    ap.add_argument("--no-cube", action="store_true", help="leave out the 3-D MLS-order-3 cube block")
    ap.add_argument("--reference-case", default=None, metavar="N,M,STAB",
                    help="with --reference-step: only this case (e.g. 1000,1,both)")
    ap.add_argument("--graph-compare", type=int, nargs="+", default=None, metavar="N",
                    help="only compare uncaptured and hipGraph-captured steps at h = 1/N for each N")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs the time step on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is designed, not built)", file=sys.stderr)
        sys.exit(2)
    if args.graph_compare:
        graph_compare(args)
        return
    if args.reference_step:
        from hcc_hip import Context
        only = None
        if args.reference_case:
            n_, m_, stab_ = args.reference_case.split(",")
            only = (int(n_), int(m_), stab_)
        print(json.dumps({"reference_step_benchmark": reference_step_run(Context(0), only=only)}), flush=True)
        return
    if args.ensemble:
        args.indicator = args.indicator or args.ens_indicator
        ensemble_main(args)
        return
    if args.variant:
        variant_main(args)
        return

    import hcc_hip
    from hcc_hip import Context, Grid, Stepper
    from hcc_hip import workloads as W

    t_setup = time.perf_counter()
    wl = W.euler_2d(h=1.0 / args.n, seed=args.seed)
    c = wl["cloud"]
    ctx = Context(0)
    sync = lambda: hcc_hip.lib().hcc_ctx_synchronize(ctx.h)   # noqa: E731
    g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
    st = Stepper(ctx, g, wl["scheme"])
    net = None
    if args.indicator:
        from hcc_hip import MLP
        from hcc_hip import indicator as IND
        net = MLP(ctx, IND.default_params())
        st.set_indicator(net, IND.DEFAULT_THRESHOLD)
    cfl = wl["cfl"]
    # The timed window sits inside the configured run: its input is the state after `first` steps from t = 0,
    # formed on the device during set-up (untimed), then W warm-up steps, then the K timed steps.
    first = max(0, args.window - args.warmup)
    st.upload(wl["U0"])
    for _ in range(first):
        st.step_cfl(cfl)
    sync()
    tuning = None
    if not args.no_autotune:   # launch choices timed on this state (restored afterwards), set-up time
        knobs, ms_best, ms_def, timed = st.autotune(steps=3, cfl=cfl)
        tuning = dict(knobs, ms_per_step_best=round(ms_best, 4), ms_per_step_default=round(ms_def, 4), timed=timed)
    t_setup = time.perf_counter() - t_setup
    print(f"# setup {t_setup:.1f} s: N={g.N} nodes, S={g.S} slots, max_nb={g.max_nb}; {first} set-up steps",
          file=sys.stderr, flush=True)

    t = 0.0
    for _ in range(args.warmup):
        st.step_cfl(cfl)
    sync()
    st.totals(reset=True)
    # the timed loop is the driver loop of the reference: each step is the CFL estimate (get_time_step) and the
    # stepper. By default both run back to back on the device (hcc_step_cfl: dt = CFL * CFL step kept on the
    # device); --host-dt returns dt to the host between them (hcc_time_step, hcc_step). The per-step HIP events are
    # read after the loop (hcc_step_totals), so timing adds no host round trip per step. The events around the gather
    # and flux passes are recorded every --pass-every-th step of the region (each is a marker packet that idles the
    # device a few microseconds); the pass averages are over those steps.
    st.set_pass_timing(args.pass_every)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        if args.host_dt:
            dt = cfl * st.time_step()
            st.step(t, dt)
            t += dt
        else:
            st.step_cfl(cfl)
    sync()
    wall = time.perf_counter() - t0
    n_timed = st.timed_passes()
    n_steps, iters, tot_ms, tot_g_ms, tot_f_ms = st.totals(reset=True)
    assert n_steps == args.steps, (n_steps, args.steps)
    assert n_timed >= 1, "no step of the timed region timed its passes"
    st.set_pass_timing(1)
    flags = int(st.mood_flags().sum())
    U = st.download()
    assert np.all(np.isfinite(U)), "non-finite state"

    N = g.N
    isb = c["is_boundary"].astype(bool)
    N_int = int((~isb).sum())
    ranges, _, _, _ = g.neighbors(weights=False, distances=False)
    counts = np.diff(ranges)
    S_int = int(counts[~isb].sum())
    nnb = S_int / max(N_int, 1)
    D, M, stages = 2, 4, 3
    value = N * stages * args.steps / wall
    flux_avg_ms = tot_f_ms / (stages * n_timed)
    gather_avg_ms = tot_g_ms / (stages * n_timed)
    fb = flux_pass_bytes(N, N_int, S_int, D, M)
    gb = gather_pass_bytes(N_int, S_int, D, M)
    achieved = fb / (flux_avg_ms * 1e-3) / 1e9
    model_b = W.bytes_per_node_stage(D, M, nnb)

    if args.timed_only:
        args.no_ensemble = args.no_cpu_baseline = args.no_reference_step = args.no_cube = True
    # the whole configured run from t = 0 (run_steps CFL steps), timed as one region
    st.upload(wl["U0"])
    sync()
    wall_run = timed_steps(st, args.run_steps, cfl, sync) if not args.timed_only else float("nan")
    # the dense state: the four quadrants times a smooth factor, so that no neighbourhood is constant and every
    # slot of every pass does its full arithmetic (the same cloud and scheme)
    x = c["pos"]
    Ud = wl["U0"].copy()
    fac = 1.0 + 0.01 * np.sin(2 * np.pi * 3 * x[:, 0]) * np.sin(2 * np.pi * 5 * x[:, 1])
    Ud[:, 0] *= fac
    Ud[:, 3] *= fac
    st.upload(Ud)
    dense_k = 0 if args.timed_only else 10
    for _ in range(2 if dense_k else 0):
        st.step_cfl(cfl)
    sync()
    st.totals(reset=True)
    wall_dense = timed_steps(st, dense_k, cfl, sync) if dense_k else float("nan")
    nd_, it_d, tot_d, g_d, f_d = st.totals(reset=True)
    nd_ = max(nd_, 1)

    traffic = profile_value("hbm_traffic_euler2d_1M.md", "traffic_bytes k_flux_staged")
    valu = profile_value("sq_counters_euler2d_1M.md", "valu_issue_share k_flux_staged")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "node-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2-D Euler four-quadrant, jittered h=1/{args.n} unit square with outflow ghost band, "
                               f"MUSCL2+VK(hard)+Rusanov+MOOD(EPD1,u1), SSP-RK3, CFL 0.4, {args.run_steps}-step run",
                   "nodes": N, "interior_nodes": N_int, "mean_neighbours": round(nnb, 3), "stages": stages,
                   "seed": args.seed, "learned_indicator": bool(args.indicator), "tuning": tuning,
                   "timed_steps": f"steps {first + args.warmup + 1}..{first + args.warmup + args.steps} of the "
                                  f"{args.run_steps}-step run from t = 0 (the first {first} formed on the GPU in "
                                  f"set-up, then {args.warmup} warm-up steps)"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/hbm_traffic_euler2d_1M.md (rocprofv3 FETCH_SIZE / WRITE_SIZE of this "
                                       "kernel, calibrated on a device copy)",
                     "fp64_valu_issue_frac": valu,
                     "fp64_valu_issue_source": "profiles/sq_counters_euler2d_1M.md (SQ_INSTS_VALU over 256 CUs x "
                                               "2.4 GHz x kernel time)",
                     "kernel": "k_flux_staged<2,4,Euler> (full pass: flux, MOOD check, stage predictor)",
                     "bytes_per_launch": fb, "avg_launch_ms": flux_avg_ms,
                     "avg_launch_source": f"HIP events around the pass on its stream, {n_timed * stages} launches: "
                                          f"every {args.pass_every}th step of the timed region"},
        "detail": {"gather_avg_ms": gather_avg_ms, "gather_bytes": gb, "flux_avg_ms": flux_avg_ms,
                   "step_event_ms": tot_ms / n_steps, "mood_iterations": iters, "flags_last_step": flags,
                   "model_bytes_per_node_stage": model_b,
                   "model_roofline_frac": value * model_b / (HBM_PEAK_GBS * 1e9)},
    }
    if not args.timed_only:
        out["detail"].update({
            "full_run": {"steps": args.run_steps, "ms_per_step": 1e3 * wall_run / args.run_steps,
                         "value": N * stages * args.run_steps / wall_run,
                         "form": "every step of the run from t = 0, one timed region"},
            "dense_state": {"steps": dense_k, "ms_per_step": 1e3 * wall_dense / dense_k,
                            "value": N * stages * dense_k / wall_dense,
                            "gather_avg_ms": g_d / (stages * nd_), "flux_avg_ms": f_d / (stages * nd_),
                            "mood_iterations": it_d,
                            "form": "the four quadrants times 1 + 0.01 sin(6 pi x) sin(10 pi y): no constant "
                                    "neighbourhood, every slot computed in full"}})
    st.close()
    if net is not None:
        net.close()
    g.close()
    if not args.no_ensemble:
        ea = argparse.Namespace(**vars(args))
        ea.indicator = args.ens_indicator
        out["ensemble"] = ensemble_run(ea, ctx)
    if not args.no_reference_step:
        out["reference_step_benchmark"] = reference_step_run(ctx)
    if not args.no_cube:
        out["cube_3d_mls3"] = cube_run(ctx)
    if not args.no_cpu_baseline:
        # the oracle on the host cores of this process: all of them available to it, at most the 16 that a one-GPU
        # share of the box holds
        host = host_cpu()
        threads = max(1, min(host["affinity"], 16))
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_steps, threads)
        out["cpu_baseline"].update(host)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
