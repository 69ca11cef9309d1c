"""The learned troubled-node indicator (2-D Euler): training data from plain MOOD runs, training on the device, the
shipped parameters, and its evaluation against the plain step.

A sample is a particle at the full order at the start of a stage's flux pass: its 16 features (stored by the full
gather) and its label, 1 if MOOD's checkpoint of that stage dropped it to first order. Negatives vastly outnumber
positives (about 1%); a seeded fraction of them is kept and the loss weights the positives accordingly.

    python -m hcc_hip.indicator --out params.npy [--seeds 32 --n 500 --steps 4]

trains the shipped network (indicator_params.npy beside this file): 32 seeded four-quadrant runs at 250 k nodes.

    python -m hcc_hip.indicator --rates [--n 1000 --eval-seeds 901 902 --steps 20]

judges the shipped network by plain MOOD on unseen seeds: its miss and false-flag rates (decision_rates).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import _lib as L
from . import workloads as W

PARAMS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "indicator_params.npy")
DEFAULT_THRESHOLD = 0.0


def default_params():
    """The shipped trained parameters (float32[HCC_MLP_PARAMS])."""
    return np.load(PARAMS_PATH, allow_pickle=False).astype(np.float32)


def init_params(seed=0):
    """Glorot-uniform weights, zero biases."""
    F, H = 16, 64
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (F, H)) * np.sqrt(6.0 / (F + H))
    W2 = rng.uniform(-1, 1, (H, H)) * np.sqrt(6.0 / (H + H))
    w3 = rng.uniform(-1, 1, H) * np.sqrt(6.0 / (H + 1))
    return np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), w3, [0.0]]).astype(np.float32)


def collect_samples(ctx, n, seeds, steps, neg_keep=1.0, cfl=0.4, sample_seed=0, log=None):
    """Features (bf16 bit rows) and labels of every stage of `steps` CFL steps of the four-quadrant problem at
    h = 1/n for each seed; negatives kept with probability neg_keep. Returns (X, y, positive rate before thinning)."""
    rng = np.random.default_rng(sample_seed)
    Xs, ys, n_all, n_pos = [], [], 0, 0
    for seed in seeds:
        wl = W.euler_2d(n=n, seed=seed)
        c = wl["cloud"]
        g = L.Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
        st = L.Stepper(ctx, g, wl["scheme"])
        st.upload(wl["U0"])
        st.collect(True)
        for _ in range(steps):
            st.step_cfl(cfl)
            for s in range(len(wl["scheme"]["b_t"])):
                f, pre, post = st.samples(s)
                live = pre == 2
                lab = (post == 1) & live
                n_all += int(live.sum())
                n_pos += int(lab.sum())
                keep = live & (lab | (rng.uniform(size=live.size) < neg_keep))
                Xs.append(f[keep])
                ys.append(lab[keep].astype(np.uint8))
        st.close()
        g.close()
        if log:
            log(f"seed {seed}: {sum(len(y) for y in ys)} samples kept, {n_pos} positives")
    X, y = np.concatenate(Xs), np.concatenate(ys)
    perm = rng.permutation(len(y))
    return X[perm], y[perm], n_pos / max(n_all, 1)


def train_indicator(ctx, X, y, epochs=6, batch=4096, lr=2e-3, seed=0, pos_weight=None):
    """An MLP trained on the device (Adam); pos_weight defaults to half the kept negative / positive ratio (at most
    50). Returns (MLP, epoch losses)."""
    pos = float(np.mean(y))
    if pos_weight is None:
        pos_weight = float(min(50.0, 0.5 * (1 - pos) / max(pos, 1e-6)))
    net = L.MLP(ctx, init_params(seed))
    losses = net.train(X, y, epochs=epochs, batch=batch, lr=lr, pos_weight=pos_weight)
    return net, losses


def compare(ctx, net, n, seed, steps, threshold=DEFAULT_THRESHOLD, cfl=0.4):
    """The plain step against the step with the indicator on the same problem (CFL steps): per form the seconds per
    step (wall and HIP events), MOOD fixed-point iterations, the particles the indicator lowered in the last step; and
    the relative L1 deviation of the final states."""
    out = {}
    for form in ("plain", "indicator"):
        wl = W.euler_2d(n=n, seed=seed)
        c = wl["cloud"]
        g = L.Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
        st = L.Stepper(ctx, g, wl["scheme"])
        st.upload(wl["U0"])
        if form == "indicator":
            st.set_indicator(net, threshold)
        st.totals(reset=True)
        L.lib().hcc_ctx_synchronize(ctx.h)
        t0 = time.perf_counter()
        for _ in range(steps):
            st.step_cfl(cfl)
        L.lib().hcc_ctx_synchronize(ctx.h)
        wall = time.perf_counter() - t0
        _, iters, ms, _, _ = st.totals(reset=True)
        out[form] = {"s_per_step": wall / steps, "event_ms_per_step": ms / steps, "mood_iterations": iters,
                     "lowered_last_step": st.indicator_count() if form == "indicator" else 0, "U": st.download()}
        st.close()
        g.close()
    U0, U1 = out["plain"].pop("U"), out["indicator"].pop("U")
    out["relative_l1_deviation"] = float(np.abs(U1 - U0).sum() / np.abs(U0).sum())
    return out


def decision_rates(ctx, net, n, seeds, steps, thresholds=(DEFAULT_THRESHOLD,), cfl=0.4):
    """The network judged by plain MOOD on the same stage states: plain-MOOD runs of the four-quadrant problem at
    h = 1/n (CFL steps) collect, for every stage, the features of each particle at the full order and MOOD's label
    (its checkpoint dropped the particle). The network's decision on those features is what it would decide at that
    stage state: flag = the particle is a candidate (its neighbourhood is not constant: a feature range is positive;
    the step runs the network on candidates only) and its logit exceeds the threshold. Per threshold: the confusion
    counts over all stage samples, the miss rate (MOOD drops the network leaves at full order: FN / (TP + FN)), the
    false-flag rate (flagged particles MOOD keeps: FP / (FP + TN)), the same over candidates only, and precision."""
    zs, labs, live_n, cand_n = [], [], 0, 0
    for seed in seeds:
        wl = W.euler_2d(n=n, seed=seed)
        c = wl["cloud"]
        g = L.Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
        st = L.Stepper(ctx, g, wl["scheme"])
        st.upload(wl["U0"])
        st.collect(True)
        for _ in range(steps):
            st.step_cfl(cfl)
            for s in range(len(wl["scheme"]["b_t"])):
                f, pre, post = st.samples(s)
                live = pre == 2
                lab = (post == 1) & live
                fv = (f[:, :4].astype(np.uint32) << 16).view(np.float32)
                idx = np.flatnonzero(live & np.any(fv > 0, axis=1))
                live_n += int(live.sum())
                cand_n += len(idx)
                # a particle that is not a candidate is never flagged, and MOOD cannot drop it (its neighbourhood is
                # constant): it counts as a true negative at every threshold
                assert int(lab.sum()) == int(lab[idx].sum()), "MOOD dropped a non-candidate"
                if len(idx):
                    zs.append(net.forward(np.ascontiguousarray(f[idx])))
                    labs.append(lab[idx])
        st.close()
        g.close()
    z, y = np.concatenate(zs), np.concatenate(labs)
    pos = int(y.sum())
    neg = live_n - pos
    out = {"nodes_per_axis": n, "seeds": list(seeds), "steps": steps, "samples": live_n, "candidates": cand_n,
           "mood_drops": pos, "thresholds": []}
    for thr in thresholds:
        flag = z > thr
        tp = int(np.sum(flag & y))
        fp = int(np.sum(flag & ~y))
        fn = pos - tp
        out["thresholds"].append({"threshold": thr, "tp": tp, "fn": fn, "fp": fp, "tn": neg - fp,
                                  "miss_rate": fn / max(pos, 1), "false_flag_rate": fp / max(neg, 1),
                                  "false_flag_rate_candidates": fp / max(cand_n - pos, 1),
                                  "precision": tp / max(tp + fp, 1)})
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--out", default=PARAMS_PATH)
    ap.add_argument("--seeds", type=int, default=32)
    ap.add_argument("--n", type=int, default=500)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--neg-keep", type=float, default=0.1)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--rates", action="store_true", help="judge the shipped network by plain MOOD (decision_rates)")
    ap.add_argument("--eval-seeds", type=int, nargs="+", default=[901, 902])
    ap.add_argument("--compare", action="store_true",
                    help="time the step with the shipped network at each threshold against the plain step (compare)")
    ap.add_argument("--thresholds", type=float, nargs="+", default=[DEFAULT_THRESHOLD])
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    ctx = L.Context(0)
    if args.compare:
        net = L.MLP(ctx, default_params())
        for thr in args.thresholds:
            for seed in args.eval_seeds:
                r = compare(ctx, net, args.n, seed, args.steps, thr)
                r.update({"threshold": thr, "seed": seed, "nodes_per_axis": args.n, "steps": args.steps})
                print(json.dumps(r), flush=True)
        net.close()
        ctx.close()
        return
    if args.rates:
        net = L.MLP(ctx, default_params())
        t0 = time.perf_counter()
        r = decision_rates(ctx, net, args.n, args.eval_seeds, args.steps, args.thresholds)
        r["seconds"] = round(time.perf_counter() - t0, 1)
        print(json.dumps(r), flush=True)
        net.close()
        ctx.close()
        return
    t0 = time.perf_counter()
    X, y, rate = collect_samples(ctx, args.n, range(1, args.seeds + 1), args.steps, args.neg_keep, log=log)
    log(f"collected {len(y)} samples ({int(y.sum())} positives, raw positive rate {rate:.4f}) in "
        f"{time.perf_counter() - t0:.1f} s")
    net, losses = train_indicator(ctx, X, y, epochs=args.epochs)
    log(f"trained: epoch losses {np.round(losses, 4).tolist()}")
    np.save(args.out, net.params())
    ctx.close()


if __name__ == "__main__":
    main()
