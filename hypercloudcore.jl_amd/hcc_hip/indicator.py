"""The learned troubled-node indicator (2-D Euler): training data from plain MOOD runs, training on the device, the
shipped parameters, and its evaluation against the plain step.

A sample is a particle at the full order at the start of a stage's flux pass: its 16 features (stored by the full
gather) and its label, 1 if MOOD's checkpoint of that stage dropped it to first order. Negatives vastly outnumber
positives (about 1%); a seeded fraction of them is kept and the loss weights the positives accordingly.

    python -m hcc_hip.indicator --out params.npy [--seeds 32 --n 500 --steps 4]

trains the shipped network (indicator_params.npy beside this file): 32 seeded four-quadrant runs at 250 k nodes.
"""
import argparse
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


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--out", default=PARAMS_PATH)
    ap.add_argument("--seeds", type=int, default=32)
    ap.add_argument("--n", type=int, default=500)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--neg-keep", type=float, default=0.1)
    ap.add_argument("--epochs", type=int, default=8)
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    ctx = L.Context(0)
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
