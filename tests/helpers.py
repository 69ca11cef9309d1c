"""Shared set-up of the parity tests: the same workload through the HIP path and through the CPU oracle."""
import numpy as np

from hcc_hip import Grid, Stepper
from hcc_hip import workloads as W
from oracle import oracle as O


def build_pair(ctx, wl, threads=8):
    c = wl["cloud"]
    g = Grid(ctx, W.domain_desc(wl["domain"]), c["pos"], c["is_boundary"], c["tags"])
    st = Stepper(ctx, g, wl["scheme"])
    st.upload(wl["U0"])
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], wl["scheme"], threads=threads)
    o.set_state(wl["U0"])
    return g, st, o


def first_diff(a, b):
    a, b = np.asarray(a), np.asarray(b)
    bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))))
    return None if len(bad) == 0 else (tuple(bad[0]), a[tuple(bad[0])], b[tuple(bad[0])], len(bad))


def rel_linf(a, b):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) / scale
