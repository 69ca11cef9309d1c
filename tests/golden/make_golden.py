"""Writes the small golden fixtures of the hot path (tests/golden/*.npz) from the CPU oracle.

The reference (Julia) cannot run in this environment, so these vectors are the oracle's outputs on tiny seeded
inputs, after the oracle has been pinned by tests/test_oracle.py. They freeze the exact bits of every stage of the
path (neighbour lists, caches, one gather+flux pass, full RK3 steps with MOOD firing) so that a later change to
either the oracle or the kernels that moves a single bit is caught. Regenerate only deliberately:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "hypercloudcore.jl_amd")):
    sys.path.insert(0, p)

from hcc_hip import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402


def cases():
    adv = W.advection_2d(8, 3)                         # 64 nodes, periodic
    eul = W.euler_2d(n=12, seed=5)                     # about 300 nodes with the ghost band
    box = W.euler_2d(n=16, seed=9)
    return {"adv2d_64": (adv, 3), "euler2d_n12": (eul, 3), "euler2d_n16": (box, 4)}


def run(wl, steps):
    c = wl["cloud"]
    o = O.OracleSim(wl["domain"], c["pos"], c["is_boundary"], c["tags"], wl["scheme"])
    o.set_state(wl["U0"])
    r, idx, w, dist = o.neighbors()
    ig, og, idv, od, dw = o.moments()
    K, G = o.rhs(wl["U0"])
    out = dict(pos=c["pos"], is_boundary=c["is_boundary"], tags=c["tags"], U0=wl["U0"], ranges=r, indices=idx,
               weights=w, distances=dist, inv_grad=ig, ok_grad=og, inv_div=idv[:, 1:, :], ok_div=od, dweights=dw,
               K0=K, G0=G)
    t = 0.0
    for k in range(steps):
        dt = wl["cfl"] * o.time_step()
        o.step(t, dt)
        t += dt
        out[f"dt{k}"] = np.array(dt)
        out[f"U{k + 1}"] = o.state()
        out[f"orders{k + 1}"] = o.orders()
        out[f"flags{k + 1}"] = o.mood_flags()
    out["steps"] = np.array(steps)
    return out


if __name__ == "__main__":
    for name, (wl, steps) in cases().items():
        data = run(wl, steps)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **data)
        fl = sum(int(data[f"flags{k + 1}"].sum()) for k in range(steps))
        print(name, "N =", len(data["pos"]), "flags over the steps:", fl)
