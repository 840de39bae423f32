#!/usr/bin/env python3
"""TEST INFRASTRUCTURE ONLY: how far the trajectory of the reference as
shipped (glibc sin/cos) stays from the one the library computes (the shared
sin/cos of include/axions/ax_trig.h), on the seeded 64^3 plain
Abelian-Higgs run (dx = 0.5, dt = 0.1, lambda = 2, g = eta = 1, seed 42;
the step of mpi_evolution.cpp:381-501).

Both trajectories are the oracle's (oracle/oracle.cpp), which is pinned bit
for bit against the reference compiled with each trig (fixtures
evo_ah_g1_libm and evo_ah_g1_shared / evo_c1_shared / evo_c2_1000_shared),
so the difference measured here is the difference between the reference's
own output and the library's.  Reported:
  - the first TimeStep after which the largest field difference, relative
    to the largest field magnitude of the current level, exceeds 1e-12 and
    1e-6;
  - energy, mean |Gauss| and xi (winding count) of both runs and their
    relative differences at TimeStep 200 and 1000;
  - the fraction of the sin/cos arguments of the step (link angles theta_a
    and the plaquette angles) at which ax_sin / ax_cos and glibc differ
    bitwise, sampled over the trajectory.

Usage: python oracle/libm_horizon.py [--n 64] [--steps 1000] [--threads 8]
(about 2 minutes at 64^3 x 1000 steps on 8 cores).
"""
import argparse
import json
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

from axions_amd.abi import ah_params  # noqa: E402
from oracle import oracle_py as orc  # noqa: E402

SHARED, LIBM = 0, 1


def field_rel(p, a, b, n):
    ca, _ = orc.levels(p, a, n)
    cb, _ = orc.levels(p, b, n)
    return float(np.max(np.abs(ca - cb)) / np.max(np.abs(cb)))


def xi(p, st, n, mode):
    del mode   # the winding detector is the library's own (no reference counterpart)
    w = orc.winding(p, st, n).reshape(3, -1)
    W = np.abs(w.astype(np.int64)).sum(axis=1)
    L = (2.0 / 3.0) * (W[0] * p.dz + W[1] * p.dy + W[2] * p.dx)
    return float(L / (p.nx * p.ny * p.nz * p.dx * p.dy * p.dz))


def trig_args(p, st, n):
    """Every argument the AH step passes to sin/cos at TimeStep n: the three
    link angles per site (Laplacian and currents) and the six plaquette
    angles per site (mpi_evolution.cpp:418-487)."""
    cur, _ = orc.levels(p, st, n)
    X, Y, Z = cur[2], cur[3], cur[4]

    def sh(a, ax):
        return np.roll(a, -1, axis=ax)
    a_xy = X + sh(Y, 0) - sh(X, 1) - Y
    a_xz = X + sh(Z, 0) - sh(X, 2) - Z
    a_yz = Y + sh(Z, 1) - sh(Y, 2) - Z
    a_yx = Y + sh(X, 1) - sh(Y, 0) - X
    a_zx = Z + sh(X, 2) - sh(Z, 0) - X
    a_zy = Z + sh(Y, 2) - sh(Z, 1) - Y
    return np.concatenate([v.ravel() for v in (X, Y, Z)]), \
        np.concatenate([v.ravel() for v in (a_xy, a_xz, a_yz, a_yx, a_zx, a_zy)])


def differing(x, fn_ax, fn_libm):
    a = orc.trig(fn_ax, x)
    b = orc.trig(fn_libm, x)
    return int(np.count_nonzero(a.view(np.uint64) != b.view(np.uint64)))


def run(n=64, steps=1000, threads=8, report=(200, 1000), sample_every=100, stop_at=None):
    p = ah_params(n)
    s = orc.init_random_phase(p, 42, 0)
    m = s.copy()
    out = {"lattice": n, "steps": steps, "first_above_1e-12": None, "first_above_1e-6": None, "at": {},
           "trig": {"link_args": 0, "link_sin_diff": 0, "link_cos_diff": 0, "plaq_args": 0, "plaq_sin_diff": 0,
                    "sampled_steps": []}}
    for t in range(steps + 1):
        if t % sample_every == 0:
            lk, pq = trig_args(p, s, t)
            tr = out["trig"]
            tr["link_args"] += lk.size
            tr["link_sin_diff"] += differing(lk, 0, 9)
            tr["link_cos_diff"] += differing(lk, 1, 10)
            tr["plaq_args"] += pq.size
            tr["plaq_sin_diff"] += differing(pq, 0, 9)
            tr["sampled_steps"].append(t)
        if t in report:
            ds, dl = orc.diag(p, s, t, SHARED), orc.diag(p, m, t, LIBM)
            xs, xl = xi(p, s, t, SHARED), xi(p, m, t, LIBM)
            out["at"][str(t)] = {
                "energy_shared": ds["energy"], "energy_libm": dl["energy"],
                "energy_rel": abs(ds["energy"] - dl["energy"]) / abs(dl["energy"]),
                "gauss_shared": ds["gauss"], "gauss_libm": dl["gauss"],
                "gauss_rel": abs(ds["gauss"] - dl["gauss"]) / abs(dl["gauss"]),
                "xi_shared": xs, "xi_libm": xl, "xi_rel": abs(xs - xl) / xl,
                "field_rel": field_rel(p, s, m, t)}
        if t == steps:
            break
        orc.step(p, s, t, 1, SHARED, threads)
        orc.step(p, m, t, 1, LIBM, threads)
        d = field_rel(p, s, m, t + 1)
        if out["first_above_1e-12"] is None and d > 1e-12:
            out["first_above_1e-12"] = t + 1
        if out["first_above_1e-6"] is None and d > 1e-6:
            out["first_above_1e-6"] = t + 1
            if stop_at == "1e-6":
                break
    tr = out["trig"]
    tr["link_sin_frac"] = tr["link_sin_diff"] / max(tr["link_args"], 1)
    tr["link_cos_frac"] = tr["link_cos_diff"] / max(tr["link_args"], 1)
    tr["plaq_sin_frac"] = tr["plaq_sin_diff"] / max(tr["plaq_args"], 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--threads", type=int, default=8)
    a = ap.parse_args()
    print(json.dumps(run(a.n, a.steps, a.threads), indent=1))


if __name__ == "__main__":
    main()
