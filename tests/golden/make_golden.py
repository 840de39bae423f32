#!/usr/bin/env python3
"""Write the golden fixtures in tests/golden/ from the reference programs.

Runs in the build container only (the reference is not available elsewhere).
Each program is compiled from sjcott/Axions by oracle/ref/build_ref.py into
oracle/_ref/<name>/, fed the seeded random-phase initial condition of
mpi_evolution.cpp:287-312 (written at 17 significant digits as Data/ic.txt,
the 5-column stationary format of evolution.cpp:102-117, or the 7-column
Data/moore_ic.txt of mooreEvol.cpp:73-91) and run single-threaded so its
OpenMP energy/Gauss sums have a fixed order.  The fixture records, per
program: energy and Gauss-deviation per TimeStep (valsPerLoop), the newest
field level after the run (finalField.txt) as a SHA-256 plus sampled sites
, and every string-crossing frame
(GifData) as count + SHA-256 (+ full list when small).

Usage: python tests/golden/make_golden.py [name ...]
"""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from axions_amd.abi import (ah_params, global_params, moore_params, AX_BC_ABSORBING, AX_BC_FIXED,  # noqa: E402
                             AX_BC_NEUMANN, AX_BC_PERIODIC)
from oracle import oracle_py as orc  # noqa: E402
sys.path.insert(0, os.path.join(REPO, "oracle", "ref"))
import build_ref  # noqa: E402

SEED = 42

# name -> (params builder, nt, saveFreq, detection start step (damped_nt), trig)
SPECS = {
    "evo_ah_g0": (lambda: ah_params(16, dx=0.5, dt=0.05, lam=1.0, eta=1.0, g=0.0, alpha=3.0,
                                    damped_nt=20, dampFac=0.5, theta_damp_form=1), 60, 5, 20, "libm"),
    "evo_ah_g1_shared": (lambda: ah_params(16, dx=0.5, dt=0.1, lam=2.0, eta=1.0, g=1.0, alpha=3.0,
                                           damped_nt=20, dampFac=0.5, theta_damp_form=1), 60, 5, 20, "shared"),
    "evo_ah_g1_libm": (lambda: ah_params(16, dx=0.5, dt=0.1, lam=2.0, eta=1.0, g=1.0, alpha=3.0,
                                         damped_nt=20, dampFac=0.5, theta_damp_form=1), 60, 5, 20, "libm"),
    "evo_ah_aniso": (lambda: ah_params(12, ny=10, nz=14, dx=0.7, dy=0.6, dz=0.55, dt=0.13, lam=1.5,
                                       eta=0.9, g=0.8, alpha=2.0, damped_nt=10, dampFac=0.3,
                                       theta_damp_form=1), 40, 5, 10, "shared"),
    "evo_c1_shared": (lambda: ah_params(64, dx=0.5, dt=0.1, lam=2.0, eta=1.0, g=1.0, alpha=0.0,
                                        damped_nt=0, dampFac=0.5, theta_damp_form=1), 200, 10, 0, "shared"),
    "evo_c2_1000_shared": (lambda: ah_params(64, dx=0.5, dt=0.1, lam=2.0, eta=1.0, g=1.0, alpha=0.0,
                                             damped_nt=0, dampFac=0.5, theta_damp_form=1), 1000, 20, 0, "shared"),
    "evo_bc_neumann": (lambda: ah_params(14, ny=12, nz=10, dx=0.5, dt=0.1, lam=2.0, g=1.0, alpha=2.0, damped_nt=10,
                                         dampFac=0.5, theta_damp_form=1, boundary=AX_BC_NEUMANN,
                                         boundary_z=AX_BC_PERIODIC), 40, 5, 10, "shared"),
    "evo_bc_neumann_z": (lambda: ah_params(13, ny=11, nz=9, dx=0.6, dy=0.5, dz=0.55, dt=0.1, lam=1.5, g=0.8,
                                           alpha=2.0, damped_nt=8, dampFac=0.4, theta_damp_form=1,
                                           boundary=AX_BC_NEUMANN, boundary_z=AX_BC_NEUMANN), 40, 5, 8, "shared"),
    "evo_bc_absorbing": (lambda: ah_params(14, ny=13, nz=10, dx=0.5, dt=0.1, lam=2.0, g=1.0, alpha=2.0,
                                           damped_nt=10, dampFac=0.5, theta_damp_form=1, boundary=AX_BC_ABSORBING,
                                           boundary_z=AX_BC_PERIODIC), 40, 5, 10, "shared"),
    "evo_bc_absorbing_zfixed": (lambda: ah_params(12, ny=14, nz=11, dx=0.6, dy=0.55, dz=0.5, dt=0.1, lam=1.5,
                                                  eta=0.9, g=0.8, alpha=2.0, damped_nt=8, dampFac=0.4,
                                                  theta_damp_form=1, boundary=AX_BC_ABSORBING,
                                                  boundary_z=AX_BC_FIXED), 40, 5, 8, "shared"),
    "evo_bc_fixed_zper": (lambda: ah_params(13, ny=12, nz=10, dx=0.5, dt=0.1, lam=2.0, g=1.0, alpha=2.0,
                                            damped_nt=10, dampFac=0.5, theta_damp_form=1, boundary=AX_BC_FIXED,
                                            boundary_z=AX_BC_PERIODIC), 40, 5, 10, "shared"),
    "evo_bc_fixed_zneu": (lambda: ah_params(12, ny=11, nz=12, dx=0.5, dt=0.1, lam=2.0, g=1.0, alpha=2.0,
                                            damped_nt=10, dampFac=0.5, theta_damp_form=1, boundary=AX_BC_FIXED,
                                            boundary_z=AX_BC_NEUMANN), 40, 5, 10, "shared"),
    "moore_neumann": (lambda: moore_params(12, ny=11, nz=10, boundary=AX_BC_NEUMANN), 30, None, None, "shared"),
    "moore_shared": (lambda: moore_params(12, boundary=AX_BC_FIXED), 30, None, None, "shared"),
    "moore_libm": (lambda: moore_params(12, boundary=AX_BC_FIXED), 30, None, None, "libm"),
    "global_neumann": (lambda: global_params(12, ny=10, nz=14), 40, None, None, "none"),
    "global_aniso": (lambda: global_params(9, ny=11, nz=13, dx=0.3, dy=0.35, dz=0.28, dt=0.07), 30, None, None,
                     "none"),
}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def params_dict(p):
    return {f: getattr(p, f) for f, _ in p._fields_}


def write_ic(p, st, path):
    cur, _ = orc.levels(p, st, 0)
    cols = cur.reshape(cur.shape[0], -1).T
    np.savetxt(path, cols, fmt="%.17g")


def run_global(name, p, nt):
    """leapfrog.cpp: reads ./ic.txt (Re phi, Im phi per site), writes one
    energy per TimeStep to ./valsPerLoop.txt and ./finalField.txt."""
    dest = os.path.join(REPO, "oracle", "_ref", name)
    if not os.path.exists(os.path.join(dest, "prog")):
        build_ref.build(name, "/root/reference")
    st = orc.init_random_phase(p, SEED, 0)      # phi = (cos, sin) of the seeded phases
    V = p.nx * p.ny * p.nz
    np.savetxt(os.path.join(dest, "ic.txt"), st[:2 * V].reshape(2, V).T, fmt="%.17g")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.run(["./prog"], cwd=dest, text=True, env=env, check=True, stdout=subprocess.DEVNULL)
    energy = [float(line) for line in open(os.path.join(dest, "valsPerLoop.txt"))]
    assert len(energy) == nt, (name, len(energy))
    fin = np.loadtxt(os.path.join(dest, "finalField.txt"), dtype=np.float64).T.copy()
    rng = np.random.default_rng(7)
    idx = sorted(int(i) for i in rng.choice(V, size=min(64, V), replace=False))
    fx = {
        "name": name, "program": build_ref.CONFIGS[name][0], "trig": "none", "seed": SEED,
        "nt": nt, "params": params_dict(p), "energy": energy,
        "final_sha256": sha(fin), "final_sample_idx": idx, "final_sample": fin[:, idx].T.tolist(),
        "frames": [],
    }
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(fx, f)
    for f in ("ic.txt", "valsPerLoop.txt", "finalField.txt"):
        os.remove(os.path.join(dest, f))
    print(name, "energy[0..2]", energy[:3], "final", fx["final_sha256"][:12])


ICW = {   # name -> (nx, ny, nz, dx, dy, dz, ic seed, split)
    "icw_len_split": (15, 14, 16, 0.35, 0.35, 0.35, 3, True),
    "icw_len_total": (12, 13, 11, 0.5, 0.45, 0.4, 5, False),
}


def run_icw(name):
    """ic_w_evolution.cpp with ic_func.cpp's seeded "random" IC (minstd_rand,
    uniform in (-1, 1), Re then Im per site in x-major order, links zero):
    the string length it prints for TimeStep 0."""
    nx, ny, nz, dx, dy, dz, seed, split = ICW[name]
    dest = os.path.join(REPO, "oracle", "_ref", name)
    if not os.path.exists(os.path.join(dest, "prog")):
        build_ref.build(name, "/root/reference")
    for d in ("Data", "GifData"):
        os.makedirs(os.path.join(dest, d), exist_ok=True)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.run(["./prog"], cwd=dest, text=True, env=env, check=True, stdout=subprocess.DEVNULL)
    f = open(os.path.join(dest, "Data", "valsPerLoop_test.txt")).readline().split()
    length = [float(v) for v in f[2:]]
    assert len(length) == (2 if split else 1)
    fx = {"name": name, "program": "ic_w_evolution.cpp + ic_func.cpp", "lattice": [nx, ny, nz],
          "spacing": [dx, dy, dz], "ic": {"engine": "minstd_rand", "seed": seed, "uniform": [-1.0, 1.0]},
          "split": split, "length": length}
    with open(os.path.join(HERE, f"{name}.json"), "w") as fo:
        json.dump(fx, fo)
    for d in ("Data", "GifData"):
        for fn in os.listdir(os.path.join(dest, d)):
            os.remove(os.path.join(dest, d, fn))
    for fn in ("test1.txt", "test2.txt"):
        if os.path.exists(os.path.join(dest, fn)):
            os.remove(os.path.join(dest, fn))
    print(name, "length", length)


# straight-string pipeline: name -> SOR constants (singleSemiAnalytic.cpp:19-32)
SOR = {
    "sor_ref": dict(nx=30001, h=0.01, w=1.5, tol=1e-3, n=1.0, lam=1.0, eta=1.0, g=0.0, max_iter=2000000),
    "sor_gauged": dict(nx=1501, h=0.02, w=1.6, tol=1e-4, n=1.0, lam=2.0, eta=1.0, g=1.0, max_iter=2000000),
    "sor_cap": dict(nx=801, h=0.03, w=1.5, tol=1e-3, n=2.0, lam=1.0, eta=1.0, g=0.5, max_iter=300),
}
# name -> straight_ic.cpp:20-31 constants and the SOR run whose output it reads
STRAIGHT = {
    "straight_ref": dict(lattice=[101, 101, 51], spacing=[0.5, 0.5, 0.5], n=1, g=0.5, SORnx=20001, SORa=0.01,
                         profile="sor_ref"),
    "straight_gauged": dict(lattice=[40, 33, 3], spacing=[0.45, 0.55, 0.5], n=1, g=1.0, SORnx=403, SORa=0.02,
                            profile="sor_gauged"),
}


def sampled(a, k, seed=7):
    """64 sampled rows of a [rows, cols] array: indices and values."""
    rng = np.random.default_rng(seed)
    idx = sorted(int(i) for i in rng.choice(a.shape[0], size=min(k, a.shape[0]), replace=False))
    return idx, a[idx].tolist()


def run_sor(name):
    """singleSemiAnalytic.cpp, one thread, ic_type 0: the profile it writes
    to Data/SOR_Fields.txt and the sweep count it prints."""
    dest = os.path.join(REPO, "oracle", "_ref", name)
    if not os.path.exists(os.path.join(dest, "prog")):
        build_ref.build(name, "/root/reference")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = subprocess.run(["./prog"], cwd=dest, text=True, env=env, check=True, capture_output=True)
    iters = None
    for part in r.stdout.replace("\r", "\n").splitlines():
        if part.startswith("Number of iterations:"):
            iters = int(part.split(":")[1])
    prof = np.loadtxt(os.path.join(dest, "Data", "SOR_Fields.txt"), dtype=np.float64)
    idx, rows = sampled(prof, 64)
    fx = {"name": name, "program": "singleSemiAnalytic.cpp", "consts": SOR[name], "iterations": iters,
          "profile_sha256": sha(prof), "sample_idx": idx, "sample": rows,
          "tail": prof[-3:].tolist(), "head": prof[:3].tolist()}
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(fx, f)
    print(name, "iterations", iters, "sha", fx["profile_sha256"][:12])


def run_straight(name):
    """straight_ic.cpp on the profile of STRAIGHT[name]["profile"] (read at
    17 digits): its five ic.txt columns and the sites it reports beyond the
    profile."""
    spec = STRAIGHT[name]
    dest = os.path.join(REPO, "oracle", "_ref", name)
    if not os.path.exists(os.path.join(dest, "prog")):
        build_ref.build(name, "/root/reference")
    src = os.path.join(REPO, "oracle", "_ref", spec["profile"], "Data", "SOR_Fields.txt")
    if not os.path.exists(src):
        run_sor(spec["profile"])
    with open(src) as fi, open(os.path.join(dest, "SOR_Fields.txt"), "w") as fo:
        fo.write(fi.read())
    r = subprocess.run(["./prog"], cwd=dest, text=True, check=True, capture_output=True)
    off = sum(1 for line in r.stdout.splitlines() if line.startswith("Off straight string solution grid"))
    ic = np.loadtxt(os.path.join(dest, "ic.txt"), dtype=np.float64)
    cols = ic.T.copy()                                   # [5, nx*ny*nz]
    idx, rows = sampled(ic, 64)
    fx = {"name": name, "program": "straight_ic.cpp", **spec, "off_grid": off,
          "ic_sha256": sha(cols), "sample_idx": idx, "sample": rows}
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(fx, f)
    for fn in ("ic.txt", "SOR_Fields.txt"):
        os.remove(os.path.join(dest, fn))
    print(name, "off-grid", off, "sha", fx["ic_sha256"][:12])


def run(name):
    if name in ICW:
        return run_icw(name)
    if name in SOR:
        return run_sor(name)
    if name in STRAIGHT:
        return run_straight(name)
    builder, nt, save, det0, trig = SPECS[name]
    p = builder()
    if p.model == 2:
        return run_global(name, p, nt)
    dest = os.path.join(REPO, "oracle", "_ref", name)
    if not os.path.exists(os.path.join(dest, "prog")):
        build_ref.build(name, "/root/reference")
    moore = p.model == 1
    st = orc.init_random_phase(p, SEED, 0)
    write_ic(p, st, os.path.join(dest, "Data", "moore_ic.txt" if moore else "ic.txt"))
    for d in ("Data", "GifData"):
        for f in os.listdir(os.path.join(dest, d)):
            if f.startswith(("valsPerLoop", "gifStringPos", "finalField")):
                os.remove(os.path.join(dest, d, f))
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.run(["./prog"], cwd=dest, input="golden\n", text=True, env=env, check=True,
                   stdout=subprocess.DEVNULL)
    vals = os.path.join(dest, "Data", "valsPerLoop.txt" if moore else "valsPerLoop_golden.txt")
    energy, gauss = [], []
    for line in open(vals):
        f = line.split()
        energy.append(float(f[0]))
        gauss.append(float(f[1]))
    assert len(energy) == nt, (name, len(energy))
    fin = np.loadtxt(os.path.join(dest, "Data", "finalField.txt"), dtype=np.float64).T.copy()
    fin = fin.reshape(fin.shape[0], p.nx, p.ny, p.nz)
    rng = np.random.default_rng(7)
    V = p.nx * p.ny * p.nz
    idx = sorted(int(i) for i in rng.choice(V, size=min(64, V), replace=False))
    flat = fin.reshape(fin.shape[0], -1)
    fx = {
        "name": name, "program": build_ref.CONFIGS[name][0], "trig": trig, "seed": SEED,
        "nt": nt, "params": params_dict(p),
        "energy": energy, "gauss": gauss,
        "final_sha256": sha(fin), "final_sample_idx": idx,
        "final_sample": flat[:, idx].T.tolist(),
    }
    frames = []
    if save is not None:
        frame = 0
        for n in range(nt):
            if n % save == 0 and n >= det0:
                path = os.path.join(dest, "GifData", f"gifStringPosData_{frame}.txt")
                pts = np.loadtxt(path, dtype=np.float64, ndmin=2) if os.path.getsize(path) else np.zeros((0, 3))
                ent = {"step": n, "count": int(pts.shape[0]), "sha256": sha(pts)}
                if pts.shape[0] <= 400:
                    ent["points"] = pts.tolist()
                frames.append(ent)
                frame += 1
    fx["frames"] = frames
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(fx, f)
    for d in ("Data", "GifData"):   # run outputs are not kept beside the binary
        for f in os.listdir(os.path.join(dest, d)):
            os.remove(os.path.join(dest, d, f))
    print(name, "energy[0..2]", energy[:3], "frames", len(frames), "final", fx["final_sha256"][:12])


def main():
    for n in (sys.argv[1:] or list(SPECS) + list(ICW) + list(SOR) + list(STRAIGHT)):
        run(n)


if __name__ == "__main__":
    main()
