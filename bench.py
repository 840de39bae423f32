#!/usr/bin/env python3
"""bench.py - throughput of the lattice-evolution hot path on one MI355X.

Default workload (the reference-scale single-GPU configuration): plain
Abelian-Higgs leapfrog, 512^3 periodic lattice, FP64, dx=0.5, dt=0.1,
lambda=2, g=eta=1, seeded random-phase IC (mt19937 seed 42), energy +
Gauss-law + winding-number string count every 20 steps.  A "step" is one
TimeStep of the fused in-place update over the whole lattice.

Prints ONE JSON line:
  value      whole-job site-updates/s over the timed steps (state resident in
             HBM; the diagnostic passes every --diag-every steps are inside
             the timed region)
  roofline   the step kernel: algorithmic bytes per site-update (120 B FP64,
             60 B FP32: 10 field values read, 5 written) x sites / the
             kernel's average duration from HIP events on its stream,
             against 8 TB/s
  cpu_baseline  the oracle (C++/OpenMP restatement of the reference step,
             glibc sin/cos like the reference) timed on this host over a
             bounded sample of the same lattice
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
METRIC = "site-updates/s at 512³ FP64/FP32, % HBM roofline; ensemble p50/p99 latency"


def measured_traffic(n, precision):
    """HBM bytes per step launch from the committed rocprofv3 PMC summary of
    this exact workload (FETCH_SIZE/WRITE_SIZE scaled by a copy of known
    size), or None when no summary covers it."""
    path = os.path.join(REPO, "profiles", f"ah_step_tiled_{n}_fp{precision}.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def physical_cores():
    try:
        import psutil
        n = psutil.cpu_count(logical=False)
        if n:
            return n
    except Exception:
        pass
    return os.cpu_count() or 1


def cpu_baseline(n, steps, threads, log):
    """Oracle step (reference arithmetic, libm trig) at n^3 for `steps` steps."""
    from axions_amd.abi import ah_params
    from oracle import oracle_py as orc
    p = ah_params(n)
    log(f"cpu baseline: oracle {n}^3, {steps} steps, {threads} threads")
    st = orc.init_random_phase(p, 42, 0)
    orc.step(p, st, 0, 1, 1, threads)  # warm caches / page in
    t0 = time.perf_counter()
    orc.step(p, st, 1, steps, 1, threads)
    dt = time.perf_counter() - t0
    del st
    return {"value": n ** 3 * steps / dt, "unit": "site-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C++/OpenMP restatement of mpi_evolution.cpp:381-501, glibc sin/cos) "
                      f"AH {n}^3 FP64, {steps} steps after 1 untimed, {dt:.2f} s"}


def copy_calibration(nbytes, reps):
    """Device-to-device copies of a known size (HIP runtime copy kernel), run
    under rocprofv3 --pmc to turn FETCH_SIZE/WRITE_SIZE into bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}))


def main():
    if "--copy-calibration" in sys.argv:
        copy_calibration(1 << 30, 3)
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--precision", type=int, default=64, choices=(64, 32))
    ap.add_argument("--diag-every", type=int, default=20)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--cpu-baseline", choices=("on", "off"), default="on")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--cpu-n", type=int, default=0, help="lattice for the CPU sample (default: --n)")
    ap.add_argument("--quiet", action="store_true")
    a = ap.parse_args()

    def log(*x):
        if not a.quiet:
            print(*x, file=sys.stderr, flush=True)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only: this build keeps one whole lattice per GPU "
              f"(requested --gpus {a.gpus}, WORLD_SIZE {world}); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import axions_amd as ax
    p = ax.ah_params(a.n, precision=a.precision)
    log(f"lattice {a.n}^3 FP{a.precision}: IC seed {a.seed}")
    lat = ax.Lattice(p)
    lat.init_random_phase(a.seed)

    def run(nsteps):
        done = 0
        while done < nsteps:
            k = min(a.diag_every - (lat.timestep % a.diag_every), nsteps - done)
            if lat.timestep % a.diag_every == 0:
                d = lat.diagnostics(winding=True)
                if not d["finite"]:
                    raise RuntimeError(f"non-finite state at step {d['timestep']}")
            lat.step(k)
            done += k
        lat.sync()

    log(f"warmup {a.warmup} steps")
    run(a.warmup)
    lat.kernel_timer(True)
    lat.kernel_time()
    log(f"timed {a.steps} steps (diagnostics every {a.diag_every})")
    t0 = time.perf_counter()
    run(a.steps)
    t1 = time.perf_counter()
    kms, ksteps = lat.kernel_time()
    lat.kernel_timer(False)
    dlast = lat.diagnostics(winding=True)
    elapsed = t1 - t0
    sites = a.n ** 3
    value = sites * a.steps / elapsed
    bytes_per_site = 120 if a.precision == 64 else 60
    avg_kernel_s = kms / 1e3 / max(ksteps, 1)
    achieved = bytes_per_site * sites / avg_kernel_s / 1e9
    lat.close()

    cpu = None
    if a.cpu_baseline == "on":
        cpu = cpu_baseline(a.cpu_n or a.n, a.cpu_steps, min(16, physical_cores()), log)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == 64 else "f32",
        "data": f"synthetic: seeded random-phase IC (std::mt19937 seed {a.seed}), pi=0, A=E=0",
        "config": {"workload": f"Abelian-Higgs leapfrog {a.n}^3 periodic FP{a.precision}, dx=0.5 dt=0.1 "
                               f"lambda=2 g=eta=1, energy+Gauss+winding every {a.diag_every} steps",
                   "n": a.n, "precision": a.precision, "diag_every": a.diag_every},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(a.n, a.precision),
                     "kernel_ms": avg_kernel_s * 1e3, "bytes_per_site": bytes_per_site},
        "cpu_baseline": cpu,
        "final_diag": {"timestep": dlast["timestep"], "energy": dlast["energy"],
                       "gauss_mean": dlast["gauss_mean"], "wound_plaquettes": dlast["wound_plaquettes"]},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
