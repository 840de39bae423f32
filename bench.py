#!/usr/bin/env python3
"""bench.py - throughput of the lattice-evolution hot path on one MI355X.

Headline workload (`value`): plain Abelian-Higgs leapfrog, 512^3 periodic
lattice, FP64, dx=0.5, dt=0.1, lambda=2, g=eta=1, seeded random-phase IC
(mt19937 seed 42), 1000 steps with energy + Gauss-law + winding-number string
count every 20 steps.  A "step" is one TimeStep of the fused in-place update
over the whole lattice.

Prints ONE JSON line:
  value         whole-job site-updates/s over the timed steps (state resident
                in HBM; the diagnostic passes every --diag-every steps are
                inside the timed region)
  roofline      the step kernel: algorithmic bytes per site-update (120 B
                FP64, 60 B FP32: 10 field values read, 5 written) x sites / the
                kernel's mean duration from HIP events on its stream, against
                8 TB/s; traffic = HBM bytes per launch measured by rocprofv3
                (committed summary under profiles/), null if none
  cpu_baseline  the reference program as shipped (evolution.cpp, glibc sin/cos,
                OpenMP; built from its sources into oracle/_ref/) timed on this
                host's CPU share over 20 TimeSteps of a 128^3 lattice of the
                same workload, with the oracle (the C++/OpenMP restatement of
                the fused mpi_evolution.cpp step) on the same sample beside it
  fp32          the same run in FP32 storage/arithmetic (FP64 reductions), its
                throughput and roofline, and its deviation from the FP64 run in
                energy, Gauss residual and xi(t) over TimeSteps 0..1000
                (--series-steps; both runs continue untimed past the timed
                window)
  moore         the Moore model (two scalars, charges 2 and 1, one U(1) field)
                at 256^3 FP64, mooreEvol.cpp couplings, 800 steps with winding
                counts every 20: site-updates/s and the step kernel's roofline
                (168 B per site-update)
  ensemble      ensemble serving: Poisson arrivals of realizations with N drawn
                from {64,96,128,192,256}, steps = N*dx/(2*dt), served by the
                FP32 pool (continuous batching, persistent batched kernel,
                hipGraph chunks); realizations/hour, site-updates/s, p50/p99
                latency; the same stream run one by one through the FP64
                single-lattice path for comparison (realizations/hour and xi)
                and a second load point offered at 0.7x the pool's capacity;
                the pool's step kernel against the same roofline
  boundaries    evolution.cpp's own schedule (absorbing x/y faces, PRS
                friction, damping window) at 512^3
  sor_profile   the straight-string SOR profile on the GPU and the oracle's
                one-thread loop beside it

Two auxiliary modes print their own line instead: --copy-calibration (HIP
runtime device-to-device copies of 1 GiB, run under rocprofv3 --pmc to turn
FETCH_SIZE / WRITE_SIZE into bytes) and --stream-probe (what plain streaming
kernels -- torch's sum, copy and add over 1 GiB FP64 tensors -- reach on this
GPU's HBM).
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
ENSEMBLE_NS = (64, 96, 128, 192, 256)


def measured_traffic(n, precision, model="ah"):
    """HBM bytes per step launch from the committed rocprofv3 PMC summary of
    this exact workload (FETCH_SIZE/WRITE_SIZE scaled by a copy of known
    size), or None when no summary covers it."""
    path = os.path.join(REPO, "profiles", f"{model}_step_tiled_{n}_fp{precision}.md")
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


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def logical_cpus():
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def cpu_threads():
    """Threads for the CPU baseline: OMP_NUM_THREADS when the host sets it
    (a share of a shared host), else every physical core."""
    v = os.environ.get("OMP_NUM_THREADS", "")
    return int(v) if v.isdigit() and int(v) > 0 else physical_cores()


REF_CPU_N = 128          # lattice of the reference CPU sample
REF_CPU_STEPS = (2, 22)  # its two builds' TimeStep counts


def reference_cpu_baseline(threads, log):
    """The reference program as shipped (evolution.cpp, glibc sin/cos,
    OpenMP, compiled from its own sources into oracle/_ref/ by
    oracle/ref/build_ref.py) on the seeded random-phase 128^3 lattice with
    the benchmark's couplings.  Its two builds differ only in nt; the
    difference of their wall times is 20 TimeSteps of its loop (the update,
    its energy/Gauss sums and the valsPerLoop line) without the ic.txt read.
    None when the builds are absent."""
    import subprocess
    import numpy as np
    from axions_amd.abi import ah_params
    from oracle import oracle_py as orc
    dirs = [os.path.join(REPO, "oracle", "_ref", f"cpu_bench_nt{k}") for k in REF_CPU_STEPS]
    if not all(os.path.exists(os.path.join(d, "prog")) for d in dirs):
        log("cpu baseline: reference builds absent")
        return None
    p = ah_params(REF_CPU_N)
    st = orc.init_random_phase(p, 42, 0)
    cur, _ = orc.levels(p, st, 0)
    log(f"cpu baseline: writing the {REF_CPU_N}^3 Data/ic.txt")
    for d in dirs:
        os.makedirs(os.path.join(d, "Data"), exist_ok=True)
    np.savetxt(os.path.join(dirs[0], "Data", "ic.txt"), cur.reshape(5, -1).T, fmt="%.17g")
    ic1 = os.path.join(dirs[1], "Data", "ic.txt")
    if os.path.exists(ic1):
        os.remove(ic1)
    os.link(os.path.join(dirs[0], "Data", "ic.txt"), ic1)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    wall = []
    for d, k in zip(dirs, REF_CPU_STEPS):
        log(f"cpu baseline: reference evolution.cpp {REF_CPU_N}^3 nt={k}, {threads} threads")
        t0 = time.perf_counter()
        subprocess.run(["./prog"], cwd=d, input="bench\n", text=True, env=env, check=True,
                       stdout=subprocess.DEVNULL)
        wall.append(time.perf_counter() - t0)
    for d in dirs:
        for f in os.listdir(os.path.join(d, "Data")):
            os.remove(os.path.join(d, "Data", f))
    nsteps = REF_CPU_STEPS[1] - REF_CPU_STEPS[0]
    dt = wall[1] - wall[0]
    return {"value": REF_CPU_N ** 3 * nsteps / dt, "unit": "site-updates/s", "cores": threads,
            "kind": "reference",
            "sample": f"reference evolution.cpp as shipped (glibc sin/cos, g++ -O3 -fopenmp), "
                      f"{REF_CPU_N}^3 periodic FP64, seeded random-phase IC, benchmark couplings; "
                      f"{nsteps} TimeSteps = run(nt={REF_CPU_STEPS[1]}) {wall[1]:.2f} s - "
                      f"run(nt={REF_CPU_STEPS[0]}) {wall[0]:.2f} s = {dt:.2f} s"}


def port_cpu_baseline(n, steps, threads, log):
    """The oracle's step (the reference's arithmetic in its fused
    mpi_evolution.cpp form, glibc sin/cos, OpenMP) at n^3 for `steps` steps."""
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


def cpu_baseline(threads, log):
    """The reference program timed on this host (the reported baseline),
    with the oracle on the same lattice beside it."""
    host = {"cpu": cpu_model(), "physical_cores": physical_cores(), "logical_cpus": logical_cpus(),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}
    port = port_cpu_baseline(REF_CPU_N, REF_CPU_STEPS[1] - REF_CPU_STEPS[0], threads, log)
    ref = reference_cpu_baseline(threads, log)
    out = dict(ref if ref is not None else port)
    out.update(host)
    if ref is not None:
        out["port"] = port
        out["port_over_reference"] = port["value"] / ref["value"]
    return out


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


def run_lattice(ax, n, precision, steps, warmup, diag_every, seed, log, moore=False, moore_dt=0.3, series_to=0,
                fused=True):
    """The headline loop on one lattice: warmup, then `steps` timed steps with
    diagnostics every diag_every steps, then (untimed) on to TimeStep
    series_to if that is further; returns throughput, the step kernel's
    mean time and the diagnostic series."""
    p = ax.moore_params(n, dt=moore_dt, precision=precision) if moore else ax.ah_params(n, precision=precision)
    log(f"lattice {n}^3 FP{precision}{' Moore' if moore else ''}: IC seed {seed}")
    series = []
    with ax.Lattice(p) as lat:
        lat.init_random_phase(seed)

        def diag():
            d = lat.diagnostics(winding=True)
            if not d["finite"]:
                raise RuntimeError(f"non-finite state at step {d['timestep']}")
            series.append((d["timestep"], d["energy"], d["gauss_mean"], d["wound_plaquettes"], d["xi"], d["finite"]))

        def run(nsteps):
            if fused:
                # the diagnostics of every diag_every-th TimeStep taken by its
                # step (ax_step_diag; the separate pass where that does not apply)
                for d in lat.step_diag(nsteps, diag_every, winding=True):
                    if not d["finite"]:
                        raise RuntimeError(f"non-finite state at step {d['timestep']}")
                    series.append((d["timestep"], d["energy"], d["gauss_mean"], d["wound_plaquettes"], d["xi"],
                                   d["finite"]))
                lat.sync()   # ax_step_diag waits at its end; the timed region ends with the device idle
                return
            done = 0
            while done < nsteps:
                k = min(diag_every - (lat.timestep % diag_every), nsteps - done)
                if lat.timestep % diag_every == 0:
                    diag()
                lat.step(k)
                done += k
            lat.sync()

        log(f"warmup {warmup} steps")
        run(warmup)
        lat.kernel_timer(True)
        lat.kernel_time()
        log(f"timed {steps} steps (diagnostics every {diag_every})")
        t0 = time.perf_counter()
        run(steps)
        t1 = time.perf_counter()
        kms, ksteps = lat.kernel_time()
        lat.kernel_timer(False)
        if lat.timestep < series_to:
            log(f"untimed steps {lat.timestep} -> {series_to} for the diagnostic series")
            run(series_to - lat.timestep)
        diag()
    elapsed = t1 - t0
    sites = n ** 3
    bps = (168 if moore else 120) if precision == 64 else (84 if moore else 60)
    avg_kernel_s = kms / 1e3 / max(ksteps, 1)
    achieved = bps * sites / avg_kernel_s / 1e9
    return {
        "value": sites * steps / elapsed,
        "ms_per_step": elapsed / steps * 1e3,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(n, precision, "moore" if moore else "ah"),
                     "kernel_ms": avg_kernel_s * 1e3, "bytes_per_site": bps},
        "series": series,
    }


def boundary_leg(ax, n, steps, seed, log):
    """evolution.cpp's own schedule at scale: global string (g = 0), lambda = 1,
    dt = 0.05, PRS friction alpha = 3 with damping 0.5 for the first 200
    steps, absorbing x/y faces, periodic z (evolution.cpp:23-53)."""
    import time as _t
    from axions_amd.abi import AX_BC_ABSORBING, AX_BC_PERIODIC
    p = ax.ah_params(n, dx=0.5, dt=0.05, lam=1.0, eta=1.0, g=0.0, alpha=3.0, scaling=1.0, damped_nt=200,
                     dampFac=0.5, theta_damp_form=1, boundary=AX_BC_ABSORBING, boundary_z=AX_BC_PERIODIC)
    log(f"boundary leg: {n}^3 absorbing x/y, PRS alpha=3, damping to step 200, {steps} steps")
    with ax.Lattice(p) as lat:
        lat.init_random_phase(seed)
        lat.step(2)
        lat.sync()
        lat.kernel_timer(True)
        lat.kernel_time()
        t0 = _t.perf_counter()
        done = 0
        while done < steps:
            if (lat.timestep % 20) == 0:
                d = lat.diagnostics(winding=False)
                if not d["finite"]:
                    raise RuntimeError(f"boundary leg: non-finite state at step {lat.timestep}")
            k = min(20 - lat.timestep % 20, steps - done)
            lat.step(k)
            done += k
        lat.sync()
        dt = _t.perf_counter() - t0
        kms, ksteps = lat.kernel_time()
        lat.kernel_timer(False)
        d = lat.diagnostics(winding=False)
    kernel_s = kms / 1e3 / max(ksteps, 1)
    achieved = 120 * n ** 3 / kernel_s / 1e9
    return {"workload": f"evolution.cpp schedule {n}^3 FP64: g=0, lambda=1, dt=0.05, alpha=3 (PRS), "
                        f"damped_nt=200, absorbing x/y, periodic z, {steps} steps from step 2, energy every 20",
            "value": n ** 3 * steps / dt, "unit": "site-updates/s", "ms_per_step": dt / steps * 1e3,
            "kernel_ms": kernel_s * 1e3,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "bytes_per_site": 120},
            "final_energy": d["energy"], "finite": bool(d["finite"])}


def sor_leg(ax, cpu_sweeps, log):
    """singleSemiAnalytic.cpp's straight-string profile at its own constants
    (30001 points, h = 0.01, w = 1.5, tol = 1e-3) on the GPU, and the oracle's
    one-thread sweeps on a bounded sample of the same problem."""
    import time as _t
    log("sor leg: singleSemiAnalytic.cpp constants")
    p = ax.sor_params()
    ax.sor_profile(ax.sor_params(nx=301, max_iter=10))           # load / warm up
    t0 = _t.perf_counter()
    prof, it, res = ax.sor_profile(p)
    dt = _t.perf_counter() - t0
    out = {"workload": "straight-string SOR profile, singleSemiAnalytic.cpp constants (nx=30001, h=0.01, w=1.5, "
                       "tol=1e-3, n=1, lambda=1, eta=1, g=0)",
           "sweeps": it, "last_residual": res, "gpu_seconds": dt,
           "value": p.nx * it / dt, "unit": "point-updates/s"}
    if cpu_sweeps:
        from oracle import oracle_py as orc
        t0 = _t.perf_counter()
        orc.sor_profile(p.nx, p.h, p.w, 0.0, p.n, p.lambda_, p.eta, p.g, cpu_sweeps)
        cdt = _t.perf_counter() - t0
        out["cpu_baseline"] = {"value": p.nx * cpu_sweeps / cdt, "unit": "point-updates/s", "cores": 1,
                               "kind": "port", "cpu": cpu_model(),
                               "sample": f"oracle one-thread SOR (singleSemiAnalytic.cpp:82-134), first "
                                         f"{cpu_sweeps} sweeps of the same problem, {cdt:.2f} s"}
    return out


def compare_series(a, b):
    """FP32 series b against FP64 series a at the same timesteps."""
    import numpy as np
    A = {r[0]: r for r in a}
    rows = [(A[r[0]], r) for r in b if r[0] in A]
    E0a, E0b = rows[0][0][1], rows[0][1][1]
    drift_a = np.array([(x[1] - E0a) / E0a for x, _ in rows])
    drift_b = np.array([(y[1] - E0b) / E0b for _, y in rows])
    eg = np.array([abs(y[1] - x[1]) / abs(x[1]) for x, y in rows])
    xa = np.array([x[4] for x, _ in rows])
    xb = np.array([y[4] for _, y in rows])
    xr = np.abs(xb - xa) / np.maximum(np.abs(xa), 1e-300)
    return {
        "timesteps_compared": len(rows),
        "energy_rel_diff_max": float(eg.max()),
        "energy_drift_fp64_final": float(drift_a[-1]), "energy_drift_fp32_final": float(drift_b[-1]),
        "gauss_fp64_final": rows[-1][0][2], "gauss_fp32_final": rows[-1][1][2],
        "xi_rel_diff_max": float(xr.max()), "xi_rel_diff_final": float(xr[-1]),
        "xi_fp64_final": float(xa[-1]), "xi_fp32_final": float(xb[-1]),
    }


def serve(ax, Ns, seeds, steps, arrivals, budget_gb, chunk, threads, log, tuned=None):
    """One pass of the request stream through the FP32 pool: submissions at
    the given arrival times, then wait for all.  Returns the load point's
    figures, the requests' xi series and the tuned tiles."""
    import numpy as np
    count = len(Ns)
    total_su = float(sum(N ** 3 * s for N, s in zip(Ns, steps)))
    with ax.Pool(ax.ah_params(8, dx=0.5, dt=0.1, precision=32), hbm_budget=int(budget_gb * (1 << 30)),
                 chunk_steps=chunk, autotune=1, ic_threads=threads) as pool:
        if tuned is None:
            tuned = {N: pool.tune(N)[0] for N in ENSEMBLE_NS}
            log(f"tuned tiles: {tuned}")
        else:
            for N in ENSEMBLE_NS:
                pool.tune(N)
        ids = []
        t0 = time.perf_counter()
        for i in range(count):
            d = t0 + arrivals[i] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            ids.append(pool.submit(seeds[i], Ns[i], steps[i]))
        while not pool.wait(timeout=20):
            s = pool.stats()
            log(f"  pool: {s['completed']}/{count} done, {s['resident']} resident, {s['queued']} queued")
        wall = time.perf_counter() - t0
        infos = [pool.info(i) for i in ids]
        xi32 = [pool.xi(i) for i in ids]
        st = pool.stats()
    lat = np.array([r["t_done"] - r["t_submit"] for r in infos])
    failed = sum(1 for r in infos if r["status"] != 2)
    # the batched step kernel against the HBM roofline: 60 algorithmic bytes
    # per site-update (FP32), over the GPU time of the chunks' step launches
    step_s = st["step_ms"] / 1e3
    achieved = 60.0 * st["site_updates"] / step_s / 1e9 if step_s > 0 else None
    out = {
        "arrival_rate_per_s": float(count / arrivals[-1]) if arrivals[-1] > 0 else None,
        "failed": failed,
        "wall_s": wall,
        "realizations_per_hour": count / wall * 3600,
        "site_updates_per_s": total_su / wall,
        "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
        "latency_mean_s": float(lat.mean()),
        "service_p50_s": float(np.percentile([r["t_done"] - r["t_admit"] for r in infos], 50)),
        "queue_wait_p50_s": float(np.percentile([r["t_admit"] - r["t_submit"] for r in infos], 50)),
        "max_resident": st["max_resident"], "peak_hbm_bytes": st["peak_in_use"], "graph_chunks": st["chunks"],
        "step_kernel_share_of_wall": step_s / wall,
        "pool_step_roofline": None if achieved is None else {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "bytes_per_site": 60,
            "step_kernel_s": step_s, "site_updates": st["site_updates"]},
    }
    log(f"pool @ {out['arrival_rate_per_s']:.1f}/s: {out['realizations_per_hour']:.0f} realizations/h, "
        f"{out['site_updates_per_s']:.3e} SU/s, p50 {out['latency_p50_s']:.2f} s, p99 {out['latency_p99_s']:.2f} s")
    return out, xi32, tuned


def ensemble(ax, count, rate, seed, budget_gb, chunk, fp64_compare, log, light_frac=0.7):
    import numpy as np
    rng = np.random.default_rng(seed)
    dx, dt = 0.5, 0.1
    Ns = [int(v) for v in rng.choice(ENSEMBLE_NS, size=count)]
    seeds = [int(v) for v in rng.integers(1, 2 ** 31, size=count)]
    steps = [int(round(N * dx / (2 * dt))) for N in Ns]
    gaps = rng.exponential(1.0, size=count)          # unit-rate Poisson process, scaled per load point
    unit = np.cumsum(gaps) - gaps[0]
    total_su = float(sum(N ** 3 * s for N, s in zip(Ns, steps)))
    threads = max(2, min(12, cpu_threads() - 2))
    log(f"ensemble: {count} requests, Poisson rate {rate}/s, {total_su:.3e} site-updates, {threads} IC threads")
    out = {"requests": count, "n_choices": list(ENSEMBLE_NS), "chunk_steps": chunk,
           "site_updates_total": total_su, "dtype": "f32", "reductions": "f64"}
    sat, xi32, tuned = serve(ax, Ns, seeds, steps, unit / rate, budget_gb, chunk, threads, log)
    out["tuned_xseg"] = {str(N): x for N, x in tuned.items()}
    out.update(sat)
    # the same stream below saturation: arrivals at light_frac of the
    # capacity just measured (completions per second of the saturated pass)
    capacity = count / sat["wall_s"]
    light_rate = light_frac * capacity
    log(f"ensemble at {light_frac:.0%} of capacity: {light_rate:.2f} arrivals/s")
    light, _, _ = serve(ax, Ns, seeds, steps, unit / light_rate, budget_gb, chunk, threads, log, tuned=tuned)
    light["offered_load_frac_of_capacity"] = light_frac
    out["below_saturation"] = light
    out["capacity_per_s"] = capacity
    if not fp64_compare:
        return out
    arrivals = unit / rate
    # the same stream, one realization at a time through the FP64 plain path
    log("fp64 plain path: the same requests one by one")
    xi64 = []
    service = []
    t0 = time.perf_counter()
    for i in range(count):
        a = time.perf_counter()
        ts_, x_ = [], []
        with ax.Lattice(ax.ah_params(Ns[i], dx=dx, dt=dt)) as L:
            L.init_random_phase(seeds[i])
            t = 0
            while True:
                d = L.diagnostics(winding=True)
                ts_.append(t)
                x_.append(d["xi"])
                if t == steps[i]:
                    break
                k = min(chunk, steps[i] - t)
                L.step(k)
                t += k
        xi64.append((np.array(ts_), np.array(x_)))
        service.append(time.perf_counter() - a)
        if (i + 1) % 32 == 0:
            log(f"  fp64 plain: {i + 1}/{count} done, {time.perf_counter() - t0:.1f} s")
    wall64 = time.perf_counter() - t0
    # FIFO single-server latencies for the same arrival times
    fin, lat64 = 0.0, []
    for i in range(count):
        fin = max(fin, arrivals[i]) + service[i]
        lat64.append(fin - arrivals[i])
    lat64 = np.array(lat64)
    rel = []
    by_n = {}
    for i in range(count):
        ts32, _, x32 = xi32[i]
        ts64, x64 = xi64[i]
        assert np.array_equal(ts32, ts64)
        rel.append(abs(x32[-1] - x64[-1]) / x64[-1])
        by_n.setdefault(Ns[i], []).append((x32, x64))
    mean_rel = {}
    for N, rows in by_n.items():
        m32 = np.mean([r[0] for r in rows], axis=0)
        m64 = np.mean([r[1] for r in rows], axis=0)
        mean_rel[str(N)] = float(np.max(np.abs(m32 - m64) / m64))
    all32 = np.mean([x[2][-1] for x in xi32])
    all64 = np.mean([x[1][-1] for x in xi64])
    out["fp64_plain"] = {
        "wall_s": wall64, "realizations_per_hour": count / wall64 * 3600,
        "site_updates_per_s": total_su / wall64,
        "latency_p50_s": float(np.percentile(lat64, 50)), "latency_p99_s": float(np.percentile(lat64, 99)),
    }
    out["speedup_vs_fp64_plain"] = out["realizations_per_hour"] / out["fp64_plain"]["realizations_per_hour"]
    out["xi_vs_fp64"] = {
        "final_rel_diff_median": float(np.median(rel)), "final_rel_diff_max": float(np.max(rel)),
        "ensemble_mean_final_rel_diff": float(abs(all32 - all64) / all64),
        "per_n_ensemble_mean_rel_diff_max_over_t": mean_rel,
    }
    log(f"fp64 plain: {out['fp64_plain']['realizations_per_hour']:.0f} realizations/h; "
        f"speedup {out['speedup_vs_fp64_plain']:.2f}x; xi ensemble diff {out['xi_vs_fp64']['ensemble_mean_final_rel_diff']:.4f}")
    return out


def stream_probe(nbytes, reps):
    """What plain streaming kernels reach on this GPU's HBM: torch's own
    sum (read only), copy and add (two reads, one write: the step's ratio of
    read to written bytes) over FP64 tensors of nbytes each, best of reps
    runs timed with events on the current stream; one JSON line of GB/s."""
    import torch
    n = nbytes // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.ones(n, dtype=torch.float64, device="cuda")
    c = torch.empty(n, dtype=torch.float64, device="cuda")
    acc = torch.empty((), dtype=torch.float64, device="cuda")
    kinds = {"read": (lambda: torch.sum(a, 0, out=acc), 1), "copy": (lambda: c.copy_(a), 2),
             "add_2r1w": (lambda: torch.add(a, b, out=c), 3)}
    out = {"bytes_per_tensor": nbytes}
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for name, (fn, moved) in kinds.items():
        fn()
        best = float("inf")
        for _ in range(reps):
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1))
        out[name + "_GBps"] = moved * nbytes / (best * 1e-3) / 1e9
    print(json.dumps(out))


def main():
    if "--copy-calibration" in sys.argv:
        copy_calibration(1 << 30, 3)
        return
    if "--stream-probe" in sys.argv:
        stream_probe(1 << 30, 10)
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--precision", type=int, default=64, choices=(64, 32))
    ap.add_argument("--diag-every", type=int, default=20)
    ap.add_argument("--series-steps", type=int, default=1000,
                    help="the FP64 and FP32 runs continue (untimed) to this TimeStep for their diagnostic "
                         "series and the FP32-vs-FP64 comparison")
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--fused-diag", choices=("on", "off"), default="on",
                    help="diagnostics taken by the step kernels (ax_step_diag) or by the separate pass")
    ap.add_argument("--cpu-baseline", choices=("on", "off"), default="on")
    ap.add_argument("--fp32", choices=("on", "off"), default="on", help="FP32 run of the same workload")
    ap.add_argument("--moore", choices=("on", "off"), default="on", help="Moore-model 256^3 run")
    ap.add_argument("--moore-n", type=int, default=256)
    ap.add_argument("--moore-steps", type=int, default=800)
    ap.add_argument("--moore-dt", type=float, default=0.2,
                    help="Moore time step (the program's 0.3 is unstable from a random-phase IC)")
    ap.add_argument("--ensemble", choices=("on", "off"), default="on")
    ap.add_argument("--ensemble-count", type=int, default=256)
    ap.add_argument("--ensemble-rate", type=float, default=50.0, help="Poisson arrivals per second")
    ap.add_argument("--ensemble-seed", type=int, default=2024)
    ap.add_argument("--ensemble-budget-gb", type=float, default=96.0)
    ap.add_argument("--ensemble-chunk", type=int, default=20,
                    help="TimeSteps per graph chunk = xi(t) cadence (the 512^3 run's diagnostic cadence)")
    ap.add_argument("--ensemble-fp64", choices=("on", "off"), default="on")
    ap.add_argument("--boundary-leg", choices=("on", "off"), default="on",
                    help="evolution.cpp schedule with absorbing x/y faces at --boundary-n")
    ap.add_argument("--boundary-n", type=int, default=512)
    ap.add_argument("--boundary-steps", type=int, default=300)
    ap.add_argument("--sor", choices=("on", "off"), default="on", help="straight-string SOR profile")
    ap.add_argument("--sor-cpu-sweeps", type=int, default=1500)
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
    fused = a.fused_diag == "on"
    main_run = run_lattice(ax, a.n, a.precision, a.steps, a.warmup, a.diag_every, a.seed, log,
                           series_to=a.series_steps, fused=fused)
    fp32 = None
    if a.fp32 == "on" and a.precision == 64:
        r32 = run_lattice(ax, a.n, 32, a.steps, a.warmup, a.diag_every, a.seed, log, series_to=a.series_steps,
                          fused=fused)
        fp32 = {"value": r32["value"], "unit": "site-updates/s", "ms_per_step": r32["ms_per_step"],
                "roofline": r32["roofline"], "vs_fp64": compare_series(main_run["series"], r32["series"])}
    moore = None
    if a.moore == "on":
        rm = run_lattice(ax, a.moore_n, 64, a.moore_steps, 10, 20, a.seed, log, moore=True, moore_dt=a.moore_dt)
        last_m = rm["series"][-1]
        moore = {"workload": f"Moore {a.moore_n}^3 periodic FP64, mooreEvol.cpp couplings (q=2,1, dx=0.7), "
                             f"dt={a.moore_dt}, {a.moore_steps} steps, energy+Gauss+winding every 20",
                 "value": rm["value"], "unit": "site-updates/s", "ms_per_step": rm["ms_per_step"],
                 "roofline": rm["roofline"], "final_wound_plaquettes": last_m[3],
                 "finite_to_end": bool(all(r[5] for r in rm["series"])),
                 "dt_note": "mooreEvol.cpp's own dt = 0.3 overflows from this random-phase IC within "
                            "16-22 steps (oracle, 32^3-96^3; tests/test_moore_timestep.py)"}
    ens = None
    if a.ensemble == "on":
        ens = ensemble(ax, a.ensemble_count, a.ensemble_rate, a.ensemble_seed, a.ensemble_budget_gb,
                       a.ensemble_chunk, a.ensemble_fp64 == "on", log)
    bnd = boundary_leg(ax, a.boundary_n, a.boundary_steps, a.seed, log) if a.boundary_leg == "on" else None
    sor = sor_leg(ax, a.sor_cpu_sweeps if a.cpu_baseline == "on" else 0, log) if a.sor == "on" else None
    cpu = None
    if a.cpu_baseline == "on":
        cpu = cpu_baseline(cpu_threads(), log)

    last = main_run["series"][-1]
    out = {
        "metric": METRIC,
        "value": main_run["value"],
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": main_run["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == 64 else "f32",
        "data": f"synthetic: seeded random-phase IC (std::mt19937 seed {a.seed}), pi=0, A=E=0",
        "config": {"workload": f"Abelian-Higgs leapfrog {a.n}^3 periodic FP{a.precision}, dx=0.5 dt=0.1 "
                               f"lambda=2 g=eta=1, energy+Gauss+winding every {a.diag_every} steps",
                   "n": a.n, "precision": a.precision, "diag_every": a.diag_every,
                   "diagnostics": "fused into the step" if fused else "separate pass"},
        "roofline": main_run["roofline"],
        "cpu_baseline": cpu,
        "fp32": fp32,
        "moore": moore,
        "ensemble": ens,
        "boundaries": bnd,
        "sor_profile": sor,
        "final_diag": {"timestep": last[0], "energy": last[1], "gauss_mean": last[2], "wound_plaquettes": last[3],
                       "xi": last[4]},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
