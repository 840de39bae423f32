#!/usr/bin/env python3
"""Build the reference programs (sjcott/Axions) for the oracle.  TEST
INFRASTRUCTURE ONLY.

The reference keeps its run parameters as file-scope consts
(evolution.cpp:23-57, mooreEvol.cpp:19-41) that cannot be overridden from
the command line.  This recipe reads the reference source where it lies,
substitutes the values of named consts in the text it pipes to g++ (nothing
is copied to disk), and compiles it with the reference's own recipe
(g++ -O3 -fopenmp, submit_job.sh:27-31) plus:
  -fmacro-prefix-map  so the program's __FILE__-relative Data/ and GifData/
                      paths (evolution.cpp:80-91) resolve in its run folder;
  -include ref_shim.h full-precision ofstream output; optional shared trig.
Outputs go to oracle/_ref/<name>/ only.  Every other byte of the compiled
program is the reference's.

Usage: build_ref.py [--ref /root/reference] [name ...]   (default: all)
"""
import argparse
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(REPO, "oracle", "_ref")

# name -> (source file, const overrides, shared trig?)
_AH_SMALL = dict(nx=16, ny=16, nz=16, nt=60, saveFreq=5, stationary_ic="true",
                 finalOut="true", stringDetect="true", detectBuffer="true",
                 makeStringPosGif="true")
CONFIGS = {
    # global string (g = 0) with PRS friction and early damping; theta stays
    # exactly 0, so sin/cos only ever see 0 and libm is irrelevant.
    "evo_ah_g0": ("evolution.cpp", dict(_AH_SMALL, dx=0.5, dy=0.5, dz=0.5, dt=0.05,
                                        **{"lambda": 1}, eta=1, g=0, damped_nt=20,
                                        dampFac=0.5, alpha=3, scaling=1), False),
    # gauged, friction + damping, shared trig and libm trig
    "evo_ah_g1_shared": ("evolution.cpp", dict(_AH_SMALL, dx=0.5, dy=0.5, dz=0.5, dt=0.1,
                                               **{"lambda": 2}, eta=1, g=1, damped_nt=20,
                                               dampFac=0.5, alpha=3, scaling=1), True),
    "evo_ah_g1_libm": ("evolution.cpp", dict(_AH_SMALL, dx=0.5, dy=0.5, dz=0.5, dt=0.1,
                                             **{"lambda": 2}, eta=1, g=1, damped_nt=20,
                                             dampFac=0.5, alpha=3, scaling=1), False),
    # non-cubic, non-power-of-two spacings (true divisions everywhere)
    "evo_ah_aniso": ("evolution.cpp", dict(_AH_SMALL, nx=12, ny=10, nz=14, nt=40,
                                           dx=0.7, dy=0.6, dz=0.55, dt=0.13,
                                           **{"lambda": 1.5}, eta=0.9, g=0.8,
                                           damped_nt=10, dampFac=0.3, alpha=2,
                                           scaling=1), True),
    # the plain AH workload at 64^3 (dx=0.5 dt=0.1 lambda=2 g=eta=1), 200 steps
    "evo_c1_shared": ("evolution.cpp", dict(nx=64, ny=64, nz=64, nt=200, dx=0.5, dy=0.5, dz=0.5,
                                            dt=0.1, **{"lambda": 2}, eta=1, g=1, damped_nt=0,
                                            dampFac=0.5, alpha=0, scaling=1, saveFreq=10,
                                            stationary_ic="true", finalOut="true",
                                            stringDetect="true", detectBuffer="true",
                                            makeStringPosGif="true"), True),
    # the 512^3 run's couplings and cadence (plain AH, dx=0.5 dt=0.1 lambda=2
    # g=eta=1, energy and Gauss every step, crossings every 20 steps) at 64^3
    # over 1000 TimeSteps
    "evo_c2_1000_shared": ("evolution.cpp", dict(nx=64, ny=64, nz=64, nt=1000, dx=0.5, dy=0.5, dz=0.5,
                                                 dt=0.1, **{"lambda": 2}, eta=1, g=1, damped_nt=0,
                                                 dampFac=0.5, alpha=0, scaling=1, saveFreq=20,
                                                 stationary_ic="true", finalOut="true",
                                                 stringDetect="true", detectBuffer="true",
                                                 makeStringPosGif="true"), True),
    # the CPU baseline of the benchmark: the program as shipped (glibc trig,
    # OpenMP) on a 128^3 periodic lattice with the benchmark's couplings,
    # no string output; two step counts, so that the difference of their run
    # times is the time of 20 TimeSteps without the ic.txt read
    **{name: ("evolution.cpp", dict(nx=128, ny=128, nz=128, nt=nt, dx=0.5, dy=0.5, dz=0.5, dt=0.1,
                                    **{"lambda": 2}, eta=1, g=1, damped_nt=0, dampFac=0.5, alpha=0, scaling=1,
                                    stationary_ic="true", finalOut="false", stringDetect="false",
                                    makeStringPosGif="false", makeGif="false"), False)
       for name, nt in (("cpu_bench_nt2", 2), ("cpu_bench_nt22", 22))},
    # Moore model as shipped (fixed edges), reference couplings
    "moore_shared": ("mooreEvol.cpp", dict(nx=12, ny=12, nz=12, nt=30, makeGif="false"), True),
    "moore_libm": ("mooreEvol.cpp", dict(nx=12, ny=12, nz=12, nt=30, makeGif="false"), False),
    # global string, momentum-form leapfrog (no trigonometry): the program's
    # own spacings, and an anisotropic set with non-power-of-two spacings
    "global_neumann": ("leapfrog.cpp", dict(nx=12, ny=10, nz=14, nt=40), False),
    "global_aniso": ("leapfrog.cpp", dict(nx=9, ny=11, nz=13, nt=30, dx=0.3, dy=0.35, dz=0.28, dt=0.07), False),
    # string length by chaining (ic_w_evolution.cpp:961-1087) on the crossings
    # of ic_func.cpp's seeded "random" initial condition, first TimeStep only
    "icw_len_split": ("ic_w_evolution.cpp", dict(nx=15, ny=14, nz=16, nt=1, dx=0.35, dy=0.35, dz=0.35,
                                                 splitLength="true"), False),
    "icw_len_total": ("ic_w_evolution.cpp", dict(nx=12, ny=13, nz=11, nt=1, dx=0.5, dy=0.45, dz=0.4,
                                                 splitLength="false"), False),
    # non-periodic boundaries of evolution.cpp (xyBC x zBC) and mooreEvol.cpp
    "evo_bc_neumann": ("evolution.cpp", dict(_AH_SMALL, nx=14, ny=12, nz=10, nt=40, dx=0.5, dy=0.5, dz=0.5, dt=0.1,
                                             **{"lambda": 2}, eta=1, g=1, damped_nt=10, dampFac=0.5, alpha=2,
                                             scaling=1, xyBC='"neumann"', zBC='"periodic"'), True),
    "evo_bc_neumann_z": ("evolution.cpp", dict(_AH_SMALL, nx=13, ny=11, nz=9, nt=40, dx=0.6, dy=0.5, dz=0.55,
                                               dt=0.1, **{"lambda": 1.5}, eta=1, g=0.8, damped_nt=8, dampFac=0.4,
                                               alpha=2, scaling=1, xyBC='"neumann"', zBC='"neumann"'), True),
    "evo_bc_absorbing": ("evolution.cpp", dict(_AH_SMALL, nx=14, ny=13, nz=10, nt=40, dx=0.5, dy=0.5, dz=0.5,
                                               dt=0.1, **{"lambda": 2}, eta=1, g=1, damped_nt=10, dampFac=0.5,
                                               alpha=2, scaling=1, xyBC='"absorbing"', zBC='"periodic"'), True),
    "evo_bc_absorbing_zfixed": ("evolution.cpp", dict(_AH_SMALL, nx=12, ny=14, nz=11, nt=40, dx=0.6, dy=0.55,
                                                      dz=0.5, dt=0.1, **{"lambda": 1.5}, eta=0.9, g=0.8,
                                                      damped_nt=8, dampFac=0.4, alpha=2, scaling=1,
                                                      xyBC='"absorbing"', zBC='"fixed"'), True),
    "evo_bc_fixed_zper": ("evolution.cpp", dict(_AH_SMALL, nx=13, ny=12, nz=10, nt=40, dt=0.1, **{"lambda": 2},
                                                eta=1, g=1, damped_nt=10, dampFac=0.5, alpha=2, scaling=1,
                                                xyBC='"fixed"', zBC='"periodic"'), True),
    "evo_bc_fixed_zneu": ("evolution.cpp", dict(_AH_SMALL, nx=12, ny=11, nz=12, nt=40, dt=0.1, **{"lambda": 2},
                                                eta=1, g=1, damped_nt=10, dampFac=0.5, alpha=2, scaling=1,
                                                xyBC='"fixed"', zBC='"neumann"'), True),
    "moore_neumann": ("mooreEvol.cpp", dict(nx=12, ny=11, nz=10, nt=30, makeGif="false", xyBC='"neumann"'), True),
    # straight-string pipeline: SOR radial profile (singleSemiAnalytic.cpp)
    # as shipped (global string, 30001 points), a gauged one, and one that
    # ends at its sweep cap; then the lattice interpolation (straight_ic.cpp)
    # as shipped and on an anisotropic lattice reaching past its profile
    "sor_ref": ("singleSemiAnalytic.cpp", {}, False),
    "sor_gauged": ("singleSemiAnalytic.cpp", dict(nx=1501, h=0.02, g=1, **{"lambda": 2}, w=1.6, tol=1e-4), False),
    "sor_cap": ("singleSemiAnalytic.cpp", dict(nx=801, h=0.03, maxIter=300, g=0.5, n=2), False),
    "straight_ref": ("straight_ic.cpp", {}, False),
    "straight_gauged": ("straight_ic.cpp", dict(nx=40, ny=33, nz=3, dx=0.45, dy=0.55, dz=0.5, g=1, SORnx=403,
                                                SORa=0.02), False),
}

# extra translation units compiled from their own substituted text
EXTRA = {
    "icw_len_split": {"ic_func.cpp": dict(ic_type='"random"', seed=3)},
    "icw_len_total": {"ic_func.cpp": dict(ic_type='"random"', seed=5)},
}


def substituted_source(path, overrides):
    text = open(path).read()
    for name, val in overrides.items():
        if isinstance(val, float) and val == int(val):
            val = repr(val)
        pat = re.compile(r"^(\s*const\s+(?:int|double|bool|string)\s+" + re.escape(name) + r"\s*=\s*)([^;]*)(;)",
                         re.M)
        text, n = pat.subn(lambda m: m.group(1) + str(val) + m.group(3), text)
        if n != 1:
            raise SystemExit(f"{os.path.basename(path)}: const {name} matched {n} times")
    return text


def build(name, ref, cxx="g++"):
    src, over, shared = CONFIGS[name]
    path = os.path.join(ref, src)
    dest = os.path.join(OUT, name)
    os.makedirs(os.path.join(dest, "Data"), exist_ok=True)
    os.makedirs(os.path.join(dest, "GifData"), exist_ok=True)
    text = substituted_source(path, over)
    objs = []
    for extra_src, extra_over in EXTRA.get(name, {}).items():
        obj = os.path.join(dest, extra_src.replace(".cpp", ".o"))
        r = subprocess.run([cxx, "-x", "c++", "-", "-c", "-I", ref, "-I", os.path.join(REPO, "include"),
                            "-include", os.path.join(HERE, "ref_shim.h"), "-O3", "-fopenmp", "-o", obj],
                           input=substituted_source(os.path.join(ref, extra_src), extra_over), text=True,
                           capture_output=True)
        if r.returncode != 0:
            sys.stderr.write(r.stderr)
            raise SystemExit(f"build of {name}/{extra_src} failed")
        objs.append(obj)
    cmd = [cxx, "-x", "c++", "-", "-x", "none", os.path.join(ref, "array.cpp"), *objs,
           "-I", ref, "-I", os.path.join(REPO, "include"),
           "-include", os.path.join(HERE, "ref_shim.h"),
           f"-fmacro-prefix-map=<stdin>=./{src}",
           "-O3", "-fopenmp", "-o", os.path.join(dest, "prog")]
    if shared:
        cmd.insert(1, "-DAX_REF_SHARED_TRIG")
    r = subprocess.run(cmd, input=text, text=True, capture_output=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise SystemExit(f"build of {name} failed")
    with open(os.path.join(dest, "overrides.txt"), "w") as f:
        f.write(f"{src} shared_trig={shared}\n")
        for k, v in over.items():
            f.write(f"{k}={v}\n")
    return dest


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("names", nargs="*")
    a = ap.parse_args()
    if not os.path.isdir(a.ref):
        print("reference sources not present; nothing to build")
        return 0
    for n in (a.names or list(CONFIGS)):
        print("built", build(n, a.ref))
    return 0


if __name__ == "__main__":
    sys.exit(main())
