"""Build libaxions.so in-tree with hipcc for gfx950.

Flags that matter for results:
  -ffp-contract=off   no a*b+c -> fma contraction, so every expression rounds
                      like the reference's x86-64 build (explicit fma() calls
                      in ax_trig.h stay fused on host and device alike);
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions (v_pk_*_f32) are generated.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
SRC = ["ax_api.hip", "ax_step.hip", "ax_step_tiled.hip", "ax_diag.hip", "ax_pool.hip", "ax_global.hip", "ax_length.hip",
       "ax_profile.hip", "ax_bound.hip", "ax_noalloc.hip"]
LIB = os.path.join(HERE, "libaxions.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AX_OFFLOAD_ARCH", "gfx950")


def flags():
    return ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(REPO, "include"), "-Wall", "-Wno-unused-function",
            "-Wno-unused-result"]


def build(verbose=False):
    objs = []
    odir = os.path.join(HERE, "_obj")
    os.makedirs(odir, exist_ok=True)
    procs = []
    for s in SRC:
        src = os.path.join(HERE, "csrc", s)
        obj = os.path.join(odir, s.replace(".hip", ".o"))
        cmd = [HIPCC, *flags(), "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    bad = []
    for s, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            bad.append(s)
            sys.stderr.write(out)
        elif verbose and out.strip():
            sys.stderr.write(out)
    if bad:
        raise RuntimeError(f"hipcc failed for {bad}")
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"], check=True)
    os.replace(tmp, LIB)
    build_cli(verbose)
    return LIB


CLI = os.path.join(HERE, "axions_run")


def build_cli(verbose=False):
    """axions_run: the reference programs' main() over libaxions.so (plain C++)."""
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-I", os.path.join(REPO, "include"),
           os.path.join(HERE, "cli", "axions_run.cpp"), "-o", CLI + ".tmp", "-L", HERE, "-laxions",
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(CLI + ".tmp", CLI)
    return CLI


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
