"""axions_run (the reference programs' main() over the library) against the
reference's own outputs: the fixture runs of evolution.cpp, mooreEvol.cpp
and leapfrog.cpp (energy/Gauss series to 1e-12, crossing frames and final
fields bit for bit after a %.17g round trip), and mpi_evolution's random-IC
path against the oracle."""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from axions_amd.abi import AxParams, ah_params
from oracle import oracle_py as orc

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
CLI = os.path.join(os.path.dirname(HERE), "axions_amd", "axions_run")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def load(name):
    fx = json.load(open(os.path.join(GOLDEN, name + ".json")))
    p = AxParams()
    for k, v in fx["params"].items():
        setattr(p, k, v)
    return fx, p


def run(args, cwd):
    say("[cli]", " ".join(args))
    r = subprocess.run([CLI, "--dir", cwd, "--digits", "17", *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def write_ic(path, st, p, ncomp):
    V = p.nx * p.ny * p.nz
    cur = st.reshape(ncomp, 2, V)[:, 0]            # level 0 = the stationary IC
    np.savetxt(path, cur.T, fmt="%.17g")


def consts_ah(p, fx, extra):
    return [f"nx={p.nx}", f"ny={p.ny}", f"nz={p.nz}", f"nt={fx['nt']}", f"dx={p.dx!r}", f"dy={p.dy!r}",
            f"dz={p.dz!r}", f"dt={p.dt!r}", f"lambda={p.lambda_!r}", f"eta={p.eta!r}", f"g={p.g!r}",
            f"damped_nt={p.damped_nt}", f"dampFac={p.dampFac!r}", f"alpha={p.alpha!r}",
            f"scaling={p.scaling!r}", *extra]


BC_NAME = {0: "periodic", 1: "fixed", 2: "neumann", 3: "absorbing"}


@pytest.mark.parametrize("name,split", [("evo_ah_g1_shared", False), ("evo_ah_g1_shared", True),
                                        ("evo_bc_absorbing_zfixed", False), ("evo_bc_neumann_z", True)])
def test_evolution_program(tmp_path, name, split):
    """Energy and Gauss columns against the reference's valsPerLoop, crossing
    frames and final field bit for bit; the string-length column(s) -- the
    chained length of ic_w_evolution.cpp:961-1087 (total, or inner and
    outer with splitLength) of the GPU crossings of every TimeStep after the
    damping window -- against the oracle's all-pairs chaining of the
    oracle's crossings on the oracle's trajectory, bit for bit."""
    fx, p = load(name)
    d = str(tmp_path)
    os.makedirs(os.path.join(d, "Data"), exist_ok=True)
    st = orc.init_random_phase(p, fx["seed"])
    write_ic(os.path.join(d, "Data", "ic.txt"), st, p, 5)
    run(["--program", "evolution", "--tag", "golden",
         *consts_ah(p, fx, ["saveFreq=5", "stationary_ic=true", "finalOut=true", f"xyBC={BC_NAME[p.boundary]}",
                            f"zBC={BC_NAME[p.boundary_z]}", f"splitLength={'true' if split else 'false'}"])], d)
    vals = np.loadtxt(os.path.join(d, "Data", "valsPerLoop_golden.txt"))
    assert vals.shape == (fx["nt"], 4 if split else 3)
    pxy, pz = p.boundary == 0, p.boundary_z == 0
    nonzero = 0
    for n in range(fx["nt"]):
        assert rel(vals[n, 0], fx["energy"][n]) < 1e-12, n
        assert rel(vals[n, 1], fx["gauss"][n]) < 1e-12, n
        want = (0.0, 0.0, 0.0)
        if n >= p.damped_nt:                       # detectBuffer: no detection while damped
            pts, _ = orc.crossings(p, st, n)
            want = orc.string_length(pts, p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, pxy, pz)
        got = tuple(vals[n, 2:])
        assert got == ((want[1], want[2]) if split else (want[0],)), (n, got, want)
        nonzero += want[0] > 0
        orc.step(p, st, n, 1, 0, 16)
    assert nonzero > 0
    for k, fr in enumerate(fx["frames"]):
        path = os.path.join(d, "GifData", f"gifStringPosData_{k}.txt")
        pts = np.loadtxt(path, ndmin=2) if os.path.getsize(path) else np.zeros((0, 3))
        assert pts.shape[0] == fr["count"] and sha(pts) == fr["sha256"], k
    fin = np.loadtxt(os.path.join(d, "Data", "finalField.txt")).T
    assert sha(fin) == fx["final_sha256"]


@pytest.mark.parametrize("name", ["moore_shared", "moore_neumann"])
def test_moore_program(tmp_path, name):
    fx, p = load(name)
    d = str(tmp_path)
    os.makedirs(os.path.join(d, "Data"), exist_ok=True)
    st = orc.init_random_phase(p, fx["seed"])
    write_ic(os.path.join(d, "Data", "moore_ic.txt"), st, p, 7)
    run(["--program", "moore", f"nx={p.nx}", f"ny={p.ny}", f"nz={p.nz}", f"nt={fx['nt']}", "makeGif=false",
         f"xyBC={BC_NAME[p.boundary]}"], d)
    vals = np.loadtxt(os.path.join(d, "Data", "valsPerLoop.txt"))
    for n in range(fx["nt"]):
        assert rel(vals[n, 0], fx["energy"][n]) < 1e-12, n
        if fx["gauss"][n] != 0:
            assert rel(vals[n, 1], fx["gauss"][n]) < 1e-12, n
    fin = np.loadtxt(os.path.join(d, "Data", "finalField.txt")).T
    assert sha(fin) == fx["final_sha256"]


def test_leapfrog_program(tmp_path):
    fx, p = load("global_neumann")
    d = str(tmp_path)
    st = orc.init_random_phase(p, fx["seed"])
    V = p.nx * p.ny * p.nz
    np.savetxt(os.path.join(d, "ic.txt"), st[:2 * V].reshape(2, V).T, fmt="%.17g")
    run(["--program", "leapfrog", f"nx={p.nx}", f"ny={p.ny}", f"nz={p.nz}", f"nt={fx['nt']}"], d)
    e = np.loadtxt(os.path.join(d, "valsPerLoop.txt"))
    for n in range(fx["nt"]):
        assert rel(e[n], fx["energy"][n]) < 1e-12, n
    fin = np.loadtxt(os.path.join(d, "finalField.txt")).T
    assert sha(fin) == fx["final_sha256"]


def test_mpi_evolution_random_ic(tmp_path):
    """Seeded random IC, MPI-form energy every step, crossing frames after the
    damping window, file names with outTag/nx/seed."""
    d = str(tmp_path)
    n, nt, seed = 12, 16, 7
    p = ah_params(n, dx=0.5, dt=0.1, lam=2.0, g=1.0, alpha=3.0, damped_nt=5, dampFac=0.5, energy_form=1)
    run(["--program", "mpi_evolution", f"nx={n}", f"ny={n}", f"nz={n}", f"nt={nt}", "dx=0.5", "dy=0.5", "dz=0.5",
         "dt=0.1", "lambda=2", "g=1", "damped_nt=5", f"seed={seed}", "saveFreq=3", "calcEnergy=true",
         "outTag=t"], d)
    tag = f"t_nx{n}_seed{seed}"
    e = np.loadtxt(os.path.join(d, "Data", f"valsPerLoop_{tag}.txt"))
    st = orc.init_random_phase(p, seed)
    frame = 0
    for t in range(nt):
        o = orc.diag(p, st, t, 0)
        assert rel(e[t], o["energy"]) < 1e-12, t
        if t % 3 == 0 and t >= 5:
            pts = np.loadtxt(os.path.join(d, "GifData", f"gifStringPosData_{tag}_{frame}.txt"), ndmin=2)
            po, co = orc.crossings(p, st, t)
            assert pts.shape[0] == co and np.array_equal(pts, po)
            frame += 1
        orc.step(p, st, t, 1, 0, 16)


def test_straight_string_pipeline(tmp_path):
    """sor -> straight_ic -> evolution, the reference's file chain
    (Data/SOR_Fields.txt, copied to SOR_Fields.txt, -> ic.txt -> Data/ic.txt):
    profile and ic.txt equal the compiled reference's (fixtures), and the
    evolution program runs the string it reads."""
    d = str(tmp_path)
    sor = json.load(open(os.path.join(GOLDEN, "sor_gauged.json")))
    c = sor["consts"]
    run(["--program", "sor", f"nx={c['nx']}", f"h={c['h']!r}", f"w={c['w']!r}", f"tol={c['tol']!r}",
         f"lambda={c['lam']!r}", f"g={c['g']!r}", f"n={c['n']!r}"], d)
    prof = np.loadtxt(os.path.join(d, "Data", "SOR_Fields.txt"))
    assert sha(prof) == sor["profile_sha256"]
    # ic_type 1 continues from the file: tol already met -> one more sweep
    run(["--program", "sor", f"nx={c['nx']}", f"h={c['h']!r}", f"w={c['w']!r}", f"tol={c['tol']!r}",
         f"lambda={c['lam']!r}", f"g={c['g']!r}", "ic_type=1", "maxIter=1"], d)
    more = np.loadtxt(os.path.join(d, "Data", "SOR_Fields.txt"))
    o, it, _ = orc.sor_profile(c["nx"], c["h"], c["w"], c["tol"], c["n"], c["lam"], c["eta"], c["g"], 1, init=prof)
    assert it == 1 and np.array_equal(more, o)
    fx = json.load(open(os.path.join(GOLDEN, "straight_gauged.json")))
    np.savetxt(os.path.join(d, "SOR_Fields.txt"), prof, fmt="%.17g")
    nx, ny, nz = fx["lattice"]
    dx, dy, dz = fx["spacing"]
    run(["--program", "straight_ic", f"nx={nx}", f"ny={ny}", f"nz={nz}", f"dx={dx!r}", f"dy={dy!r}",
         f"dz={dz!r}", f"g={fx['g']!r}", f"SORnx={fx['SORnx']}", f"SORa={fx['SORa']!r}"], d)
    ic = np.loadtxt(os.path.join(d, "ic.txt"))
    assert sha(ic.T.copy()) == fx["ic_sha256"]
    os.replace(os.path.join(d, "ic.txt"), os.path.join(d, "Data", "ic.txt"))
    run(["--program", "evolution", "--tag", "s", f"nx={nx}", f"ny={ny}", f"nz={nz}", f"dx={dx!r}", f"dy={dy!r}",
         f"dz={dz!r}", "dt=0.05", "nt=6", f"g={fx['g']!r}", "lambda=2", "xyBC=fixed", "damped_nt=0",
         "alpha=0", "saveFreq=5", "finalOut=true", "zBC=fixed"], d)
    fin = np.loadtxt(os.path.join(d, "Data", "finalField.txt"))
    p = ah_params(nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, dt=0.05, lam=2.0, g=fx["g"], theta_damp_form=1,
                  boundary=1)
    st = np.repeat(ic.T[:, None, :], 2, axis=1).ravel().copy()
    orc.step(p, st, 0, 6, 0, 16)
    cur, _ = orc.levels(p, st, 6)
    assert np.array_equal(fin.T, cur.reshape(5, -1))


@pytest.mark.parametrize("program", ["evolution", "mpi_evolution"])
def test_two_level_restart(tmp_path, program):
    """Restart from a two-level Data/ic.txt -- the level-0 block, then the
    level-1 block, 5 columns per site in x-major order (evolution.cpp:122-134
    with stationary_ic = false; mpi_evolution.cpp:229-285 with ic_type data):
    the run continues the trajectory the two levels were taken from, as the
    oracle continues it (energies every step to 1e-12, final field bit for
    bit)."""
    d = str(tmp_path)
    os.makedirs(os.path.join(d, "Data"), exist_ok=True)
    nx, ny, nz, k, nt = 12, 10, 14, 6, 10
    mpi = program == "mpi_evolution"
    p = ah_params(nx, ny=ny, nz=nz, g=1.0, theta_damp_form=0 if mpi else 1, energy_form=1 if mpi else 0)
    st = orc.init_random_phase(p, 5)
    orc.step(p, st, 0, k, 0, 16)        # k even: level 1 is the current level of TimeStep 0 again
    V = nx * ny * nz
    s5 = st.reshape(5, 2, V)
    np.savetxt(os.path.join(d, "Data", "ic.txt"), np.concatenate([s5[:, 0].T, s5[:, 1].T]), fmt="%.17g")
    common = [f"nx={nx}", f"ny={ny}", f"nz={nz}", f"nt={nt}", "dx=0.5", "dy=0.5", "dz=0.5", "dt=0.1", "lambda=2",
              "g=1", "damped_nt=0", "alpha=0", "stringDetect=false"]
    if mpi:
        run(["--program", "mpi_evolution", *common, "ic_type=data", "seed=3", "calcEnergy=true", "outTag=r"], d)
        e = np.loadtxt(os.path.join(d, "Data", f"valsPerLoop_r_nx{nx}_seed3.txt"))
    else:
        run(["--program", "evolution", "--tag", "r", *common, "stationary_ic=false", "finalOut=true"], d)
        e = np.loadtxt(os.path.join(d, "Data", "valsPerLoop_r.txt"))[:, 0]
    assert e.shape == (nt,)
    for t in range(nt):
        assert rel(e[t], orc.diag(p, st, t, 0)["energy"]) < 1e-12, t
        orc.step(p, st, t, 1, 0, 16)
    if not mpi:
        fin = np.loadtxt(os.path.join(d, "Data", "finalField.txt")).T
        cur, _ = orc.levels(p, st, nt)
        assert np.array_equal(fin, cur.reshape(5, -1))
