"""GPU tests of the ensemble pool (continuous batching, persistent batched
kernel, hipGraph chunks, HBM-budget admission, xi streaming).

Bar: each realization's final field and xi(t) series are bit-identical to the
same realization run alone through the single-lattice path (ax_step /
ax_diagnostics), whatever else shares the pool, with or without graphs; the
FP64 pool is bit-identical to the oracle.
"""
import sys
import time

import numpy as np
import pytest

import axions_amd as ax
from axions_amd.abi import AX_EBUDGET, AX_REQ_DONE, lattice_bytes
from oracle import oracle_py as orc

pytestmark = pytest.mark.gpu
GB = 1 << 30


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def single_run(n, seed, steps, chunk, precision=32):
    """Final state and xi samples (t = 0, every chunk end, last step) of one
    realization through the single-lattice API."""
    p = ax.ah_params(n, precision=precision)
    ts, wound, xi = [], [], []
    with ax.Lattice(p) as lat:
        lat.init_random_phase(seed)
        t = 0
        while True:
            d = lat.diagnostics(winding=True)
            ts.append(t)
            wound.append(d["wound_plaquettes"])
            xi.append(d["xi"])
            if t == steps:
                break
            k = min(chunk, steps - t)
            lat.step(k)
            t += k
        return lat.get_state(), np.array(ts), np.array(wound), np.array(xi)


def check_same(pool, rid, ref):
    st, ts, wound, xi = ref
    pts, pw, px = pool.xi(rid)
    assert np.array_equal(pts, ts), (pts, ts)
    assert np.array_equal(pw, wound)
    assert np.array_equal(px.view(np.uint64), xi.view(np.uint64))
    got = pool.get_state(rid)
    assert np.array_equal(got.view(np.uint64), st.view(np.uint64))


@pytest.mark.parametrize("use_graph", [1, 0])
def test_pool_matches_single_lattice(use_graph):
    say(f"[pool] vs single, graph={use_graph}")
    reqs = [(64, 11, 25), (96, 12, 17), (40, 13, 31)]
    with ax.Pool(ax.ah_params(8, precision=32), hbm_budget=2 * GB, chunk_steps=10, use_graph=use_graph,
                 autotune=0, ic_threads=2) as pool:
        ids = [pool.submit(seed, n, steps, keep_state=True) for n, seed, steps in reqs]
        assert pool.wait(timeout=120)
        for rid, (n, seed, steps) in zip(ids, reqs):
            info = pool.info(rid)
            assert info["status"] == AX_REQ_DONE
            assert info["t_submit"] <= info["t_admit"] <= info["t_done"]
            check_same(pool, rid, single_run(n, seed, steps, 10))
            pool.release(rid)


def test_alone_equals_mixed():
    """One realization alone, then inside a stream of others of uneven N that
    arrive while it runs (continuous batching): identical results."""
    say("[pool] alone vs mixed")
    target = (64, 99, 40)
    params = ax.ah_params(8, precision=32)
    with ax.Pool(params, hbm_budget=2 * GB, chunk_steps=5, autotune=0, ic_threads=2) as pool:
        rid = pool.submit(target[1], target[0], target[2], keep_state=True)
        assert pool.wait(rid, 120)
        alone = (pool.get_state(rid),) + pool.xi(rid)
    # budget for about three resident lattices, so admissions wait on completions
    budget = 3 * lattice_bytes(96, 32)
    with ax.Pool(params, hbm_budget=budget, chunk_steps=5, autotune=0, ic_threads=3) as pool:
        first = [pool.submit(s, n, st) for s, n, st in [(1, 96, 30), (2, 48, 45)]]
        time.sleep(0.05)
        rid = pool.submit(target[1], target[0], target[2], keep_state=True)
        later = []
        for k, (n, st) in enumerate([(32, 20), (80, 15), (96, 25), (64, 10), (48, 35)]):
            time.sleep(0.01)
            later.append(pool.submit(100 + k, n, st))
        assert pool.wait(timeout=180)
        s = pool.stats()
        assert s["completed"] == 8 and s["failed"] == 0
        assert s["max_resident"] >= 2                      # the work overlapped
        assert s["peak_in_use"] <= budget
        st = pool.get_state(rid)
        ts, w, x = pool.xi(rid)
        assert np.array_equal(st.view(np.uint64), alone[0].view(np.uint64))
        assert np.array_equal(ts, alone[1]) and np.array_equal(w, alone[2])
        assert np.array_equal(x.view(np.uint64), alone[3].view(np.uint64))


def test_fp64_pool_matches_oracle():
    say("[pool] fp64 vs oracle")
    n, seed, steps = 20, 5, 12
    p = ax.ah_params(n)
    st = orc.init_random_phase(p, seed)
    orc.step(p, st, 0, steps, 0, 16)
    with ax.Pool(ax.ah_params(8), hbm_budget=GB // 4, chunk_steps=5, autotune=0, ic_threads=1) as pool:
        rid = pool.submit(seed, n, steps, keep_state=True)
        assert pool.wait(rid, 120)
        got = pool.get_state(rid)
        assert np.array_equal(got.view(np.uint64), st.view(np.uint64))
        ts, w, x = pool.xi(rid)
        assert list(ts) == [0, 5, 10, 12]
        wo = orc.winding(p, st, steps, 0)
        assert w[-1] == int(np.abs(wo.astype(np.int64)).sum())


def test_budget_and_autotune():
    say("[pool] budget + autotune")
    params = ax.ah_params(8, precision=32)
    budget = 2 * lattice_bytes(64, 32) + 1024
    with ax.Pool(params, hbm_budget=budget, chunk_steps=10, autotune=1, ic_threads=2) as pool:
        with pytest.raises(ax.AxError) as e:
            pool.submit(1, 128, 10)
        assert e.value.code == AX_EBUDGET
        ids = [pool.submit(s, 64, 20) for s in range(5)]
        assert pool.wait(timeout=120)
        s = pool.stats()
        assert s["completed"] == 5 and s["max_resident"] <= 2 and s["peak_in_use"] <= budget
        xs = {pool.info(i)["xseg"] for i in ids}
        assert len(xs) == 1 and xs.pop() in (4, 8, 16, 32, 64)
        x, ms = pool.tune(32)
        assert x in (4, 8, 16, 32) and ms > 0


def test_failed_requests_free_their_memory():
    """Lattices whose state turns non-finite end as failed; their arena slots
    are released only after every chunk that may still touch them has
    finished, also when the caller releases the requests before that."""
    from axions_amd.abi import AX_REQ_FAILED
    say("[pool] failed requests")
    # dt = 5 at dx = 0.5 is far beyond the leapfrog's stability limit: FP32
    # states overflow within a few steps
    params = ax.ah_params(8, dt=5.0, precision=32)
    budget = 3 * lattice_bytes(32, 32) + 1024
    with ax.Pool(params, hbm_budget=budget, chunk_steps=4, autotune=0, ic_threads=2) as pool:
        ids = [pool.submit(s, 32, 200) for s in range(7)]
        assert pool.wait(timeout=120)
        for i in ids:
            assert pool.info(i)["status"] == AX_REQ_FAILED
            pool.release(i)
        t0 = time.time()
        while pool.stats()["in_use"] and time.time() - t0 < 30:
            time.sleep(0.01)
        s = pool.stats()
        assert s["failed"] == 7 and s["completed"] == 0
        assert s["in_use"] == 0 and s["peak_in_use"] <= budget and s["max_resident"] <= 3


@pytest.mark.parametrize("precision", [32, 64])
def test_pool_short_z_chunks(precision):
    """Lattice edges whose last z chunk is short -- one site (65), two sites
    (66) -- or a whole row of 33: the pool marches them (the single-lattice
    path steps such rows with the site kernel), and each realization matches
    its single-lattice run bit for bit."""
    say(f"[pool] short chunks fp{precision}")
    reqs = [(65, 21, 9), (66, 22, 7), (33, 23, 12)]
    with ax.Pool(ax.ah_params(8, precision=precision), hbm_budget=2 * GB, chunk_steps=4, autotune=0,
                 ic_threads=2) as pool:
        ids = [pool.submit(seed, n, steps, keep_state=True) for n, seed, steps in reqs]
        assert pool.wait(timeout=120)
        for rid, (n, seed, steps) in zip(ids, reqs):
            assert pool.info(rid)["status"] == AX_REQ_DONE
            check_same(pool, rid, single_run(n, seed, steps, 4, precision))
            pool.release(rid)
