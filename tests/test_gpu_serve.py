"""GPU checks of the request server (afmm_server_*): continuous batching of
force and integration jobs of different N and order p.

  * every job's result is bit-identical to the same job run alone (a plain
    evaluation or a resident leapfrog of that system), whatever batch it ran
    in, with a memory budget small enough that jobs queue;
  * admission follows arrival order and the admitted footprint never exceeds
    the budget;
  * a job that arrives while others run is admitted and finishes without
    waiting for the earlier ones to drain;
  * argument errors are reported, not thrown past the ABI.
"""
import threading
import time

import numpy as np
import pytest

import cabi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def afmm():
    return cabi.product()


def make_jobs(seed, count):
    orc = cabi.oracle()
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(500), np.log(20000))))
        pos, vel, m = orc.plummer(n, seed=int(rng.integers(1 << 30)))
        jobs.append(dict(pos=pos, vel=vel, mass=m, G=float(rng.uniform(0.5, 2.0)),
                         eps=float(rng.choice([0.005, 0.01, 0.02])), dt=float(rng.choice([0.005, 0.01])),
                         p=int(rng.choice([4, 6, 8])), steps=int(rng.integers(0, 5))))
    return jobs


def run_alone(afmm, j, precision):
    s = afmm.AdaptiveFMM3D(G=j["G"], eps=j["eps"], p=j["p"], threshold=64, precision=precision)
    if j["steps"] == 0:
        return j["pos"], j["vel"], s.computeAccelerations(j["pos"], j["mass"])
    st = afmm.LeapfrogState(s, j["pos"], j["vel"], j["mass"])
    st.step(j["dt"], j["steps"])
    return st.download()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_results_equal_alone(afmm, prec):
    precision = afmm.PRECISION_FP64 if prec == "fp64" else afmm.PRECISION_MIXED
    jobs = make_jobs(5 if prec == "fp64" else 6, 14)
    budget = 3 * max(afmm.Server.job_bytes(len(j["mass"]), j["p"]) for j in jobs)
    sv = afmm.Server(hbm_budget=budget, precision=precision)
    print(f"server {prec}: {len(jobs)} jobs, budget {budget / 2**20:.0f} MiB", flush=True)
    ids = [sv.submit(j["pos"], j["vel"], j["mass"], j["G"], j["eps"], j["dt"], j["p"], j["steps"]) for j in jobs]
    sv.drain()
    infos = [sv.info(i) for i in ids]
    assert all(i["state"] == afmm.JOB_DONE for i in infos)
    admits = [i["t_admit"] for i in infos]
    assert admits == sorted(admits)  # arrival order
    assert max(i["batches"] for i in infos) > 1
    for j, i, info in zip(jobs, ids, infos):
        assert info["steps_done"] == j["steps"]
        got = sv.result(i)
        ref = run_alone(afmm, j, precision)
        for a, b, name in zip(got, ref, ("pos", "vel", "acc")):
            assert np.array_equal(a, b), (i, name)
    sv.close()


def test_budget_and_late_arrival(afmm):
    orc = cabi.oracle()
    big = [orc.plummer(30000, seed=s) for s in (1, 2)]
    small = orc.plummer(2000, seed=3)
    b_big = afmm.Server.job_bytes(30000, 8)
    sv = afmm.Server(hbm_budget=int(1.5 * b_big))
    sv.warm(8)
    sv.warm(4)
    seen = []
    stop = threading.Event()

    def watch():
        while not stop.is_set():
            seen.append(sv.hbm_in_use())
            time.sleep(0.001)

    t = threading.Thread(target=watch)
    t.start()
    i1 = sv.submit(*big[0], G=1.0, eps=0.01, dt=0.01, p=8, steps=20)
    i2 = sv.submit(*big[1], G=1.0, eps=0.01, dt=0.01, p=8, steps=2)   # must wait for i1 (budget)
    time.sleep(0.05)
    i3 = sv.submit(*small, G=1.0, eps=0.01, dt=0.01, p=4, steps=1)    # queued behind i2 (arrival order)
    sv.drain()
    stop.set()
    t.join()
    inf = {k: sv.info(k) for k in (i1, i2, i3)}
    assert max(seen) <= int(1.5 * b_big)
    assert inf[i2]["t_admit"] >= inf[i1]["t_done"]
    assert inf[i3]["t_admit"] >= inf[i2]["t_admit"]
    # a late arrival joins a running batch of its order: submit while a long job
    # runs (300 steps of 30k bodies: several hundred ms, far past the 50 ms wait)
    sv2 = afmm.Server()
    sv2.warm(6)
    a = sv2.submit(*big[0], G=1.0, eps=0.01, dt=0.01, p=6, steps=300)
    time.sleep(0.05)
    b = sv2.submit(*small, G=1.0, eps=0.01, dt=0.01, p=6, steps=2)
    sv2.wait(b)
    ia, ib = sv2.info(a), sv2.info(b)
    assert ib["t_done"] < ia["t_done"] or ia["state"] != afmm.JOB_DONE
    sv2.drain()
    sv.close()
    sv2.close()


def test_server_argument_errors(afmm):
    sv = afmm.Server()
    pos = np.zeros((3, 3))
    m = np.ones(3)
    with pytest.raises(afmm.AfmmError) as e:
        sv.submit(pos, None, m, p=0)
    assert e.value.code == -2
    with pytest.raises(afmm.AfmmError) as e:
        sv.submit(pos, None, m, p=4, steps=2)  # integration needs velocities
    assert e.value.code == -1
    with pytest.raises(afmm.AfmmError):
        sv.submit(np.zeros((0, 3)), None, np.zeros(0), p=4)
    sv.close()
