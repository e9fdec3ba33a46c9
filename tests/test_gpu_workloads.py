"""The batched and served workloads at their own sizes, against the oracle.

* The 256-box ensemble (gen_box(4, k), k = 0..255: N log-uniform in [1k, 64k],
  3.6M spheres, every box sheared) stepped 5 times in one captured-graph
  batch: every box equals the oracle stepping it alone, bit for bit.
* Twelve served-stream jobs (gen_box(5, k), chosen to span the N range, with
  their generated step counts) through the server together: a job whose
  oracle run fits the test budget equals the oracle bit for bit; a larger one
  equals the same job stepped alone on the device, whose end state (stepped
  on to a rebuild) holds the k-d tree image-pair set and exact force balance.
"""
import ctypes as C
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import soil_amd
from oracle_ref import RefBox
from paritylib import assert_same, make_pair
from propcheck import check_csr_well_formed, check_force_balance, check_image_pair_set
from soil_amd.serve import Server

pytestmark = pytest.mark.gpu

SEED = 0x50494C
ORACLE_BUDGET = 3.0e7  # particle-steps of one job's oracle run
WORKERS = max(1, min(8, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def ctx():
    return soil_amd.Context(0)


def test_c4_ensemble_matches_the_oracle(ctx):
    print("case: C4 256-box ensemble, 5 graph steps", flush=True)
    gens = [soil_amd.gen_box(4, k) for k in range(256)]
    assert sum(len(g[1]) for g in gens) > 3_000_000
    pairs = [make_pair(ctx, *g, k=1000.0) for g in gens]
    batch = soil_amd.Batch(ctx, [p[0] for p in pairs])
    batch.step(5, graph=True)
    batch.sync()
    batch.close()
    print("device done; oracle stepping", flush=True)
    with ThreadPoolExecutor(WORKERS) as ex:  # the oracle releases the interpreter while it steps
        list(ex.map(lambda p: p[1].step(5), pairs))
    for k, (g, r) in enumerate(pairs):
        assert_same(g, r, what=f"C4 box {k}")
        g.close()


def c5_jobs(count=12, pool=96):
    """`count` served-stream jobs spanning the N range of the first `pool`."""
    sizes = []
    for k in range(pool):
        d = soil_amd._ffi.BoxDesc()
        n = C.c_size_t(0)
        soil_amd._ffi.check(soil_amd._ffi.lib().soil_hip_gen_box(5, k, SEED, C.byref(d), C.byref(n), None, None,
                                                                 None, None))
        sizes.append(n.value)
    order = np.argsort(sizes)
    pick = sorted(set(int(order[int(round(q))]) for q in np.linspace(0, pool - 1, count)))
    _, steps = soil_amd.gen_arrivals(SEED, pool, 6.0)
    return [(k, soil_amd.gen_box(5, k, SEED), int(steps[k])) for k in pick]


def test_c5_served_jobs_match_the_oracle_or_hold_the_invariants(ctx):
    print("case: C5 served jobs", flush=True)
    jobs = c5_jobs()
    ns = [len(g[1]) for _, g, _ in jobs]
    print(f"jobs N {ns} steps {[s for _, _, s in jobs]}", flush=True)
    assert min(ns) < 4000 and max(ns) > 100_000
    assert any(g[0].has_shear for _, g, _ in jobs) and any(not g[0].has_shear for _, g, _ in jobs)
    sv = Server(ctx, budget_gb=200.0, max_slots=256, chunk_steps=16)
    sv.start()
    ids = [sv.submit(d, pos, vel, mass, rad, steps, keep_result=True) for _, (d, pos, vel, mass, rad), steps in jobs]
    sv.drain()
    st = sv.stats()
    print({k: st[k] for k in ("jobs_done", "jobs_failed", "mean_active_jobs", "step_p99_ms")}, flush=True)
    assert st["jobs_done"] == len(jobs) and st["jobs_failed"] == 0 and st["mean_active_jobs"] > 1.0
    results = {jid: sv.result(jid, len(g[1][3])) for jid, g in zip(ids, jobs)}
    sv.close()

    small = [(jid, job) for jid, job in zip(ids, jobs) if len(job[1][1]) * job[2] <= ORACLE_BUDGET]
    large = [(jid, job) for jid, job in zip(ids, jobs) if len(job[1][1]) * job[2] > ORACLE_BUDGET]
    assert small and large

    def oracle(job):
        _, (d, pos, vel, mass, rad), steps = job
        r = RefBox(d)
        r.add_atoms(pos, mass, rad, vel=vel)
        r.set_consumer(1000.0)
        r.setup()
        r.step(steps)
        a = r.atoms()
        n = len(mass)
        return a["pos"][:n], a["vel"][:n]

    print(f"oracle runs: {len(small)} jobs", flush=True)
    with ThreadPoolExecutor(WORKERS) as ex:
        refs = list(ex.map(lambda t: oracle(t[1]), small))
    for (jid, job), (p, v) in zip(small, refs):
        pm, vm = results[jid]
        assert np.array_equal(pm, p), f"job {job[0]} (N={len(p)}): positions differ from the oracle"
        assert np.array_equal(vm, v), f"job {job[0]} (N={len(p)}): velocities differ from the oracle"

    for jid, (k, (d, pos, vel, mass, rad), steps) in large:
        print(f"job {k} N={len(mass)} steps={steps}: alone on the device + invariants", flush=True)
        b = soil_amd.Box(ctx, d)
        b.upload(pos, mass, rad, vel=vel)
        b.set_consumer(1000.0)
        b.setup()
        b.step(steps)
        a = b.atoms()
        n = len(mass)
        pm, vm = results[jid]
        assert np.array_equal(pm, a["pos"][:n]) and np.array_equal(vm, a["vel"][:n]), f"job {k}: mix != alone"
        before = b.counts()["rebuilds"]
        for _ in range(60):  # on to a rebuild: the stored positions are then the build's
            b.step(1)
            if b.counts()["rebuilds"] > before:
                break
        assert b.counts()["rebuilds"] > before
        a = b.atoms()
        off, idx = b.csr()
        check_csr_well_formed(off, idx, b.counts())
        check_image_pair_set(a, off, idx, d.skin_fraction)
        check_force_balance(a)
        b.close()
