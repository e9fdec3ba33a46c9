"""Served stream on the GPU: each job's result is the same alone and in the mix."""
import numpy as np
import pytest

import soil_amd
from soil_amd.serve import Server

pytestmark = pytest.mark.gpu


def test_served_jobs_match_alone():
    print("case: server alone vs mix", flush=True)
    ctx = soil_amd.Context(0)
    rng = np.random.default_rng(77)
    jobs = []
    for k in range(7):
        n = int(rng.integers(300, 3000))
        shear = float(rng.uniform(0.01, 0.2)) if k % 2 else 0.0
        gen = soil_amd.gen_custom(9000 + k, n, 0.4, 0.6, float(rng.uniform(0.2, 0.55)), (1, 1, 1), 0, 0.0, 0.0,
                                  0.1, shear)
        jobs.append((gen, int(rng.integers(20, 90))))
    sv = Server(ctx, budget_gb=0.2, max_slots=3, chunk_steps=8)
    sv.start()
    ids = []
    for (d, pos, vel, mass, rad), steps in jobs:
        ids.append(sv.submit(d, pos, vel, mass, rad, steps, keep_result=True))
    sv.drain()
    st = sv.stats()
    print(st, flush=True)
    assert st["jobs_done"] == len(jobs) and st["jobs_failed"] == 0
    assert st["mean_active_jobs"] > 1.0  # jobs really overlapped on the device
    for jid, ((d, pos, vel, mass, rad), steps) in zip(ids, jobs):
        p_mix, v_mix = sv.result(jid, len(mass))
        b = soil_amd.Box(ctx, d)
        b.upload(pos, mass, rad, vel=vel)
        b.set_consumer(1000.0)
        b.setup()
        b.step(steps)
        a = b.atoms()
        n = len(mass)
        assert np.array_equal(p_mix, a["pos"][:n]), f"job {jid} positions differ"
        assert np.array_equal(v_mix, a["vel"][:n]), f"job {jid} velocities differ"
    sv.close()


def test_served_jobs_evict_resume_and_write_results(tmp_path):
    # finished jobs write RestartData; an evicted job's restart file resumes in
    # the server exactly like a box loaded from that file and stepped alone
    print("case: server evict / resume / output", flush=True)
    from soil_amd import files
    ctx = soil_amd.Context(0)
    rng = np.random.default_rng(78)
    jobs = []
    for k in range(4):
        n = int(rng.integers(800, 2000))
        gen = soil_amd.gen_custom(9100 + k, n, 0.4, 0.6, float(rng.uniform(0.25, 0.5)), (1, 1, 1), 0, 0.0, 0.0,
                                  0.1, 0.0)
        jobs.append((gen, int(rng.integers(120, 240))))
    out = tmp_path / "out"
    out.mkdir()
    sv = Server(ctx, budget_gb=0.2, max_slots=4, chunk_steps=8)
    sv.set_output(out, files.BINCODE)
    sv.start()
    ids = [sv.submit(d, pos, vel, mass, rad, steps, keep_result=True) for (d, pos, vel, mass, rad), steps in jobs]
    victim = ids[2]
    ev = tmp_path / "evicted.json"
    sv.evict(victim, ev, files.JSON)
    sv.drain()
    st = sv.stats()
    print(st, flush=True)
    assert st["jobs_evicted"] == 1 and st["jobs_failed"] == 0 and st["jobs_done"] == 3
    for jid, ((d, pos, vel, mass, rad), steps) in zip(ids, jobs):
        if jid == victim:
            continue
        p, v = sv.result(jid, len(mass))
        r = files.read_restart(out / f"restart_{steps}_job{jid}.bin", files.BINCODE)
        assert r["step"] == steps
        assert np.array_equal(r["pos"], p) and np.array_equal(r["vel"], v)
    sv.close()

    r = files.read_restart(ev, files.JSON)
    (d, pos, vel, mass, rad), steps = jobs[2]
    left = steps - int(r["step"])
    assert 0 < left <= steps
    sv = Server(ctx, budget_gb=0.2, max_slots=4, chunk_steps=8)
    sv.set_output(out, files.BINCODE)
    sv.start()
    jid = sv.submit_restart(d, ev, files.JSON, left, keep_result=True)
    sv.drain()
    p_mix, v_mix = sv.result(jid, len(mass))
    assert (out / f"restart_{steps}_job{jid}.bin").exists()
    sv.close()
    b = soil_amd.Box(ctx, d)
    assert files.box_read_restart(b, ev, files.JSON) == r["step"]
    b.set_consumer(1000.0)
    b.setup()
    b.step(left)
    a = b.atoms()
    n = len(mass)
    assert np.array_equal(p_mix, a["pos"][:n]) and np.array_equal(v_mix, a["vel"][:n])
    b.close()
