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
