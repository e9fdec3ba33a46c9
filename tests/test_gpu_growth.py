"""Capacity growth of batched and served boxes, against the oracle.

A box whose ghosts or pairs outgrow their set-up capacity inside a batched
run freezes at that step; the host grows it, finishes the step and steps it
up to the rest of the batch. The results must be the oracle's, bit for bit.
Two ways to outgrow the set-up estimate (which assumes uniform density):
atoms placed only in slabs at the x faces (every one of them is ghosted at the
first build), and an x axis compressed step by step (the ghosts grow as the
box thins).
"""
import numpy as np
import pytest

import soil_amd
from oracle_ref import RefBox
from paritylib import assert_same, make_pair
from soil_amd.serve import Server

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return soil_amd.Context(0)


def slab_box(seed, n=2000, L=40.0, width=1.1):
    """n spheres r~U[0.45,0.55] with x within `width` of the two x faces."""
    rng = np.random.default_rng(seed)
    d = soil_amd.make_desc((0, 0, 0), (L, L, L), skin_fraction=1.1, bin_min_size=1.0, dt=1e-3)
    x = rng.uniform(0.0, width, n)
    x = np.where(rng.random(n) < 0.5, x, L - 1e-9 - x)
    pos = np.stack([x, rng.uniform(0, L, n), rng.uniform(0, L, n)], axis=1)
    vel = rng.normal(0.0, 0.1, (n, 3))
    return d, pos, vel, np.ones(n), rng.uniform(0.45, 0.55, n)


def compress_x(d, nsteps, frac=0.85):
    L = d.hi[0] - d.lo[0]

    def f(b):
        b.set_axis_deform(0, "vel", -frac * L / (nsteps * d.dt))
        b.set_stage_steps(nsteps)
    return f


@pytest.mark.parametrize("graph", [False, True])
def test_batched_boxes_outgrow_their_capacity(ctx, graph):
    print(f"case: batch growth graph={graph}", flush=True)
    gens = [slab_box(5100), soil_amd.gen_custom(5101, 1500, 0.4, 0.6, 0.3, (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.1),
            soil_amd.gen_custom(5102, 800, 0.4, 0.6, 0.3, (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.0),
            soil_amd.gen_custom(5103, 3000, 0.4, 0.6, 0.45, (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.05)]
    pairs = [make_pair(ctx, *g, k=1000.0) for g in gens[:3]]
    # the compressed box: its x axis thins by 85 % over 70 steps
    pairs.append(make_pair(ctx, *gens[3], k=1000.0, pre_setup=compress_x(gens[3][0], 70)))
    cap0 = [p[0].device_bytes()[1] for p in pairs]
    batch = soil_amd.Batch(ctx, [p[0] for p in pairs])
    # 40 steps (two runs), then 30 more: the second call first settles what
    # the first call's last run left (a mid-call settle at a half-ring wait)
    batch.step(40, graph=graph)
    batch.step(30, graph=graph)
    batch.sync()
    batch.close()
    for g, r in pairs:
        r.step(70)
    for i, (g, r) in enumerate(pairs):
        assert_same(g, r, what=f"box {i}")
    grew = [p[0].device_bytes()[1] > c for p, c in zip(pairs, cap0)]
    print(f"grew: {grew}", flush=True)
    assert grew[0] and grew[3]


def test_served_job_outgrows_its_capacity(ctx):
    print("case: served job growth", flush=True)
    jobs = [(slab_box(5200), 60), (slab_box(5201, n=1200, L=30.0), 45),
            (soil_amd.gen_custom(5202, 2000, 0.4, 0.6, 0.35, (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.1), 50)]
    sv = Server(ctx, budget_gb=0.5, max_slots=4, chunk_steps=8)
    sv.start()
    ids = [sv.submit(d, pos, vel, mass, rad, steps, keep_result=True) for (d, pos, vel, mass, rad), steps in jobs]
    sv.drain()
    st = sv.stats()
    print(st, flush=True)
    assert st["jobs_done"] == len(jobs) and st["jobs_failed"] == 0
    for jid, ((d, pos, vel, mass, rad), steps) in zip(ids, jobs):
        p_mix, v_mix = sv.result(jid, len(mass))
        r = RefBox(d)
        r.add_atoms(pos, mass, rad, vel=vel)
        r.set_consumer(1000.0)
        r.setup()
        r.step(steps)
        a = r.atoms()
        n = len(mass)
        assert np.array_equal(p_mix, a["pos"][:n]), f"job {jid} positions differ from the oracle"
        assert np.array_equal(v_mix, a["vel"][:n]), f"job {jid} velocities differ from the oracle"
    sv.close()
