"""GPU parity of the opt-in pull transpose (SOIL_HIP_TRANSPOSE=pull).

The harmonic consumer gathers every atom's incoming pairs through the
transpose, so its forces, compared bit for bit with the oracle, check that
each column holds exactly the atom's incoming pairs; a force hook that
gathers over the transpose checks the view handed to hooks. The cases cover
a dense box, a sheared clustered box, a box compressing until its rows grow
eightfold and its arenas regrow (the strategy then switches under
SOIL_HIP_TRANSPOSE=auto), a batch with the autotuner (which switches CSR
strategies, and with them the transpose launches), and the full list.
"""
import pytest

import soil_amd
from paritylib import assert_same, make_pair

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="opt-in strategy not yet run on the device with its two host-side fixes "
                            "(the batch's transpose launches recomputed after a CSR-strategy switch; the "
                            "strategy flag written only after the box's stream drains); two earlier device "
                            "runs without them faulted"),
]


@pytest.fixture(scope="module")
def ctx():
    return soil_amd.Context(0)


def small(seed, n, rmin, rmax, phi, aspect=(1, 1, 1), nblobs=0, blob_frac=0.0, sigma=0.0, shear=0.0):
    return soil_amd.gen_custom(seed, n, rmin, rmax, phi, aspect, nblobs, blob_frac, sigma, 0.1, shear)


def run(ctx, gen, nsteps, every, k=1000.0, desc_mod=None, pre_setup=None):
    d, pos, vel, mass, rad = gen
    if desc_mod:
        desc_mod(d)
    g, r = make_pair(ctx, d, pos, vel, mass, rad, k=k, pre_setup=pre_setup)
    done = 0
    while done < nsteps:
        s = min(every, nsteps - done)
        g.step(s)
        r.step(s)
        done += s
        assert_same(g, r, what=f"step {done}")


@pytest.mark.parametrize("case", ["dense", "sheared_clustered", "newton_off"])
def test_pull_transpose_matches_oracle(ctx, monkeypatch, case):
    print(f"case: pull transpose, {case}", flush=True)
    monkeypatch.setenv("SOIL_HIP_TRANSPOSE", "pull")
    if case == "dense":
        run(ctx, small(4921, 5000, 0.4, 0.6, 0.5), 20, 5)
    elif case == "sheared_clustered":
        run(ctx, small(4922, 4000, 0.4, 0.6, 0.45, (2, 1, 1), 4, 0.5, 2.0, shear=0.2), 20, 5)
    else:
        def mod(d):
            d.newton = 0
        run(ctx, small(4923, 3000, 0.4, 0.6, 0.45), 20, 5, desc_mod=mod)


def test_auto_transpose_switches_on_a_compressing_box(ctx, monkeypatch):
    print("case: auto transpose through compression and arena growth", flush=True)
    monkeypatch.setenv("SOIL_HIP_TRANSPOSE", "auto")
    d, pos, vel, mass, rad = small(4401, 500, 0.4, 0.6, 0.3)
    L = d.hi[0] - d.lo[0]

    def pre(b):
        b.set_axis_deform(0, "vel", -L / (10 * d.dt), (0, 0))
        b.set_deform_remap(True)
        b.set_stage_steps(100)
    run(ctx, (d, pos, vel, mass, rad), 9, 1, pre_setup=pre)


def test_pull_transpose_in_an_autotuned_batch(ctx, monkeypatch):
    print("case: pull transpose in an autotuned batch", flush=True)
    monkeypatch.setenv("SOIL_HIP_TRANSPOSE", "pull")
    pairs = [make_pair(ctx, *small(4930 + k, 1500 + 500 * k, 0.4, 0.6, 0.4), k=1000.0) for k in range(4)]
    batch = soil_amd.Batch(ctx, [g for g, _ in pairs])
    batch.autotune()
    batch.step(10)
    batch.sync()
    for k, (g, r) in enumerate(pairs):
        r.step(10)
        assert_same(g, r, what=f"box {k}")
