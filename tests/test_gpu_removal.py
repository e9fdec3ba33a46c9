"""Atoms leaving through a non-periodic face (pbc removal, domain.rs:633-661,
700-706), on the device against the oracle: the survivors, their order (the
reference's descending swap_remove) and every column, bit for bit."""
import numpy as np
import pytest

import soil_amd
from paritylib import assert_same, make_pair

pytestmark = pytest.mark.gpu

FWD = soil_amd._ffi.COL_FORWARD


@pytest.fixture(scope="module")
def ctx():
    return soil_amd.Context(0)


def walled(seed, n, shear=0.0, outward=3.0):
    """A box whose z faces are fixed walls; atoms within 0.08 of a z face move
    outward at `outward`, so they leave over the next steps."""
    d, pos, vel, mass, rad = soil_amd.gen_custom(seed, n, 0.4, 0.6, 0.35, (1, 1, 1), 0, 0.0, 0.0, 0.1, shear)
    lo, hi = d.lo[2], d.hi[2]
    near_lo = pos[:, 2] < lo + 0.08
    near_hi = pos[:, 2] >= hi - 0.08
    vel[near_lo, 2] = -outward
    vel[near_hi, 2] = outward
    return d, pos, vel, mass, rad


def wall_z(b):
    b.set_boundary(2, "f")


@pytest.mark.parametrize("shear", [0.0, 0.12])
def test_atoms_leave_through_fixed_walls(ctx, shear):
    print(f"case: removal through fixed z walls, shear={shear}", flush=True)
    gen = walled(9301 + int(shear * 100), 3000, shear)
    n = len(gen[1])
    tags = np.arange(n, dtype=np.float64)
    g, r = make_pair(ctx, *gen, k=1000.0, columns=((2, FWD),),
                     col_data=(np.stack([tags, tags + 0.25], axis=1).ravel(),), pre_setup=wall_z)
    for s in range(6):
        g.step(5)
        r.step(5)
        assert_same(g, r, what=f"removal step {5 * (s + 1)}")
        assert np.array_equal(g.column(0), r.column(0))
    left = n - g.counts()["nlocal"]
    print(f"{left} atoms left", flush=True)
    assert left > 20


def test_removal_phase_by_phase(ctx):
    # the pbc system alone removes, like the reference's, and sets FullRebuild
    print("case: removal per phase", flush=True)
    d, pos, vel, mass, rad = walled(9310, 2000)
    rng = np.random.default_rng(2)
    out = rng.random(len(pos)) < 0.08
    pos[out, 2] = np.where(rng.random(out.sum()) < 0.5, d.lo[2] - 0.2, d.hi[2] + 0.1)
    g, r = make_pair(ctx, d, pos, vel, mass, rad, k=1000.0, pre_setup=wall_z)
    for step in range(3):
        g.apply_deform(); r.phase("apply_deform")
        g.initial_integration(); r.phase("initial_integration")
        if step > 0:
            fr = g.decide_rebuild(); r.phase("decide_rebuild")
            assert fr == (r.state == 0)
        if r.state == 0:
            g.remove_ghost_atoms(); r.phase("remove_ghost_atoms")
        g.zero_all_forces(); r.phase("zero_all_forces")
        if r.state == 0:
            g.pbc(); r.phase("pbc")
            assert g.counts()["nlocal"] == r.counts()["nlocal"]
        g.sort_atoms_by_bin(); r.phase("sort_atoms_by_bin")
        if r.state == 0:
            g.borders(); r.phase("borders")
            g.bin_neighbor_list(); r.phase("bin_neighbor_list")
        else:
            g.forward_comm_borders(); r.phase("forward_comm_borders")
        g.harmonic_force(); r.phase("harmonic_force")
        g.reverse_send_force(); r.phase("reverse_send_force")
        g.final_integration(); r.phase("final_integration")
        ag, ar = g.atoms(), r.atoms()
        for key in ("tag", "pos", "vel", "force"):
            assert np.array_equal(ag[key], ar[key]), f"step {step}: {key}"
    assert g.counts()["nlocal"] == r.counts()["nlocal"] <= len(pos) - out.sum()


def test_removal_in_a_batch(ctx):
    # a walled box among periodic ones: its pbc removes inside the batched step
    print("case: removal in a batch", flush=True)
    gens = [walled(9320, 2500), soil_amd.gen_custom(9321, 1500, 0.4, 0.6, 0.3, (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.1),
            walled(9322, 1200, shear=0.05)]
    pairs = [make_pair(ctx, *gens[0], k=1000.0, pre_setup=wall_z), make_pair(ctx, *gens[1], k=1000.0),
             make_pair(ctx, *gens[2], k=1000.0, pre_setup=wall_z)]
    batch = soil_amd.Batch(ctx, [p[0] for p in pairs])
    batch.step(25, graph=True)
    batch.sync()
    batch.close()
    for i, (g, r) in enumerate(pairs):
        r.step(25)
        assert_same(g, r, what=f"batched box {i}")
    assert pairs[0][0].counts()["nlocal"] < len(gens[0][1])


def test_every_atom_leaves(ctx):
    # a box that empties: the steps after go on with no atoms, as the reference's
    print("case: box emptied through its walls", flush=True)
    d, pos, vel, mass, rad = soil_amd.gen_custom(9330, 40, 0.4, 0.6, 0.2, (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.0)
    vel[:, 2] = np.where(pos[:, 2] < 0.5 * (d.lo[2] + d.hi[2]), -4000.0, 4000.0)
    g, r = make_pair(ctx, d, pos, vel, mass, rad, pre_setup=wall_z)
    for s in range(4):
        g.step(1)
        r.step(1)
        assert_same(g, r, what=f"emptying step {s + 1}")
    assert g.counts()["nlocal"] == 0
