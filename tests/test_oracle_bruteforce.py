"""Pins the oracle's neighbor build, ghosts, bins and consumer by brute force."""
from collections import Counter

import numpy as np
import pytest

from bruteforce import brute_ghosts, brute_pairs, csr_pairs
from oracle_ref import RefBox, ref_desc


def random_box(seed, n, rmin, rmax, phi, aspect=(1, 1, 1)):
    rng = np.random.default_rng(seed)
    rad = rng.uniform(rmin, rmax, n) if rmax > rmin else np.full(n, rmin)
    V = np.sum(4.0 / 3.0 * np.pi * rad ** 3) / phi
    a = (V / np.prod(aspect)) ** (1.0 / 3.0)
    L = np.array(aspect, dtype=float) * a
    pos = rng.uniform(0, 1, (n, 3)) * L
    vel = rng.normal(0, 0.1, (n, 3))
    return L, pos, vel, rad


@pytest.mark.parametrize("seed,n,rmin,rmax,phi", [(1, 500, 0.5, 0.5, 0.3), (2, 700, 0.4, 0.6, 0.45),
                                                  (3, 400, 0.45, 0.55, 0.2)])
def test_half_list_equals_minimum_image_pairs(seed, n, rmin, rmax, phi):
    L, pos, vel, rad = random_box(seed, n, rmin, rmax, phi)
    b = RefBox(ref_desc([0, 0, 0], L, skin_fraction=1.1, dt=1e-3))
    b.add_atoms(pos, np.ones(n), rad, vel=vel)
    b.setup()
    b.step(1)
    a = b.atoms()
    off, idx = b.csr()
    keys = csr_pairs(off, idx, a, L)
    assert len(keys) == len(set(keys)), "a physical pair appears twice in the half list"
    nl = a["nlocal"]
    ref, amb = brute_pairs(a["pos"][:nl], a["radius"][:nl], a["tag"][:nl], L, 1.1)
    got = set(keys) - amb
    assert got == ref - amb


@pytest.mark.parametrize("xy", [0.7, -2.3, 4.9])
def test_triclinic_half_list_equals_minimum_image_pairs(xy):
    n = 600
    L, pos, vel, rad = random_box(11, n, 0.4, 0.6, 0.35)
    b = RefBox(ref_desc([0, 0, 0], L, skin_fraction=1.1, dt=1e-3))
    b.add_atoms(pos, np.ones(n), rad, vel=vel)
    b.setup()
    b.set_tilt(xy, 0.0)
    b.step(1)
    a = b.atoms()
    off, idx = b.csr()
    keys = csr_pairs(off, idx, a, L, xy)
    assert len(keys) == len(set(keys))
    nl = a["nlocal"]
    ref, amb = brute_pairs(a["pos"][:nl], a["radius"][:nl], a["tag"][:nl], L, 1.1, xy)
    assert set(keys) - amb == ref - amb


def test_ghost_images_match_swap_rules():
    n = 800
    L, pos, vel, rad = random_box(5, n, 0.4, 0.6, 0.4)
    b = RefBox(ref_desc([0, 0, 0], L, skin_fraction=1.1, dt=1e-3))
    b.add_atoms(pos, np.ones(n), rad, vel=vel)
    b.setup()
    b.step(1)
    a = b.atoms()
    nl, ng = a["nlocal"], a["nghost"]
    gc = b.grid()[9]
    want = brute_ghosts(a["pos"][:nl], [0, 0, 0], L, gc)
    origin = a["origin"]

    def root(j):
        while j >= nl:
            j = int(origin[j])
        return j

    got = []
    for g in range(nl, nl + ng):
        r = root(g)
        nv = np.round((a["pos"][g] - a["pos"][r]) / L).astype(int)
        got.append((r, *[int(v) for v in nv]))
        # ghost carries the owner's tag, radius, mass; inv_mass = 1/mass
        assert a["tag"][g] == a["tag"][r] and a["radius"][g] == a["radius"][r]
        assert a["inv_mass"][g] == 1.0 / a["mass"][g]
    assert len(got) == len(set(got)) == ng
    assert set(got) == want
    # scan order: within each swap, the sources are strictly increasing
    sw = b.swaps()
    for s in range(6):
        start, cnt = int(sw[s][0]), int(sw[s][1])
        src = origin[start:start + cnt]
        assert np.all(np.diff(src) > 0)


def test_bins_are_a_stable_counting_sort():
    n = 900
    L, pos, vel, rad = random_box(7, n, 0.4, 0.6, 0.5)
    b = RefBox(ref_desc([0, 0, 0], L, skin_fraction=1.1, dt=1e-3))
    b.add_atoms(pos, np.ones(n), rad, vel=vel)
    b.setup()
    b.step(1)
    bs, sa, si, cell = b.bins()
    a = b.atoms()
    nall = a["nlocal"] + a["nghost"]
    assert bs[-1] == nall
    assert np.array_equal(sa, np.argsort(cell, kind="stable"))
    assert np.array_equal(si[sa], np.arange(nall))
    # locals were spatially sorted: their cells are non-decreasing
    assert np.all(np.diff(cell[:a["nlocal"]].astype(np.int64)) >= 0)
    # the cell formula (neighbor.rs:1065-1076) recomputed from positions
    g = b.grid()
    nb = g[:3].astype(int)
    c3 = np.floor((a["pos"] - g[6:9]) * (1.0 / g[3:6])).astype(int)
    c3 = np.clip(c3, 0, nb - 1)
    assert np.array_equal(cell, c3[:, 0] * nb[1] * nb[2] + c3[:, 1] * nb[2] + c3[:, 2])


def test_sort_permutation_matches_numpy_stable_argsort():
    n = 500
    L, pos, vel, rad = random_box(9, n, 0.4, 0.6, 0.4)
    b = RefBox(ref_desc([0, 0, 0], L, skin_fraction=1.1, dt=1e-3))
    b.add_atoms(pos, np.ones(n), rad, vel=vel)
    b.setup()
    before = b.atoms()["pos"].copy()
    b.phase("pbc")
    b.phase("sort_atoms_by_bin")
    perm = b.sort_perm()
    g = b.grid()
    nb = g[:3].astype(int)
    c3 = np.clip(np.floor((before - g[6:9]) * (1.0 / g[3:6])).astype(int), 0, nb - 1)
    cell = c3[:, 0] * nb[1] * nb[2] + c3[:, 1] * nb[2] + c3[:, 2]
    assert np.array_equal(perm, np.argsort(cell, kind="stable"))
    assert np.array_equal(b.atoms()["pos"], before[perm])


def test_consumer_conserves_momentum_after_reverse_comm():
    n = 600
    L, pos, vel, rad = random_box(13, n, 0.4, 0.6, 0.5)
    b = RefBox(ref_desc([0, 0, 0], L, skin_fraction=1.1, dt=1e-3))
    b.add_atoms(pos, np.ones(n), rad, vel=vel)
    b.set_consumer(1000.0)
    b.setup()
    b.step(1)
    a = b.atoms()
    f = a["force"][:a["nlocal"]]
    assert np.abs(f).max() > 1.0
    assert np.all(np.abs(f.sum(axis=0)) < 1e-9 * np.abs(f).sum())


def test_rebuild_trigger_threshold():
    # neighbor.rs:693-758: no rebuild while every atom moved <= (sf-1)*min_r
    n = 300
    L, pos, vel, rad = random_box(17, n, 0.5, 0.5, 0.2)
    vel[:] = 0.0
    vel[0] = [1.0, 0.0, 0.0]
    b = RefBox(ref_desc([0, 0, 0], L, skin_fraction=1.1, dt=1e-2))
    b.add_atoms(pos, np.ones(n), rad, vel=vel)
    b.setup()
    b.step(1)  # first step is a FullRebuild (neighbor.rs:545)
    assert b.counts()["rebuilds"] == 1
    # atom 0 moves 0.01 per step; threshold 0.05 -> rebuild when displacement > 0.05 (6th step)
    b.step(5)
    assert b.counts()["rebuilds"] == 1
    b.step(1)
    assert b.counts()["rebuilds"] == 2
