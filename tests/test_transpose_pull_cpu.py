"""The pull transpose's traversal, restated on the CPU and checked against the oracle's CSR.

k_tpull_count / k_tpull_place (soil_amd/csrc/kernels.hip) find every incoming pair of a
target without atomics: row i holds target j exactly when j's cell lies in one of i's
windows (its own cell, or cells [ci + o, ci + o + len) of a stencil run (o, len); the
reference's row walk, neighbor.rs:1175-1304). So the rows that can hold a target in cells
[cl, ch] lie in cells [cl, ch] and [cl - o - len + 1, ch - o]; merged, those cell ranges are
row ranges of the cell-ordered locals, each a contiguous span of the CSR. A tile of 256
targets walks its local targets' merged ranges; a ghost target (ghosts are not in cell
order) walks its own cell's ranges, which are disjoint. This test restates that traversal
with the same range rules and requires, on boxes stepped by the oracle, that every atom's
pulled count equals its true incoming count (np.bincount of the CSR's column indices), that
no row is walked twice for a ghost, and that every row range stays inside the locals. It
checks the method, which the kernels follow; the device strategy is opt-in
(SOIL_HIP_TRANSPOSE=pull) and its device test is tests/test_gpu_transpose_pull.py.
"""
import numpy as np
import pytest

import soil_amd
from oracle_ref import RefBox

BLK = 256


def runs_of(stencil):
    runs = []
    for o in stencil:
        if runs and runs[-1][0] + runs[-1][1] == o:
            runs[-1][1] += 1
        else:
            runs.append([int(o), 1])
    return runs


def pulled_counts(r, shift=0):
    c = r.counts()
    nl, na = c["nlocal"], c["nlocal"] + c["nghost"]
    bs, _, _, ac = r.bins()
    ncell = bs.size - 1
    off, idx = r.csr()
    if r.desc.newton:
        stencil = r.stencil(True)
    else:
        stencil = np.array([o for o in r.stencil(False) if o != 0], dtype=np.int64)
    runs = runs_of(stencil)
    cell = ac.astype(np.int64)
    lc = cell[:nl]
    assert np.all(np.diff(lc) >= 0), "the locals are in cell order at a build"

    def row_start(cc):
        return int(np.searchsorted(lc, cc, side="left"))

    got = np.zeros(na, np.int64)
    for a0 in range(0, na, BLK):
        an = min(BLK, na - a0)
        nloc = min(nl - a0, an) if a0 < nl else 0
        if nloc > 0:
            cl, ch = int(cell[a0]), int(cell[a0 + nloc - 1])
            rg = []
            for (a, z) in [(cl, ch)] + [(cl - o - ln + 1 + shift, ch - o + shift) for o, ln in runs]:
                a, z = max(a, 0), min(z, ncell - 1)
                if a <= z:
                    rg.append([a, z])
            rg.sort()
            merged = []
            for a, z in rg:
                if merged and a <= merged[-1][1] + 1:
                    merged[-1][1] = max(merged[-1][1], z)
                else:
                    merged.append([a, z])
            assert len(merged) <= 64  # TP_MAXR
            for a, z in merged:
                r0, r1 = row_start(a), row_start(z + 1)
                assert 0 <= r0 <= r1 <= nl
                t = idx[int(off[r0]):int(off[r1])].astype(np.int64) - a0
                t = t[(t >= 0) & (t < nloc)]
                np.add.at(got, a0 + t, 1)
        for a in range(max(a0, nl), a0 + an):
            cg = int(cell[a])
            seen = set()
            for (lo, hi) in [(cg, cg)] + [(cg - o - ln + 1 + shift, cg - o + shift) for o, ln in runs]:
                lo, hi = max(lo, 0), min(hi, ncell - 1)
                if lo > hi:
                    continue
                r0, r1 = row_start(lo), row_start(hi + 1)
                rows = set(range(r0, r1))
                assert shift or not (rows & seen), f"ghost {a}: its cell's ranges overlap"
                seen |= rows
                got[a] += int(np.count_nonzero(idx[int(off[r0]):int(off[r1])] == a))
    true = np.bincount(idx.astype(np.int64), minlength=na)[:na]
    return got, true, idx.size


def oracle(gen, k=1000.0, desc_mod=None, pre=None):
    d, pos, vel, mass, rad = gen
    if desc_mod:
        desc_mod(d)
    r = RefBox(d)
    r.add_atoms(pos, mass, rad, vel=vel)
    r.set_consumer(k)
    if pre:
        pre(r)
    r.setup()
    return r


def small(seed, n, rmin, rmax, phi, aspect=(1, 1, 1), nblobs=0, blob_frac=0.0, sigma=0.0, shear=0.0):
    return soil_amd.gen_custom(seed, n, rmin, rmax, phi, aspect, nblobs, blob_frac, sigma, 0.1, shear)


def check_steps(r, steps, what):
    for s in range(1, steps + 1):
        r.step(1)
        got, true, npairs = pulled_counts(r)
        bad = np.nonzero(got != true)[0]
        assert bad.size == 0, f"{what} step {s}: {bad.size} targets differ, first {bad[:5]}"
        assert got.sum() == npairs


@pytest.mark.parametrize("case", ["dense", "sheared_clustered", "newton_off", "collapsing_axis"])
def test_pull_traversal_finds_every_incoming_pair(case):
    if case == "dense":
        check_steps(oracle(small(4901, 5000, 0.4, 0.6, 0.55)), 3, case)
    elif case == "sheared_clustered":
        check_steps(oracle(small(4902, 4000, 0.4, 0.6, 0.45, (2, 1, 1), 4, 0.5, 2.0, shear=0.2)), 3, case)
    elif case == "newton_off":
        def mod(d):
            d.newton = 0
        check_steps(oracle(small(4903, 3000, 0.4, 0.6, 0.45), desc_mod=mod), 2, case)
    else:
        # x shrinks by a tenth of its length per step, to a fifth: rows grow
        # eightfold and the thinner grid adds stencil runs
        d, pos, vel, mass, rad = small(4401, 500, 0.4, 0.6, 0.3)
        L = d.hi[0] - d.lo[0]

        def pre(b):
            b.set_axis_deform(0, "vel", -L / (10 * d.dt), (0, 0))
            b.set_deform_remap(True)
            b.set_stage_steps(100)
        check_steps(oracle((d, pos, vel, mass, rad), pre=pre), 9, case)


def test_pull_traversal_catches_a_planted_range_mistake():
    # the run ranges shifted by one cell (an off-by-one in [c - o - len + 1,
    # c - o]) must miss incoming pairs: the comparison is not vacuous
    r = oracle(small(4904, 2000, 0.4, 0.6, 0.45))
    r.step(1)
    got, true, _ = pulled_counts(r)
    assert np.array_equal(got, true)
    bad, true2, _ = pulled_counts(r, shift=1)
    assert np.array_equal(true2, true) and not np.array_equal(bad, true)
