"""The oracle against the rules the reference's text fixes (see pins.py):
forward comm, the AtomData column contract and the rebuild schedule. These
are the oracle functions that the KATs and brute-force checks leave open."""
import numpy as np
import pytest

import soil_amd
from oracle_ref import RefBox
from pins import (cells, chain_depth, chain_roots, check_forward, ghost_swaps, rebuild_steps_every,
                  sort_is_identity)

FWD, REV, ZERO = 1, 2, 4


def small_box(seed, n, phi, vel_sigma, shear=0.0, every=0, check=1, sort_every=1000, columns=()):
    d, pos, vel, mass, rad = soil_amd.gen_custom(seed, n, 0.45, 0.55, phi, (1, 1, 1), 0, 0.0, 0.0, vel_sigma, shear)
    d.every, d.check, d.sort_every = every, check, sort_every
    r = RefBox(d)
    for w, f in columns:
        r.register_column(w, f)
    r.add_atoms(pos, mass, rad, vel=vel)
    return r, d


def test_forward_comm_ghosts_are_source_plus_fixed_offset():
    # a small dense cube: many corner ghosts (three swaps deep); fast atoms
    # cross faces between builds, where only the build's fixed offset is right
    # (a sheared box rebuilds every step, so the tilted case is a static tilt
    # with a boundary velocity: tilted y shifts and velocity offsets)
    for tilt in (0.0, 1.7):
        r, d = small_box(7001, 1200, 0.35, 2.0)
        r.setup()
        if tilt:
            r.set_tilt(tilt, 0.3)
        r.step(1)
        comm_only = 0
        crossed = False
        for s in range(8):
            before = r.counts()["rebuilds"]
            r.step(1)
            if r.counts()["rebuilds"] != before:
                continue
            comm_only += 1
            a = r.atoms()
            depth = check_forward(a, r.swaps())
            assert depth.max() == 3
            nl = a["nlocal"]
            lo, hi = np.array(d.lo), np.array(d.hi)
            crossed |= bool(((a["pos"][:nl] < lo) | (a["pos"][:nl] >= hi)).any())
        assert comm_only >= 3, "too few CommunicateOnly steps to check"
        assert crossed, "no local crossed a face between builds"


def test_forward_comm_rejects_a_planted_offset_mistake():
    # the check is sensitive: recomputing a ghost from its source's wrapped
    # (current-image) position instead of the fixed offset breaks it
    r, d = small_box(7002, 1200, 0.35, 2.0)
    r.setup()
    r.step(3)
    a = r.atoms()
    check_forward(a, r.swaps())
    nl = a["nlocal"]
    L = np.array(d.hi) - np.array(d.lo)
    bad = dict(a)
    bad["pos"] = a["pos"].copy()
    sw = ghost_swaps(r.swaps(), nl, a["nghost"])
    g = nl + int(np.nonzero(sw == 0)[0][0])  # a ghost of the first swap: shift +Lx
    bad["pos"][g, 0] -= L[0]
    with pytest.raises(AssertionError):
        check_forward(bad, r.swaps())


def test_column_contract_forward_reverse_zero():
    cols = ((2, FWD), (1, REV), (1, ZERO), (1, REV | ZERO))
    r, d = small_box(7003, 1500, 0.4, 0.5, columns=cols)
    rng = np.random.default_rng(3)
    n = 1500
    # a #[forward] column tied to the atom (tag): it must follow every sort and reach the ghosts
    tag0 = np.arange(n, dtype=np.float64)
    r.set_column(0, np.stack([tag0, 2.0 * tag0 + 0.5], axis=1).ravel(), 2)
    r.set_column(1, np.zeros(n), 1)
    r.set_column(2, np.zeros(n), 1)
    r.set_column(3, np.zeros(n), 1)
    r.setup()
    r.step(6)
    a = r.atoms()
    nl, ng = a["nlocal"], a["nghost"]
    assert ng > 0 and r.counts()["rebuilds"] >= 1
    c0 = r.column(0).reshape(-1, 2)
    t = a["tag"].astype(np.float64)
    assert np.array_equal(c0[:, 0], t) and np.array_equal(c0[:, 1], 2.0 * t + 0.5), "forward column left its atom"

    # #[forward]: a CommunicateOnly refresh copies each ghost's source value (chained)
    vals = rng.normal(size=(nl + ng, 2))
    vals[nl:] = 7.0  # stale ghost values, overwritten by the refresh
    r.set_column(0, vals.ravel(), 2)
    r.phase("forward_comm_borders")
    c0 = r.column(0).reshape(-1, 2)
    root = chain_roots(a["origin"], nl)
    assert np.array_equal(c0[nl:], vals[root[nl:]]), "forward column is not the source's value"
    assert np.array_equal(c0[:nl], vals[:nl])

    # #[reverse]: integer-valued ghost entries (exact in any order) land on
    # their roots, chain by chain: owner = own + sum over every ghost rooted at it
    rv = np.zeros(nl + ng)
    rv[:nl] = rng.integers(0, 1000, nl)
    rv[nl:] = rng.integers(1, 1 << 20, ng)
    r.set_column(1, rv, 1)
    r.set_column(3, rv, 1)
    r.phase("reverse_send_force")
    want = rv[:nl].copy()
    np.add.at(want, root[nl:], rv[nl:])
    for c in (1, 3):
        got = r.column(c)
        assert np.array_equal(got[:nl], want), f"reverse column {c}"
    # a ghost that is the source of others also received their values first
    depth = chain_depth(a["origin"], nl)
    assert depth.max() >= 2

    # #[zero]: zero_all_forces clears the zero columns on every atom and leaves the others
    zv = rng.normal(size=nl + ng)
    r.set_column(2, zv, 1)
    r.set_column(3, zv, 1)
    keep0, keep1 = r.column(0).copy(), r.column(1).copy()
    r.phase("zero_all_forces")
    assert not r.column(2).any() and not r.column(3).any()
    assert np.array_equal(r.column(0), keep0) and np.array_equal(r.column(1), keep1)
    assert not r.atoms()["force"].any()


@pytest.mark.parametrize("every", [1, 2, 5])
def test_every_without_check_rebuilds_on_the_schedule(every):
    # check off: the displacements (fast atoms) must not trigger anything
    nsteps = 3 * (every + 1) + 2
    r, _ = small_box(7010 + every, 1000, 0.3, 3.0, every=every, check=0, sort_every=0)
    r.setup()
    got = []
    for s in range(1, nsteps + 1):
        before = r.counts()["rebuilds"]
        r.step(1)
        if r.counts()["rebuilds"] != before:
            got.append(s)
    assert got == rebuild_steps_every(every, nsteps)


def test_every_with_check_adds_displacement_rebuilds():
    # check on: the same schedule, plus rebuilds wherever an atom moved past
    # the threshold ((sf - 1) min_r)^2 since the last build
    r, _ = small_box(7020, 1000, 0.3, 3.0, every=4, check=1, sort_every=0)
    r.setup()
    got = []
    for s in range(1, 40):
        before = r.counts()["rebuilds"]
        r.step(1)
        if r.counts()["rebuilds"] != before:
            got.append(s)
    gaps = np.diff(got)
    assert got[0] == 1 and gaps.max() <= 5  # never longer than `every` + 1
    assert gaps.min() < 5                   # displacement rebuilds come earlier


@pytest.mark.parametrize("sigma", [0.0, 0.6])
def test_periodic_sort_forces_a_rebuild_only_for_a_moved_order(sigma):
    # no forces (no consumer) and few steps: no displacement trigger, so only
    # the periodic sort can rebuild. Its stable cell order is computed here
    # from the positions the step's kick-drift will produce (x + v dt: v is
    # constant without forces).
    S = 5
    r, d = small_box(7030, 2000, 0.3, sigma, every=0, check=1, sort_every=S)
    r.setup()
    r.step(1)
    sorts = []
    for s in range(2, 4 * S + 1):
        a = r.atoms()
        nl = a["nlocal"]
        x = a["pos"][:nl] + a["vel"][:nl] * d.dt
        before = r.counts()["rebuilds"]
        r.step(1)
        rebuilt = r.counts()["rebuilds"] != before
        if s % S == 0:
            ident = sort_is_identity(x, r.grid())
            assert rebuilt == (not ident), (s, ident, rebuilt)
            sorts.append(ident)
        else:
            assert not rebuilt, f"step {s} rebuilt without a trigger"
    if sigma == 0.0:
        assert all(sorts)
    else:
        assert not all(sorts)


def test_cells_match_the_oracle_bins():
    # the cell formula used above is the oracle's own binning (bins pin it to a stable counting sort)
    r, _ = small_box(7040, 1500, 0.35, 0.5)
    r.setup()
    r.step(2)
    a = r.atoms()
    _, _, _, atom_cell = r.bins()
    assert np.array_equal(cells(a["pos"], r.grid()), atom_cell.astype(np.int64))


def test_removal_order_closed_form_equals_the_sequential_loop():
    # the device's one-pass form of the descending swap_remove loop, on random
    # removal sets of every density (and the empty and full sets)
    from pins import removal_order, removal_order_sequential
    rng = np.random.default_rng(11)
    for n in (1, 2, 5, 40, 300):
        for frac in (0.0, 0.05, 0.3, 0.7, 1.0):
            for _ in range(5):
                rem = rng.random(n) < frac
                assert np.array_equal(removal_order(rem), removal_order_sequential(rem)), (n, frac)


@pytest.mark.parametrize("tilt", [0.0, 1.3])
def test_oracle_pbc_removes_atoms_outside_fixed_faces(tilt):
    # atoms placed outside the fixed z faces leave at the first pbc: the
    # survivors are exactly the inside atoms, in the swap_remove order, with
    # every column (tag, pos, radius, AtomData) carried along
    from pins import removal_order
    d, pos, vel, mass, rad = soil_amd.gen_custom(7050, 1500, 0.45, 0.55, 0.3, (1, 1, 1), 0, 0.0, 0.0, 0.0, 0.0)
    rng = np.random.default_rng(5)
    L = d.hi[2] - d.lo[2]
    out = rng.random(len(pos)) < 0.1
    pos[out, 2] = np.where(rng.random(out.sum()) < 0.5, d.lo[2] - 0.3, d.hi[2] + 0.2)
    r = RefBox(d)
    r.register_column(1, FWD)
    r.add_atoms(pos, mass, rad, vel=vel)
    r.set_column(0, np.arange(len(pos), dtype=np.float64) * 3.0, 1)
    r.set_boundary(2, "f")
    r.setup()
    if tilt:
        r.set_tilt(tilt, 0.0)
    for ph in ("apply_deform", "initial_integration", "remove_ghost_atoms", "zero_all_forces"):
        r.phase(ph)
    a0 = r.atoms()
    z = a0["pos"][:, 2]
    gone = (z < d.lo[2]) | (z >= d.hi[2])
    assert np.array_equal(gone, out)
    r.phase("pbc")
    a1 = r.atoms()
    order = removal_order(gone)
    assert a1["nlocal"] == len(pos) - out.sum() == len(order)
    assert np.array_equal(a1["tag"], a0["tag"][order])
    assert np.array_equal(a1["radius"], a0["radius"][order])
    assert np.array_equal(r.column(0), a0["tag"][order] * 3.0)
    assert r.state == 0  # FullRebuild


def csr_case(case):
    """Oracle boxes for the row-order rule: uniform and polydisperse radii
    (the two accept paths), the full list, and a tilted grid."""
    if case == "uniform":
        g = soil_amd.gen_custom(7060, 1500, 0.5, 0.5, 0.45, (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.0)
    elif case == "tilted":
        g = soil_amd.gen_custom(7061, 1500, 0.4, 0.6, 0.4, (2, 1, 1), 3, 0.5, 2.0, 0.1, 0.2)
    else:
        g = soil_amd.gen_custom(7062, 1500, 0.4, 0.6, 0.45, (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.0)
    d, pos, vel, mass, rad = g
    if case == "full":
        d.newton = 0
    r = RefBox(d)
    r.add_atoms(pos, mass, rad, vel=vel)
    r.setup()
    r.step(2)
    return r, bool(d.newton)


def rows_stencil(r, newton):
    return r.stencil(True) if newton else np.array([o for o in r.stencil(False) if o != 0], dtype=np.int64)


@pytest.mark.parametrize("case", ["uniform", "poly", "full", "tilted"])
def test_csr_rows_follow_the_reference_loop_order(case):
    from pins import check_csr_order
    r, newton = csr_case(case)
    off, idx = r.csr()
    assert idx.size > 1000
    check_csr_order(off, idx, r.bins(), rows_stencil(r, newton), r.grid()[0:3], newton)


def test_csr_order_rule_catches_planted_mistakes():
    # the rule is not vacuous: the same rows with two entries of a row
    # swapped, a stencil walked in the reverse of the generation loops, or the
    # own cell walked last, must fail
    from pins import check_csr_order, check_stencil_order
    r, newton = csr_case("poly")
    off, idx = r.csr()
    bins, nbin, st = r.bins(), r.grid()[0:3], rows_stencil(r, newton)
    long_row = int(np.nonzero(np.diff(off) >= 3)[0][0])
    bad = idx.copy()
    a = int(off[long_row])
    bad[a], bad[a + 1] = bad[a + 1], bad[a]
    with pytest.raises(AssertionError):
        check_csr_order(off, bad, bins, st, nbin, newton)
    # a forward stencil in generation order (reach 2 in z), then reversed
    gen = [o for o in (dx * 25 + dy * 5 + dz for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-2, -1, 0, 1, 2))
           if o > 0]
    check_stencil_order(gen, (3, 5, 5), True)
    with pytest.raises(AssertionError):
        check_stencil_order(gen[::-1], (3, 5, 5), True)
    with pytest.raises(AssertionError):
        check_csr_order(off, idx, bins, st[::-1], nbin, newton)
    # own cell last: move each row's self-cell entries behind its stencil entries
    _, _, _, ac = (np.asarray(x, dtype=np.int64) for x in bins)
    moved = idx.copy()
    for i in range(len(off) - 1):
        e = idx[off[i]:off[i + 1]].astype(np.int64)
        own = ac[e] == ac[i]
        moved[off[i]:off[i + 1]] = np.concatenate([e[~own], e[own]])
    assert not np.array_equal(moved, idx)
    with pytest.raises(AssertionError):
        check_csr_order(off, moved, bins, st, nbin, newton)


def cutoff_desc():
    d = soil_amd._ffi.BoxDesc()
    for k in range(3):
        d.lo[k], d.hi[k], d.periodic[k] = 0.0, 20.0, 1
    d.newton, d.skin_fraction, d.bin_min_size = 1, 1.2, 0.0
    d.every, d.check, d.sort_every, d.dt = 0, 1, 1000, 1e-3
    return d


def test_cutoff_forms_decide_a_pair_between_their_roundings():
    # one pair sits between the uniform form's c*c and the per-pair form: it
    # is listed exactly when the locals' radii span 1e-15 or more (a third,
    # distant atom decides), and the uniform form takes the smallest radius
    from pins import CUT_POS, CUT_R, CUT_SF, cutoff_cases
    cases = cutoff_cases()
    assert [c[1] for c in cases] == [False, False, True, True]  # the rule's verdicts, as laid out in pins.py
    for r3, listed in cases:
        d = cutoff_desc()
        assert d.skin_fraction == CUT_SF
        r = RefBox(d)
        r.add_atoms(CUT_POS.copy(), np.ones(3), np.array([CUT_R, CUT_R, r3]), vel=np.zeros((3, 3)))
        r.setup()
        r.step(1)
        off, idx = r.csr()
        a = r.atoms()
        tags = a["tag"]
        pairs = {tuple(sorted((int(tags[i]), int(tags[j]))))
                 for i in range(len(off) - 1) for j in idx[off[i]:off[i + 1]]}
        assert pairs == ({(0, 1)} if listed else set()), (r3, pairs)


def fma_box_csr(make, pj):
    """CSR pairs (by tag) of a 20-cube holding the two atoms of the fma pin
    and a third, distant one; `make(desc, pos, radii)` -> stepped box."""
    from pins import FMA_PI, FMA_R, FMA_SF
    d = cutoff_desc()
    d.skin_fraction = FMA_SF
    pos = np.array([FMA_PI, pj, (10.0, 10.0, 10.0)])
    b = make(d, pos, np.full(3, FMA_R))
    off, idx = b.csr()
    tags = b.atoms()["tag"]
    return {tuple(sorted((int(tags[i]), int(tags[j])))) for i in range(len(off) - 1) for j in idx[off[i]:off[i + 1]]}


def test_distance_uses_the_reference_fma_nesting():
    # two pairs whose listing depends on how the squared distance rounds:
    # the reference's nesting lists the second and not the first, and each
    # plausible misreading flips one of them
    from pins import FMA_PI, FMA_PJ, FMA_R, FMA_SF, r2_alternatives, r2_reference
    c = 2.0 * FMA_R * FMA_SF
    cut = c * c
    verdicts = []
    for pj in FMA_PJ:
        dxyz = [pj[k] - FMA_PI[k] for k in range(3)]
        assert all(dxyz[k] + FMA_PI[k] == pj[k] for k in range(3))  # exact differences
        listed = r2_reference(*dxyz) < cut
        assert all((alt < cut) != listed for alt in r2_alternatives(*dxyz))
        verdicts.append(listed)
    assert verdicts == [False, True]

    def make(d, pos, rad):
        r = RefBox(d)
        r.add_atoms(pos, np.ones(3), rad, vel=np.zeros((3, 3)))
        r.setup()
        r.step(1)
        return r

    for pj, listed in zip(FMA_PJ, verdicts):
        assert fma_box_csr(make, pj) == ({(0, 1)} if listed else set()), pj


def verlet_expected():
    from pins import VERLET_ATOM as A, VERLET_DT, verlet_alternatives, verlet_reference
    ref = [verlet_reference(A["x"][k], A["v"][k], A["f"][k], A["m"], VERLET_DT) for k in range(3)]
    alts = [verlet_alternatives(A["x"][k], A["v"][k], A["f"][k], A["m"], VERLET_DT) for k in range(3)]
    for a in range(3):  # every misreading changes some component: the case discriminates
        assert any(alts[k][a] != ref[k] for k in range(3)), a
    return np.array([r[0] for r in ref]), np.array([r[1] for r in ref])


def test_verlet_kick_and_drift_round_as_the_reference():
    # one atom with an uploaded force: after the first step (whose final
    # half-kick sees no force) its velocity and position are the reference
    # expression's, bit for bit
    from pins import VERLET_ATOM as A, VERLET_DT
    v1, x1 = verlet_expected()
    d = cutoff_desc()
    d.dt = VERLET_DT
    r = RefBox(d)
    r.add_atoms(np.array([A["x"]]), np.array([A["m"]]), np.array([0.5]), vel=np.array([A["v"]]),
                force=np.array([A["f"]]))
    r.setup()
    r.step(1)
    a = r.atoms()
    assert np.array_equal(a["vel"][0], v1) and np.array_equal(a["pos"][0], x1)


def test_cell_uses_the_inverse_bin_size():
    # atoms on cell faces where the product by the inverse and a division
    # disagree: the oracle bins them by the product (and the case is one
    # where the division would move them)
    from pins import CELL_POS, cells_by_division
    d = cutoff_desc()
    d.skin_fraction = 1.1
    r = RefBox(d)
    r.add_atoms(CELL_POS.copy(), np.ones(3), np.full(3, 0.5), vel=np.zeros((3, 3)))
    r.setup()
    r.step(1)
    a, g = r.atoms(), r.grid()
    _, _, _, atom_cell = r.bins()
    pos = a["pos"][:3]
    assert np.array_equal(pos[np.argsort(a["tag"][:3])], CELL_POS)  # nothing moved
    want = cells(pos, g)
    assert (cells_by_division(CELL_POS[:2], g) != cells(CELL_POS[:2], g)).all()
    assert np.array_equal(atom_cell[:3].astype(np.int64), want)


def disp_desc():
    d = cutoff_desc()
    d.skin_fraction = 1.1
    return d


def test_displacement_trigger_fuses_only_in_fully_periodic_boxes():
    # an atom whose step-2 displacement sits between the fused and the plain
    # squared length: a fully periodic box decides by the fused one, a box
    # with a fixed z by the plain one (two cases, one each way)
    from pins import DISP_CASES, DISP_DT, DISP_R, disp_decisions
    seen = set()
    for x0, v in DISP_CASES:
        want = disp_decisions(x0, v)
        assert want[0] != want[1]  # the case separates the two forms
        seen.add(want)
        for periodic_z, rebuild in zip((True, False), want):
            d = disp_desc()
            d.dt = DISP_DT
            r = RefBox(d)
            if not periodic_z:
                r.set_boundary(2, "f")
            r.add_atoms(np.array([x0]), np.ones(1), np.array([DISP_R]), vel=np.array([v]))
            r.setup()
            r.step(1)
            n1 = r.counts()["rebuilds"]
            r.step(1)
            assert r.counts()["rebuilds"] - n1 == (1 if rebuild else 0), (x0, periodic_z)
    assert seen == {(True, False), (False, True)}


def test_periodic_wrap_bound_is_low_plus_size():
    # an atom at x = hi stays (hi < lo + size), an atom at y = lo + size < hi
    # wraps: the bound is computed as low + size, not read from high
    from pins import WRAP_HI, WRAP_LO, wrap_case
    pos, want_pos, want_img = wrap_case()
    d = cutoff_desc()
    d.skin_fraction = 1.1
    for k in range(3):
        d.lo[k], d.hi[k] = WRAP_LO[k], WRAP_HI[k]
    r = RefBox(d)
    r.add_atoms(pos, np.ones(2), np.full(2, 0.5), vel=np.zeros((2, 3)))
    r.setup()
    r.step(1)
    a = r.atoms()
    o = np.argsort(a["tag"][:2])
    assert np.array_equal(a["pos"][:2][o], want_pos)
    assert np.array_equal(a["image"][:2][o], want_img)


def test_ghost_faces_use_rounded_thresholds():
    # two atoms on the face thresholds of a box whose x bounds make them
    # round: both are sent (low and high swap of x), where the difference
    # form would send neither
    from pins import FACE_HI, FACE_LO, FACE_X, face_sends, sent_by_swap
    d = cutoff_desc()
    d.skin_fraction = 1.1
    for k in range(3):
        d.lo[k], d.hi[k] = FACE_LO[k], FACE_HI[k]
    r = RefBox(d)
    pos = np.array([[FACE_X[0], 10.0, 10.0], [FACE_X[1], 10.0, 10.0]])
    r.add_atoms(pos, np.ones(2), np.full(2, 0.5), vel=np.zeros((2, 3)))
    r.setup()
    r.step(1)
    ref, alt = face_sends(r.grid()[9])
    assert ref == (True, True) and alt == (False, False)
    assert sent_by_swap(r.atoms(), r.swaps()) == {0: {0}, 1: {1}}


def tilt_desc():
    d = cutoff_desc()
    d.skin_fraction, d.has_shear, d.shear_rate = 1.1, 1, 0.112785
    return d


def test_lees_edwards_tilt_product_order():
    # the tilt after 9 steps equals the left-to-right product, which every
    # regrouping of it rounds differently
    from pins import TILT_RATE, TILT_STEPS, tilt_reference, tilt_regroupings
    d = tilt_desc()
    assert d.shear_rate == TILT_RATE
    want = tilt_reference(TILT_RATE, d.dt, TILT_STEPS, 20.0, 20.0)
    assert all(alt != want for alt in tilt_regroupings(TILT_RATE, d.dt, TILT_STEPS, 20.0))
    r = RefBox(d)
    r.add_atoms(np.array([[5.0, 5.0, 5.0], [12.0, 14.0, 9.0]]), np.ones(2), np.full(2, 0.5), vel=np.zeros((2, 3)))
    r.setup()
    r.step(TILT_STEPS)
    assert r.grid()[12] == want


def tri_expected():
    from pins import TRI_POS, tri_pbc
    want = [tri_pbc(list(p)) for p in TRI_POS]
    # the case separates the misreadings: A moves only by the round trip, B
    # wraps only under the regrouped x2lamda
    assert want[0][0] != list(TRI_POS[0]) and tri_pbc(list(TRI_POS[0]), keep=True)[0] == list(TRI_POS[0])
    assert want[1][1] == [0, 0, 0] and tri_pbc(list(TRI_POS[1]), regrouped=True)[1] == [1, 0, 0]
    return np.array([w[0] for w in want]), np.array([w[1] for w in want], dtype=np.int32)


def test_triclinic_wrap_round_trips_every_local():
    from pins import TRI_BVEL, TRI_HI, TRI_LO, TRI_POS, TRI_XY
    want_pos, want_img = tri_expected()
    d = cutoff_desc()
    d.skin_fraction = 1.1
    for k in range(3):
        d.lo[k], d.hi[k] = TRI_LO[k], TRI_HI[k]
    r = RefBox(d)
    r.add_atoms(TRI_POS.copy(), np.ones(2), np.full(2, 0.5), vel=np.zeros((2, 3)))
    r.setup()
    r.set_tilt(TRI_XY, TRI_BVEL)
    r.step(1)
    a = r.atoms()
    o = np.argsort(a["tag"][:2])
    assert np.array_equal(a["pos"][:2][o], want_pos)
    assert np.array_equal(a["image"][:2][o], want_img)


def test_reverse_comm_adds_in_the_reference_order():
    # inexact #[reverse] values on a box thin in z (its two z selections
    # overlap, so some owners receive from both z swaps) and with chained
    # corner ghosts: the owners' sums equal the reference's sequence of adds,
    # which the high-swap-first order and a per-root sum both miss
    from pins import chain_roots, reverse_order, reverse_sequential
    d, pos, vel, mass, rad = soil_amd.gen_custom(7070, 700, 0.45, 0.55, 0.4, (1, 1, 0.07), 0, 0.0, 0.0, 0.3, 0.0)
    Lz = d.hi[2] - d.lo[2]
    r = RefBox(d)
    r.register_column(1, REV)
    r.add_atoms(pos, mass, rad, vel=vel)
    r.set_column(0, np.zeros(len(pos)), 1)
    r.setup()
    r.step(2)
    a, sw = r.atoms(), r.swaps()
    nl, ng = a["nlocal"], a["nghost"]
    gc = r.grid()[9]
    assert gc < Lz < 2 * gc, (gc, Lz)  # the z selections overlap and need one image each
    rng = np.random.default_rng(9)
    vals = rng.normal(size=nl + ng) * 10.0 ** rng.integers(-6, 7, nl + ng)
    r.set_column(0, vals, 1)
    r.phase("reverse_send_force")
    got = r.column(0)
    want = reverse_sequential(vals, sw, a["origin"])
    assert np.array_equal(got, want)
    assert reverse_order(6) == [4, 5, 2, 3, 0, 1]
    hi_first = reverse_sequential(vals, sw, a["origin"], order=list(range(len(sw) - 1, -1, -1)))
    per_root = vals[:nl].copy()
    np.add.at(per_root, chain_roots(a["origin"], nl)[nl:], vals[nl:])
    assert not np.array_equal(hi_first[:nl], want[:nl]) and not np.array_equal(per_root, want[:nl])


def test_periodic_sort_step_rebuilds_from_unwrapped_positions():
    # step 3 is CommunicateOnly until its periodic sort finds A's new cell
    # out of order: it rebuilds, but pbc already skipped, so A stays past the
    # face with its image unchanged (and on step 4, no rebuild, still there)
    from pins import PSORT_EVERY, PSORT_POS, PSORT_VEL, psort_x_after
    d = cutoff_desc()
    d.skin_fraction, d.sort_every = 1.1, PSORT_EVERY
    r = RefBox(d)
    r.add_atoms(PSORT_POS.copy(), np.ones(2), np.full(2, 0.5), vel=PSORT_VEL.copy())
    r.setup()
    r.step(2)
    assert r.counts()["rebuilds"] == 1
    for s, rebuilt in ((3, 2), (4, 2)):
        r.step(1)
        assert r.counts()["rebuilds"] == rebuilt, s
        a = r.atoms()
        k = int(np.nonzero(a["tag"][:2] == 0)[0][0])
        assert a["pos"][k, 0] == psort_x_after(s) and a["pos"][k, 0] >= 20.0, s
        assert a["image"][k, 0] == 0
        if s == 3:
            assert k == 1  # the sort put A after B


def test_streaming_profile_takes_the_difference_first():
    # the first sheared step gives a resting atom v_x = rate * (y - yc),
    # which the expanded form rate * y - rate * yc would round differently
    from pins import PROFILE_Y, TILT_RATE, profile_vx
    want = profile_vx(TILT_RATE, PROFILE_Y)
    assert want != TILT_RATE * PROFILE_Y - TILT_RATE * 10.0
    r = RefBox(tilt_desc())
    r.add_atoms(np.array([[10.0, PROFILE_Y, 10.0]]), np.ones(1), np.full(1, 0.5), vel=np.zeros((1, 3)))
    r.setup()
    r.step(1)
    assert r.atoms()["vel"][0, 0] == want


def test_triclinic_ghost_face_divides_before_inflating():
    from pins import TRIF_POS, TRIF_XY, sent_by_swap, trif_lx, trif_thresholds
    d = cutoff_desc()
    d.skin_fraction = 1.1
    r = RefBox(d)
    r.add_atoms(TRIF_POS.copy(), np.ones(1), np.full(1, 0.5), vel=np.zeros((1, 3)))
    r.setup()
    r.set_tilt(TRIF_XY, 0.0)
    r.step(1)
    a = r.atoms()
    t_ref, t_alt = trif_thresholds(r.grid()[9])
    lx = trif_lx(a["pos"][0])
    assert lx < t_ref and not lx < t_alt  # the case separates the two forms
    assert sent_by_swap(a, r.swaps()).get(0) == {0}
