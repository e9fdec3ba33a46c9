"""GPU parity: the HIP path against the CPU oracle on the same seeded inputs.

Integer outputs (bins, ghost lists, CSR offsets and indices, swap tables,
rebuild counts) must be bit-identical at every compared step. The state
(positions, velocities, forces) is compared bit for bit as well: the substrate
is f64 add/mul/div in the reference's order, and the harmonic consumer sums
2^-32 fixed-point pair forces exactly, so its result does not depend on the
accumulation order. The stated tolerance for state is 1e-12 relative; these
cases meet it with zero difference.
"""
import sys

import numpy as np
import pytest

import soil_amd
from paritylib import assert_same, make_pair

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return soil_amd.Context(0)


def small(seed, n, rmin, rmax, phi, aspect=(1, 1, 1), nblobs=0, blob_frac=0.0, sigma=0.0, shear=0.0):
    return soil_amd.gen_custom(seed, n, rmin, rmax, phi, aspect, nblobs, blob_frac, sigma, 0.1, shear)


def run_compare(ctx, gen, nsteps, every, k=None, exact=True, desc_mod=None, columns=(), col_data=(), pre_setup=None):
    d, pos, vel, mass, rad = gen
    if desc_mod:
        desc_mod(d)
    g, r = make_pair(ctx, d, pos, vel, mass, rad, k=k, columns=columns, col_data=col_data, pre_setup=pre_setup)
    done = 0
    while done < nsteps:
        s = min(every, nsteps - done)
        g.step(s)
        r.step(s)
        done += s
        print(f"compare step {done}", flush=True)
        assert_same(g, r, exact_state=exact, what=f"step {done}")
    return g, r


def test_c1_full_config_bit_exact_substrate(ctx):
    # C1 (4,000 monodisperse spheres, phi 0.30), 100 steps, no force: bit-exact everything
    print("case: C1 substrate only", flush=True)
    run_compare(ctx, soil_amd.gen_box(1), 100, 10)


def test_c1_full_config_with_consumer(ctx):
    print("case: C1 + harmonic k=1000", flush=True)
    run_compare(ctx, soil_amd.gen_box(1), 100, 20, k=1000.0)


def test_polydisperse_dense_with_consumer(ctx):
    # C2 shape at reduced size: r~U[0.4,0.6], phi 0.5, trigger-driven rebuilds
    print("case: polydisperse phi=0.5 n=6000", flush=True)
    g, r = run_compare(ctx, small(101, 6000, 0.4, 0.6, 0.5), 150, 25, k=1000.0)
    assert g.counts()["rebuilds"] > 3


def test_lees_edwards_clustered(ctx):
    # C3 shape at reduced size: 3:1:1, LE shear, Gaussian blobs over uniform background
    print("case: LE clustered n=12000", flush=True)
    run_compare(ctx, small(202, 12000, 0.45, 0.55, 0.4, (3, 1, 1), 16, 0.5, 2.0, 0.1), 40, 10, k=1000.0)


def test_lees_edwards_substrate_bit_exact_with_flip(ctx):
    # high strain rate so the tilt wraps (box flip) within the run
    print("case: LE flip n=3000", flush=True)

    def mod(d):
        d.dt = 1e-2
        d.shear_rate = 2.0
        d.has_shear = 1
    run_compare(ctx, small(303, 3000, 0.4, 0.6, 0.3), 60, 5, desc_mod=mod)


def test_periodic_sort_and_every_check(ctx):
    print("case: sort_every=7 every=5 check=0", flush=True)

    def mod(d):
        d.sort_every = 7
        d.every = 5
        d.check = 0
    run_compare(ctx, small(404, 2500, 0.4, 0.6, 0.35), 40, 1, desc_mod=mod)

    def mod2(d):
        d.sort_every = 3
        d.every = 4
        d.check = 1
    run_compare(ctx, small(405, 2500, 0.5, 0.5, 0.35), 30, 1, k=1000.0, desc_mod=mod2)


def test_full_list_newton_off(ctx):
    print("case: newton off", flush=True)

    def mod(d):
        d.newton = 0
    run_compare(ctx, small(505, 2000, 0.4, 0.6, 0.3), 20, 5, desc_mod=mod)


def test_atomdata_columns_forward_reverse_zero(ctx):
    # #[forward] overwrite on ghosts, #[reverse] += onto owners, #[zero] reset each step
    print("case: AtomData columns", flush=True)
    gen = small(606, 2000, 0.4, 0.6, 0.35)
    n = len(gen[1])
    rng = np.random.default_rng(6)
    cols = ((3, soil_amd._ffi.COL_FORWARD), (1, soil_amd._ffi.COL_REVERSE | soil_amd._ffi.COL_ZERO), (2, 0))
    data = (rng.normal(size=3 * n), np.zeros(n), rng.normal(size=2 * n))
    g, r = run_compare(ctx, gen, 12, 3, columns=cols, col_data=data)
    for c in range(3):
        assert np.array_equal(g.column(c), r.column(c)), f"column {c}"


def test_tiny_and_single_atom_boxes(ctx):
    print("case: tiny boxes", flush=True)
    run_compare(ctx, small(707, 1, 0.5, 0.5, 0.01), 5, 1)
    run_compare(ctx, small(708, 20, 0.4, 0.6, 0.05), 10, 2)


@pytest.mark.parametrize("shear", [0.0, 0.1])
def test_boundary_and_coincident_atoms(ctx, shear):
    # placements random boxes never hit: exactly on the lower faces, exactly
    # on the upper faces (outside [lo, hi) until the first pbc wraps them), a
    # corner (imaged through every swap), one ulp inside the upper face, on
    # bin faces (multiples of L/8), and a coincident pair (r2 = 0: a CSR pair
    # that the consumer skips)
    print(f"case: boundary placements shear={shear}", flush=True)
    d, pos, vel, mass, rad = small(909, 800, 0.4, 0.6, 0.3, shear=shear)
    lo, hi = np.array(d.lo), np.array(d.hi)
    L = hi - lo
    pos = pos.copy()
    pos[0] = [lo[0], pos[0][1], pos[0][2]]
    pos[1] = [pos[1][0], hi[1], pos[1][2]]
    pos[2] = lo
    pos[3] = hi
    pos[4] = np.nextafter(hi, lo)
    for k in range(5, 13):
        pos[k] = lo + L * ((k - 5) / 8.0)
    pos[14] = pos[13]
    vel = vel.copy()
    vel[14] = vel[13]
    run_compare(ctx, (d, pos, vel, mass, rad), 30, 3, k=1000.0)


def test_phase_by_phase_matches_oracle(ctx):
    # the per-system entry points follow the reference systems one at a time
    print("case: per-phase", flush=True)
    d, pos, vel, mass, rad = small(808, 3000, 0.4, 0.6, 0.4)
    g, r = make_pair(ctx, d, pos, vel, mass, rad, k=1000.0)
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
        g.sort_atoms_by_bin(); r.phase("sort_atoms_by_bin")
        if r.state == 0:
            g.borders(); r.phase("borders")
            g.bin_neighbor_list(); r.phase("bin_neighbor_list")
        else:
            g.forward_comm_borders(); r.phase("forward_comm_borders")
        g.harmonic_force(); r.phase("harmonic_force")
        g.reverse_send_force(); r.phase("reverse_send_force")
        g.final_integration(); r.phase("final_integration")
        cg, cr = g.counts(), r.counts()
        assert (cg["nghost"], cg["npairs"]) == (cr["nghost"], cr["npairs"])
        ag, ar = g.atoms(), r.atoms()
        for key in ("pos", "vel", "force"):
            assert np.array_equal(ag[key], ar[key]), f"step {step}: {key} differs"
        og, ig = g.csr()
        orr, ir = r.csr()
        assert np.array_equal(og, orr) and np.array_equal(ig, ir)


def test_sqrt_and_division_match_host(ctx):
    # the consumer's sqrt and division: one step on a box whose pairs overlap
    # heavily; forces must agree with the oracle bit for bit
    print("case: consumer forces", flush=True)
    d, pos, vel, mass, rad = small(909, 3000, 0.4, 0.6, 0.6)
    g, r = make_pair(ctx, d, pos, vel, mass, rad, k=1000.0)
    g.step(1)
    r.step(1)
    fg, fr = g.atoms()["force"], r.atoms()["force"]
    err = np.abs(fg - fr) / np.maximum(np.abs(fr), 1e-300)
    print("max rel force err", float(err.max()), "exact fraction", float(np.mean(fg == fr)), flush=True)
    assert float(err.max()) <= 1e-12
    assert np.array_equal(fg, fr)


def axis_deform(spec, remap=True, stage=None):
    # spec: [(dim, style, value, target)]; applied to both boxes before setup
    def f(b):
        for dim, style, value, target in spec:
            b.set_axis_deform(dim, style, value, target)
        b.set_deform_remap(remap)
        if stage is not None:
            b.set_stage_steps(stage)
    return f


@pytest.mark.parametrize("case", ["shrink_z_grow_x", "final_y_with_shear", "no_remap"])
def test_axis_deform_matches_oracle(ctx, case):
    # [deform] x/y/z styles with the affine remap (soil_deform/src/lib.rs:435-545):
    # the box and the remapped atoms follow the oracle bit for bit, the grid is
    # recomputed when the bounds move, and a growing box grows its cell arrays
    print(f"case: axis deform {case}", flush=True)
    nsteps = 120
    shear = 0.1 if case == "final_y_with_shear" else 0.0
    gen = small(4400, 5000, 0.4, 0.6, 0.45, shear=shear)
    d = gen[0]
    rate = 0.15 / (nsteps * d.dt)
    if case == "shrink_z_grow_x":
        L = d.hi[0] - d.lo[0]
        spec = [(2, "erate", -rate, (0, 0)), (0, "vel", 0.3 * L / (nsteps * d.dt), (0, 0))]
        f = axis_deform(spec, stage=nsteps)
    elif case == "final_y_with_shear":
        lo, hi = d.lo[1], d.hi[1]
        f = axis_deform([(1, "final", 0.0, (lo - 0.1 * (hi - lo), hi + 0.1 * (hi - lo)))], stage=nsteps)
    else:
        f = axis_deform([(2, "erate", -0.5 * rate, (0, 0))], remap=False, stage=nsteps)
    g, r = run_compare(ctx, gen, nsteps, 20, k=1000.0, pre_setup=f)
    for a, b in zip(g.domain(), r.domain()):
        assert np.array_equal(a, b)
    lo0 = np.array(d.lo)
    assert not np.array_equal(g.domain()[0], lo0)


def test_axis_collapse_grows_capacity_then_errors(ctx):
    # x shrinks by a tenth of its length per step: by step 9 the box is thinner
    # than the ghost cutoff, so every local is ghosted both ways and the ghost
    # and pair arenas both outgrow their set-up capacity inside one rebuild
    # (the regrown atom arena then bins every atom itself). Step 10 collapses
    # the axis, an error on both sides (soil_deform/src/lib.rs:506-515).
    print("case: axis collapse through capacity growth", flush=True)
    d, pos, vel, mass, rad = small(4401, 500, 0.4, 0.6, 0.3)
    L = d.hi[0] - d.lo[0]
    g, r = make_pair(ctx, d, pos, vel, mass, rad, k=1000.0,
                     pre_setup=axis_deform([(0, "vel", -L / (10 * d.dt), (0, 0))], stage=100))
    for s in range(1, 10):
        g.step(1)
        r.step(1)
        print(f"compare step {s}", flush=True)
        assert_same(g, r, what=f"collapse step {s}")
    assert g.counts()["nghost"] > 3 * len(mass)
    with pytest.raises(RuntimeError, match="collapsed"):
        g.step(1)
    with pytest.raises(RuntimeError, match="collapsed"):
        r.step(1)


def test_axis_compress_grows_capacity_within_one_call(ctx):
    # the collapse case's first nine steps in one call: the ghost and pair
    # arenas regrow between steps of the same call, while the box is stepping
    # ahead of its rebuild decisions (every step rebuilds)
    print("case: capacity growth inside one multi-step call", flush=True)
    d, pos, vel, mass, rad = small(4401, 500, 0.4, 0.6, 0.3)
    L = d.hi[0] - d.lo[0]
    g, r = make_pair(ctx, d, pos, vel, mass, rad, k=1000.0,
                     pre_setup=axis_deform([(0, "vel", -L / (10 * d.dt), (0, 0))], stage=100))
    g.step(9)
    r.step(9)
    assert_same(g, r, what="compressed 9 steps")
    assert g.counts()["nghost"] > 3 * len(mass)


def boundaries(kinds, padding=None):
    def f(b):
        for dim, kind in enumerate(kinds):
            b.set_boundary(dim, kind)
        if padding is not None:
            b.set_shrink_wrap_padding(padding)
    return f


@pytest.mark.parametrize("kinds,k,padding", [("pps", 1000.0, None), ("sss", 1000.0, 0.75), ("psf", None, None)])
def test_shrink_wrap_matches_oracle(ctx, kinds, k, padding):
    # shrink-wrap axes follow the locals' extent every step (domain.rs:509-589):
    # the bounds move, the grid is recomputed and every step rebuilds; the
    # device reduces the extents exactly and the host moves the bounds
    print(f"case: shrink-wrap {kinds}", flush=True)
    g, r = run_compare(ctx, small(5100, 3000, 0.4, 0.6, 0.35), 100, 10, k=k, pre_setup=boundaries(kinds, padding))
    for a, b in zip(g.domain(), r.domain()):
        assert np.array_equal(a, b)


def test_shrink_wrap_phase_by_phase(ctx):
    print("case: shrink-wrap per phase", flush=True)
    d, pos, vel, mass, rad = small(5101, 2000, 0.4, 0.6, 0.35)
    g, r = make_pair(ctx, d, pos, vel, mass, rad, k=1000.0, pre_setup=boundaries("spp"))
    for step in range(3):
        g.apply_deform(); r.phase("apply_deform")
        g.initial_integration(); r.phase("initial_integration")
        if step > 0:
            g.decide_rebuild(); r.phase("decide_rebuild")
        if r.state == 0:
            g.remove_ghost_atoms(); r.phase("remove_ghost_atoms")
        g.zero_all_forces(); r.phase("zero_all_forces")
        g.shrink_wrap(); r.phase("shrink_wrap")
        if r.state == 0:
            g.pbc(); r.phase("pbc")
        g.sort_atoms_by_bin(); r.phase("sort_atoms_by_bin")
        if r.state == 0:
            g.borders(); r.phase("borders")
            g.bin_neighbor_list(); r.phase("bin_neighbor_list")
        else:
            g.forward_comm_borders(); r.phase("forward_comm_borders")
        g.harmonic_force(); r.phase("harmonic_force")
        g.reverse_send_force(); r.phase("reverse_send_force")
        g.final_integration(); r.phase("final_integration")
        for a, b in zip(g.domain(), r.domain()):
            assert np.array_equal(a, b), f"step {step}: domain differs"
        ag, ar = g.atoms(), r.atoms()
        for key in ("pos", "vel", "force"):
            assert np.array_equal(ag[key], ar[key]), f"step {step}: {key} differs"
        og, ig = g.csr()
        orr, ir = r.csr()
        assert np.array_equal(og, orr) and np.array_equal(ig, ir)


def test_shrink_wrap_box_is_refused_by_a_batch(ctx):
    print("case: shrink-wrap in a batch", flush=True)
    d, pos, vel, mass, rad = small(5102, 500, 0.4, 0.6, 0.3)
    b = soil_amd.Box(ctx, d)
    b.upload(pos, mass, rad, vel=vel)
    b.set_boundary(2, "s")
    b.setup()
    with pytest.raises(RuntimeError, match="shrink-wrap"):
        soil_amd.Batch(ctx, [b])
    b.step(3)
    b.close()


@pytest.mark.parametrize("shear", [0.0, 0.15])
def test_classify_boundary_atoms_matches_oracle(ctx, shear):
    # neighbor.rs:1321-1336 on the device list equals the oracle's mask
    print(f"case: classify boundary shear={shear}", flush=True)
    g, r = run_compare(ctx, small(4501, 4000, 0.4, 0.6, 0.45, shear=shear), 30, 15, k=1000.0)
    mg, mr = g.classify_boundary_atoms(), r.classify_boundary_atoms()
    assert mg.dtype == bool and mg.size == mr.size and np.array_equal(mg, mr)
    assert 0 < mg.sum() < mg.size


def test_compute_ke_matches_oracle(ctx):
    # atom.rs:589-604: same per-atom terms; the device sums them in a fixed
    # tree order, the reference sequentially (stated bound: 1e-12 relative)
    print("case: compute_ke", flush=True)
    g, r = run_compare(ctx, small(4502, 5000, 0.4, 0.6, 0.45), 20, 20, k=1000.0)
    n = g.counts()["nlocal"]
    mask = (np.arange(n) % 3 == 0)
    for m in (None, mask):
        kg, kr = g.compute_ke(m), r.compute_ke(m)
        assert kr > 0 and abs(kg - kr) <= 1e-12 * kr, (kg, kr)
    assert g.compute_ke(np.zeros(n, dtype=bool)) == 0.0
    assert g.compute_ke() == g.compute_ke()  # deterministic


@pytest.mark.parametrize("slots", [4, 6])
def test_narrow_row_slots_take_the_overflow_path(ctx, monkeypatch, slots):
    # The one-pass CSR count keeps a row's first `slots` entries; a longer row
    # is walked again by the fill, which also gives its entries transpose
    # slots. Narrow slots (read at set-up) send most rows of a dense box down
    # that path; the result must not change.
    print(f"case: row slots {slots}", flush=True)
    monkeypatch.setenv("SOIL_HIP_ROW_SLOTS", str(slots))
    g, r = run_compare(ctx, small(7300 + slots, 4000, 0.4, 0.6, 0.55), 40, 8, k=1000.0)
    off, _ = r.csr()
    long_rows = int((np.diff(off.astype(np.int64)) > slots).sum())
    print(f"rows longer than {slots}: {long_rows}", flush=True)
    assert long_rows > 20


@pytest.mark.parametrize("slots", [None, 20])
def test_device_bytes_estimate_covers_setup(ctx, monkeypatch, slots):
    # the server admits jobs by the estimate; set-up must not allocate more
    # (the pair arena keeps SOIL_HIP_ROW_SLOTS entries per local)
    print(f"case: device bytes estimate slots={slots}", flush=True)
    if slots:
        monkeypatch.setenv("SOIL_HIP_ROW_SLOTS", str(slots))
    for seed, n, phi, shear in ((4601, 3000, 0.3, 0.0), (4602, 20000, 0.55, 0.1), (4603, 500, 0.2, 0.0)):
        d, pos, vel, mass, rad = soil_amd.gen_custom(seed, n, 0.4, 0.6, phi, (1, 1, 1), 0, 0.0, 0.0, 0.1, shear)
        b = soil_amd.Box(ctx, d)
        b.upload(pos, mass, rad, vel=vel)
        est0, alloc0 = b.device_bytes()
        assert alloc0 == 0
        b.setup()
        est, alloc = b.device_bytes()
        assert est == est0 and 0 < alloc <= est, (n, est, alloc)
        pair_bytes = 4 * 4 * (slots or 12) * n  # indices, row slots, transpose slot and source
        assert est > pair_bytes
        b.close()


def test_pair_growth_after_a_sort_keeps_the_columns(ctx):
    # an x axis compressed by 85 % over 70 steps: the pair arena (12 entries
    # per row) overflows at step 67, inside a fused step whose spatial sort
    # has swapped the column buffers; the growth must keep the swapped
    # descriptor (the redone build then reads the permuted columns)
    print("case: pair growth after a sort", flush=True)
    d, pos, vel, mass, rad = small(5103, 3000, 0.4, 0.6, 0.45)
    L = d.hi[0] - d.lo[0]
    g, r = make_pair(ctx, d, pos, vel, mass, rad, k=1000.0,
                     pre_setup=axis_deform([(0, "vel", -0.85 * L / (70 * d.dt), (0, 0))], stage=70))
    grew_pairs = False
    for s in range(1, 71):
        b0 = g.device_bytes()[1]
        g.step(1)
        r.step(1)
        grew_pairs |= s >= 61 and g.device_bytes()[1] > b0
        if s >= 60:
            assert_same(g, r, what=f"compress step {s}")
    assert grew_pairs


@pytest.mark.parametrize("mode", ["3", "4"])
@pytest.mark.parametrize("case", ["dense_poly", "lees_edwards", "newton_off", "narrow_regions", "uniform"])
def test_other_csr_modes_match_oracle(ctx, monkeypatch, mode, case):
    # CSR modes 3 (per-row slots, offset scan, fill) and 4 (a wave's
    # candidates flattened over its lanes, dense wave regions, the scan over
    # wave totals) give the same lists as the default one-pass mode: dense and
    # sheared boxes, the full list, slots or regions narrow enough that many
    # rows overflow into the fill's re-walk, and the uniform-radius test
    print(f"case: CSR mode {mode}, {case}", flush=True)
    monkeypatch.setenv("SOIL_HIP_CSR_MODE", mode)
    if case == "dense_poly":
        run_compare(ctx, small(4801, 5000, 0.4, 0.6, 0.55), 30, 10, k=1000.0)
    elif case == "lees_edwards":
        run_compare(ctx, small(4802, 4000, 0.4, 0.6, 0.45, (2, 1, 1), 8, 0.5, 3.0, shear=0.15), 20, 10, k=1000.0)
    elif case == "newton_off":
        def mod(d):
            d.newton = 0
        run_compare(ctx, small(4803, 2000, 0.4, 0.6, 0.3), 20, 10, desc_mod=mod)
    elif case == "narrow_regions":
        monkeypatch.setenv("SOIL_HIP_ROW_SLOTS", "4")
        run_compare(ctx, small(4804, 4000, 0.4, 0.6, 0.55), 30, 10, k=1000.0)
    else:
        run_compare(ctx, small(4805, 4000, 0.5, 0.5, 0.4), 30, 10, k=1000.0)


@pytest.mark.parametrize("case", ["clustered_half", "clustered_full", "narrow_tiles", "tiny"])
def test_tile_staged_csr_long_rows_and_small_boxes(ctx, monkeypatch, case):
    # the default build (mode 5): blobs whose rows run far past the 8 LDS
    # slots (walked again by their lanes), under the half and the full list;
    # tile regions so narrow (4 entries per row) that dense tiles stage
    # nothing and are walked again at placement; boxes of one partial tile
    print(f"case: tile-staged CSR, {case}", flush=True)
    if case == "clustered_half":
        run_compare(ctx, small(4811, 6000, 0.4, 0.6, 0.4, (2, 1, 1), 6, 0.6, 2.0, shear=0.1), 20, 10, k=1000.0)
    elif case == "clustered_full":
        def mod(d):
            d.newton = 0
        run_compare(ctx, small(4812, 3000, 0.4, 0.6, 0.35, (1, 1, 1), 4, 0.6, 2.0), 20, 10, desc_mod=mod)
    elif case == "narrow_tiles":
        monkeypatch.setenv("SOIL_HIP_ROW_SLOTS", "4")
        run_compare(ctx, small(4814, 6000, 0.4, 0.6, 0.45, (1, 1, 1), 6, 0.6, 2.0), 20, 10, k=1000.0)
    else:
        for n in (1, 63, 64, 200, 256, 257):
            run_compare(ctx, small(4813 + n, n, 0.4, 0.6, 0.3), 6, 3, k=1000.0)
