"""The device against the rules the reference's text fixes (pins.py): forward
comm, the #[forward] column contract and the rebuild schedule, asserted on
device runs directly rather than through agreement with the oracle."""
import numpy as np
import pytest

import soil_amd
from pins import check_forward, rebuild_steps_every, sort_is_identity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return soil_amd.Context(0)


def box(ctx, seed, n, phi, vel_sigma, every=0, check=1, sort_every=1000, columns=(), col_data=()):
    d, pos, vel, mass, rad = soil_amd.gen_custom(seed, n, 0.45, 0.55, phi, (1, 1, 1), 0, 0.0, 0.0, vel_sigma, 0.0)
    d.every, d.check, d.sort_every = every, check, sort_every
    b = soil_amd.Box(ctx, d)
    for w, f in columns:
        b.register_column(w, f)
    b.upload(pos, mass, rad, vel=vel)
    for c, data in enumerate(col_data):
        b.upload_column(c, data)
    b.setup()
    return b, d


def test_forward_comm_on_device():
    ctx = soil_amd.Context(0)
    print("case: device forward comm pin", flush=True)
    for tilt in (0.0, 1.7):
        b, d = box(ctx, 7001, 1200, 0.35, 2.0)
        if tilt:
            b.set_tilt(tilt, 0.3)
        b.step(1)
        comm_only = 0
        for s in range(8):
            before = b.counts()["rebuilds"]
            b.step(1)
            if b.counts()["rebuilds"] != before:
                continue
            comm_only += 1
            depth = check_forward(b.atoms(), b.swaps())
            assert depth.max() == 3
        assert comm_only >= 3
        b.close()


def test_forward_column_follows_its_atom_on_device(ctx):
    # a #[forward] column set from the tag at upload stays tied to its atom
    # through every sort, and every ghost carries its source's value
    print("case: device forward column pin", flush=True)
    n = 1500
    t = np.arange(n, dtype=np.float64)
    b, _ = box(ctx, 7003, n, 0.4, 0.5, columns=((2, soil_amd._ffi.COL_FORWARD),),
               col_data=(np.stack([t, 2.0 * t + 0.5], axis=1).ravel(),))
    for s in range(3):
        b.step(4)
        a = b.atoms()
        assert a["nghost"] > 0
        c = b.column(0).reshape(-1, 2)
        tag = a["tag"].astype(np.float64)
        assert np.array_equal(c[:, 0], tag) and np.array_equal(c[:, 1], 2.0 * tag + 0.5)
    b.close()


@pytest.mark.parametrize("every", [1, 2, 5])
def test_every_schedule_on_device(ctx, every):
    print(f"case: device every={every} schedule", flush=True)
    nsteps = 3 * (every + 1) + 2
    b, _ = box(ctx, 7010 + every, 1000, 0.3, 3.0, every=every, check=0, sort_every=0)
    got = []
    for s in range(1, nsteps + 1):
        before = b.counts()["rebuilds"]
        b.step(1)
        if b.counts()["rebuilds"] != before:
            got.append(s)
    assert got == rebuild_steps_every(every, nsteps)
    b.close()


@pytest.mark.parametrize("sigma", [0.0, 0.6])
def test_periodic_sort_rule_on_device(ctx, sigma):
    print(f"case: device periodic sort sigma={sigma}", flush=True)
    S = 5
    b, d = box(ctx, 7030, 2000, 0.3, sigma, every=0, check=1, sort_every=S)
    b.step(1)
    sorts = []
    for s in range(2, 4 * S + 1):
        a = b.atoms()
        nl = a["nlocal"]
        x = a["pos"][:nl] + a["vel"][:nl] * d.dt
        before = b.counts()["rebuilds"]
        b.step(1)
        rebuilt = b.counts()["rebuilds"] != before
        if s % S == 0:
            ident = sort_is_identity(x, b.grid())
            assert rebuilt == (not ident), (s, ident, rebuilt)
            sorts.append(ident)
        else:
            assert not rebuilt
    assert all(sorts) if sigma == 0.0 else not all(sorts)
    b.close()


@pytest.mark.parametrize("case", ["uniform", "poly", "tilted"])
def test_csr_row_order_on_device(ctx, case):
    # the device's raw CSR rows in the reference's loop order (own cell
    # first, then the forward stencil in generation order, bin order within
    # each cell), checked directly on device output
    from pins import check_csr_order
    print(f"case: device CSR row order {case}", flush=True)
    if case == "uniform":
        g = soil_amd.gen_custom(7060, 1500, 0.5, 0.5, 0.45, (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.0)
    elif case == "tilted":
        g = soil_amd.gen_custom(7061, 1500, 0.4, 0.6, 0.4, (2, 1, 1), 3, 0.5, 2.0, 0.1, 0.2)
    else:
        g = soil_amd.gen_custom(7062, 1500, 0.4, 0.6, 0.45, (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.0)
    d, pos, vel, mass, rad = g
    b = soil_amd.Box(ctx, d)
    b.upload(pos, mass, rad, vel=vel)
    b.setup()
    b.step(2)
    off, idx = b.csr()
    assert idx.size > 1000
    check_csr_order(off, idx, b.bins(), b.stencil(), b.grid()[0:3], True)
    b.close()


def test_cutoff_forms_on_device(ctx):
    # the pair between the two cutoff forms' roundings is listed exactly when
    # the radii span 1e-15 or more, on the device (pins.py lays the case out)
    from pins import CUT_POS, CUT_R, CUT_SF, cutoff_cases
    print("case: device cutoff forms", flush=True)
    for r3, listed in cutoff_cases():
        d = soil_amd._ffi.BoxDesc()
        for k in range(3):
            d.lo[k], d.hi[k], d.periodic[k] = 0.0, 20.0, 1
        d.newton, d.skin_fraction, d.bin_min_size = 1, CUT_SF, 0.0
        d.every, d.check, d.sort_every, d.dt = 0, 1, 1000, 1e-3
        b = soil_amd.Box(ctx, d)
        b.upload(CUT_POS.copy(), np.ones(3), np.array([CUT_R, CUT_R, r3]), vel=np.zeros((3, 3)))
        b.setup()
        b.step(1)
        off, idx = b.csr()
        tags = b.atoms()["tag"]
        pairs = {tuple(sorted((int(tags[i]), int(tags[j]))))
                 for i in range(len(off) - 1) for j in idx[off[i]:off[i + 1]]}
        assert pairs == ({(0, 1)} if listed else set()), (r3, pairs)
        b.close()


def test_distance_fma_nesting_on_device(ctx):
    # the device lists the pair the reference's fma nesting puts inside the
    # cutoff and not the one it puts on it (pins.py lays the pairs out)
    from pins import FMA_PI, FMA_PJ, FMA_R, FMA_SF, r2_reference
    print("case: device distance rounding", flush=True)
    c = 2.0 * FMA_R * FMA_SF
    for pj in FMA_PJ:
        listed = r2_reference(*[pj[k] - FMA_PI[k] for k in range(3)]) < c * c
        d = soil_amd._ffi.BoxDesc()
        for k in range(3):
            d.lo[k], d.hi[k], d.periodic[k] = 0.0, 20.0, 1
        d.newton, d.skin_fraction, d.bin_min_size = 1, FMA_SF, 0.0
        d.every, d.check, d.sort_every, d.dt = 0, 1, 1000, 1e-3
        b = soil_amd.Box(ctx, d)
        b.upload(np.array([FMA_PI, pj, (10.0, 10.0, 10.0)]), np.ones(3), np.full(3, FMA_R), vel=np.zeros((3, 3)))
        b.setup()
        b.step(1)
        off, idx = b.csr()
        tags = b.atoms()["tag"]
        pairs = {tuple(sorted((int(tags[i]), int(tags[j])))) for i in range(len(off) - 1) for j in idx[off[i]:off[i + 1]]}
        assert pairs == ({(0, 1)} if listed else set()), pj
        b.close()


def test_verlet_rounding_on_device(ctx):
    # the device's first step of one atom with an uploaded force: the reference
    # expression's velocity and position bit for bit (pins.py lays it out)
    from pins import VERLET_ATOM as A, VERLET_DT, verlet_reference
    print("case: device Verlet rounding", flush=True)
    ref = [verlet_reference(A["x"][k], A["v"][k], A["f"][k], A["m"], VERLET_DT) for k in range(3)]
    d = soil_amd._ffi.BoxDesc()
    for k in range(3):
        d.lo[k], d.hi[k], d.periodic[k] = 0.0, 20.0, 1
    d.newton, d.skin_fraction, d.bin_min_size = 1, 1.1, 0.0
    d.every, d.check, d.sort_every, d.dt = 0, 1, 1000, VERLET_DT
    b = soil_amd.Box(ctx, d)
    b.upload(np.array([A["x"]]), np.array([A["m"]]), np.array([0.5]), vel=np.array([A["v"]]),
             force=np.array([A["f"]]))
    b.setup()
    b.step(1)
    a = b.atoms()
    assert np.array_equal(a["vel"][0], np.array([r[0] for r in ref]))
    assert np.array_equal(a["pos"][0], np.array([r[1] for r in ref]))
    b.close()


def test_cell_formula_on_device(ctx):
    # atoms on cell faces where the product by the inverse bin size and a
    # division disagree: the device bins them by the product
    from pins import CELL_POS, cells, cells_by_division
    print("case: device cell formula", flush=True)
    d = soil_amd._ffi.BoxDesc()
    for k in range(3):
        d.lo[k], d.hi[k], d.periodic[k] = 0.0, 20.0, 1
    d.newton, d.skin_fraction, d.bin_min_size = 1, 1.1, 0.0
    d.every, d.check, d.sort_every, d.dt = 0, 1, 1000, 1e-3
    b = soil_amd.Box(ctx, d)
    b.upload(CELL_POS.copy(), np.ones(3), np.full(3, 0.5), vel=np.zeros((3, 3)))
    b.setup()
    b.step(1)
    a, g = b.atoms(), b.grid()
    _, _, _, atom_cell = b.bins()
    assert (cells_by_division(CELL_POS[:2], g) != cells(CELL_POS[:2], g)).all()
    assert np.array_equal(atom_cell[:3].astype(np.int64), cells(a["pos"][:3], g))
    b.close()


def test_displacement_trigger_forms_on_device(ctx):
    # step 2's decision for an atom between the fused and the plain squared
    # displacement: fused in a fully periodic box, plain with a fixed z
    from pins import DISP_CASES, DISP_DT, DISP_R, DISP_SF, disp_decisions
    print("case: device displacement trigger forms", flush=True)
    for x0, v in DISP_CASES:
        want = disp_decisions(x0, v)
        assert want[0] != want[1]
        for periodic_z, rebuild in zip((True, False), want):
            d = soil_amd._ffi.BoxDesc()
            for k in range(3):
                d.lo[k], d.hi[k], d.periodic[k] = 0.0, 20.0, 1
            d.newton, d.skin_fraction, d.bin_min_size = 1, DISP_SF, 0.0
            d.every, d.check, d.sort_every, d.dt = 0, 1, 1000, DISP_DT
            b = soil_amd.Box(ctx, d)
            if not periodic_z:
                b.set_boundary(2, "f")
            b.upload(np.array([x0]), np.ones(1), np.array([DISP_R]), vel=np.array([v]))
            b.setup()
            b.step(1)
            n1 = b.counts()["rebuilds"]
            b.step(1)
            assert b.counts()["rebuilds"] - n1 == (1 if rebuild else 0), (x0, periodic_z)
            b.close()


def test_periodic_wrap_bound_on_device(ctx):
    # the wrap's upper bound is low + size: an atom at x = hi (below it)
    # stays, an atom at y = lo + size (below hi) wraps
    from pins import WRAP_HI, WRAP_LO, wrap_case
    print("case: device periodic wrap bound", flush=True)
    pos, want_pos, want_img = wrap_case()
    d = soil_amd._ffi.BoxDesc()
    for k in range(3):
        d.lo[k], d.hi[k], d.periodic[k] = WRAP_LO[k], WRAP_HI[k], 1
    d.newton, d.skin_fraction, d.bin_min_size = 1, 1.1, 0.0
    d.every, d.check, d.sort_every, d.dt = 0, 1, 1000, 1e-3
    b = soil_amd.Box(ctx, d)
    b.upload(pos, np.ones(2), np.full(2, 0.5), vel=np.zeros((2, 3)))
    b.setup()
    b.step(1)
    a = b.atoms()
    o = np.argsort(a["tag"][:2])
    assert np.array_equal(a["pos"][:2][o], want_pos)
    assert np.array_equal(a["image"][:2][o], want_img)
    b.close()


def test_ghost_faces_on_device(ctx):
    # atoms on rounded face thresholds are sent by the x swaps on the device
    from pins import FACE_HI, FACE_LO, FACE_X, face_sends, sent_by_swap
    print("case: device ghost faces", flush=True)
    d = soil_amd._ffi.BoxDesc()
    for k in range(3):
        d.lo[k], d.hi[k], d.periodic[k] = FACE_LO[k], FACE_HI[k], 1
    d.newton, d.skin_fraction, d.bin_min_size = 1, 1.1, 0.0
    d.every, d.check, d.sort_every, d.dt = 0, 1, 1000, 1e-3
    b = soil_amd.Box(ctx, d)
    b.upload(np.array([[FACE_X[0], 10.0, 10.0], [FACE_X[1], 10.0, 10.0]]), np.ones(2), np.full(2, 0.5),
             vel=np.zeros((2, 3)))
    b.setup()
    b.step(1)
    ref, alt = face_sends(b.grid()[9])
    assert ref == (True, True) and alt == (False, False)
    assert sent_by_swap(b.atoms(), b.swaps()) == {0: {0}, 1: {1}}
    b.close()


def test_lees_edwards_tilt_on_device(ctx):
    # the device box's tilt after 9 steps is the reference's left-to-right product
    from pins import TILT_RATE, TILT_STEPS, tilt_reference
    print("case: device LE tilt", flush=True)
    d = soil_amd._ffi.BoxDesc()
    for k in range(3):
        d.lo[k], d.hi[k], d.periodic[k] = 0.0, 20.0, 1
    d.newton, d.skin_fraction, d.bin_min_size = 1, 1.1, 0.0
    d.every, d.check, d.sort_every, d.dt = 0, 1, 1000, 1e-3
    d.has_shear, d.shear_rate = 1, TILT_RATE
    b = soil_amd.Box(ctx, d)
    b.upload(np.array([[5.0, 5.0, 5.0], [12.0, 14.0, 9.0]]), np.ones(2), np.full(2, 0.5), vel=np.zeros((2, 3)))
    b.setup()
    b.step(TILT_STEPS)
    assert b.grid()[12] == tilt_reference(TILT_RATE, d.dt, TILT_STEPS, 20.0, 20.0)
    b.close()


def test_triclinic_wrap_on_device(ctx):
    # every local round-trips through lamda at the triclinic wrap, and
    # x2lamda subtracts left to right (pins.py lays the two atoms out)
    from pins import TRI_BVEL, TRI_HI, TRI_LO, TRI_POS, TRI_XY, tri_pbc
    print("case: device triclinic wrap", flush=True)
    want = [tri_pbc(list(p)) for p in TRI_POS]
    d = soil_amd._ffi.BoxDesc()
    for k in range(3):
        d.lo[k], d.hi[k], d.periodic[k] = TRI_LO[k], TRI_HI[k], 1
    d.newton, d.skin_fraction, d.bin_min_size = 1, 1.1, 0.0
    d.every, d.check, d.sort_every, d.dt = 0, 1, 1000, 1e-3
    b = soil_amd.Box(ctx, d)
    b.upload(TRI_POS.copy(), np.ones(2), np.full(2, 0.5), vel=np.zeros((2, 3)))
    b.setup()
    b.set_tilt(TRI_XY, TRI_BVEL)
    b.step(1)
    a = b.atoms()
    o = np.argsort(a["tag"][:2])
    assert np.array_equal(a["pos"][:2][o], np.array([w[0] for w in want]))
    assert np.array_equal(a["image"][:2][o], np.array([w[1] for w in want], dtype=np.int32))
    b.close()


def test_periodic_sort_step_on_device(ctx):
    # a periodic sort that rebuilds a CommunicateOnly step leaves the atom
    # that crossed a face unwrapped (pbc was skipped), on the device
    from pins import PSORT_EVERY, PSORT_POS, PSORT_VEL, psort_x_after
    print("case: device periodic sort without pbc", flush=True)
    d = soil_amd._ffi.BoxDesc()
    for k in range(3):
        d.lo[k], d.hi[k], d.periodic[k] = 0.0, 20.0, 1
    d.newton, d.skin_fraction, d.bin_min_size = 1, 1.1, 0.0
    d.every, d.check, d.sort_every, d.dt = 0, 1, PSORT_EVERY, 1e-3
    b = soil_amd.Box(ctx, d)
    b.upload(PSORT_POS.copy(), np.ones(2), np.full(2, 0.5), vel=PSORT_VEL.copy())
    b.setup()
    b.step(2)
    assert b.counts()["rebuilds"] == 1
    for s, rebuilt in ((3, 2), (4, 2)):
        b.step(1)
        assert b.counts()["rebuilds"] == rebuilt, s
        a = b.atoms()
        k = int(np.nonzero(a["tag"][:2] == 0)[0][0])
        assert a["pos"][k, 0] == psort_x_after(s) and a["pos"][k, 0] >= 20.0, s
        assert a["image"][k, 0] == 0
        if s == 3:
            assert k == 1
    b.close()


def test_streaming_profile_on_device(ctx):
    # the device's one-time shear profile takes y - yc first
    from pins import PROFILE_Y, TILT_RATE, profile_vx
    print("case: device streaming profile", flush=True)
    d = soil_amd._ffi.BoxDesc()
    for k in range(3):
        d.lo[k], d.hi[k], d.periodic[k] = 0.0, 20.0, 1
    d.newton, d.skin_fraction, d.bin_min_size = 1, 1.1, 0.0
    d.every, d.check, d.sort_every, d.dt = 0, 1, 1000, 1e-3
    d.has_shear, d.shear_rate = 1, TILT_RATE
    b = soil_amd.Box(ctx, d)
    b.upload(np.array([[10.0, PROFILE_Y, 10.0]]), np.ones(1), np.full(1, 0.5), vel=np.zeros((1, 3)))
    b.setup()
    b.step(1)
    assert b.atoms()["vel"][0, 0] == profile_vx(TILT_RATE, PROFILE_Y)
    b.close()


def test_triclinic_ghost_face_on_device(ctx):
    # the device sends the atom between the two forms of the tilted low-x
    # face threshold (pins.py lays it out)
    from pins import TRIF_POS, TRIF_XY, sent_by_swap, trif_lx, trif_thresholds
    print("case: device triclinic ghost face", flush=True)
    d = soil_amd._ffi.BoxDesc()
    for k in range(3):
        d.lo[k], d.hi[k], d.periodic[k] = 0.0, 20.0, 1
    d.newton, d.skin_fraction, d.bin_min_size = 1, 1.1, 0.0
    d.every, d.check, d.sort_every, d.dt = 0, 1, 1000, 1e-3
    b = soil_amd.Box(ctx, d)
    b.upload(TRIF_POS.copy(), np.ones(1), np.full(1, 0.5), vel=np.zeros((1, 3)))
    b.setup()
    b.set_tilt(TRIF_XY, 0.0)
    b.step(1)
    a = b.atoms()
    t_ref, t_alt = trif_thresholds(b.grid()[9])
    lx = trif_lx(a["pos"][0])
    assert lx < t_ref and not lx < t_alt
    assert sent_by_swap(a, b.swaps()).get(0) == {0}
    b.close()
