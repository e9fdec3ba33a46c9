"""Pins the CPU oracle on the reference's own known-answer tests and closed forms.

Each test names the reference test or expression it reproduces.
"""
import math

import numpy as np
import pytest

from oracle_ref import RefBox, ref_desc


def tilted_box():
    # domain.rs:997-1008 tilted_domain(): box [0,4]x[0,3]x[0,2], xy = 1.2
    b = RefBox(ref_desc([0, 0, 0], [4, 3, 2]))
    b.set_tilt(1.2, 0.0)
    return b


def test_lamda_roundtrip_with_tilt():
    # domain.rs:1010-1029
    b = tilted_box()
    for r in ([0.3, 0.7, 0.9], [3.9, 2.9, 0.1], [2.0, 1.5, 1.0]):
        back = b.lamda2x(b.x2lamda(r))
        assert np.all(np.abs(back - np.array(r)) < 1e-12)
    lam = b.x2lamda([1.2, 3.0, 0.0])
    assert abs(lam[1] - 1.0) < 1e-12
    assert abs(lam[0] - 0.0) < 1e-12


def test_triclinic_pbc_wraps_and_remaps_velocity():
    # domain.rs:1031-1058: y 3.03 -> 0.03, x 2.0 -> 0.8, vx 7 -> 2, image_y +1
    b = RefBox(ref_desc([0, 0, 0], [4, 3, 2]))
    b.add_atoms([[2.0, 3.03, 1.0]], [1.0], [0.1], vel=[[7.0, 0.0, 0.0]])
    b.set_tilt(1.2, 5.0)
    b.phase("pbc")
    a = b.atoms()
    assert abs(a["pos"][0, 1] - 0.03) < 1e-9
    assert abs(a["pos"][0, 0] - 0.8) < 1e-9
    assert abs(a["vel"][0, 0] - 2.0) < 1e-12
    assert a["image"][0, 1] == 1


def test_wrap_periodic_orthogonal():
    # domain.rs:594-605: p < lo -> +L, image -1; p >= lo+L -> -L, image +1
    b = RefBox(ref_desc([0, 0, 0], [10, 10, 10]))
    b.add_atoms([[-0.5, 10.0, 5.0], [9.999, 0.0, 10.25]], [1, 1], [0.1, 0.1])
    b.phase("pbc")
    a = b.atoms()
    assert np.allclose(a["pos"][0], [9.5, 0.0, 5.0], atol=0)
    assert list(a["image"][0]) == [-1, 1, 0]
    assert np.allclose(a["pos"][1], [9.999, 0.0, 0.25], atol=1e-15)
    assert list(a["image"][1]) == [0, 0, 1]


def make_verlet_atom():
    # soil_verlet/src/lib.rs:209-219 make_atom(): dt 0.01, v=(1,0,0), F=(2,0,0), m=1
    b = RefBox(ref_desc([-10, -10, -10], [10, 10, 10], dt=0.01))
    b.add_atoms([[0.0, 0.0, 0.0]], [1.0], [0.001], vel=[[1.0, 0.0, 0.0]], force=[[2.0, 0.0, 0.0]])
    return b


def test_initial_integration_kat():
    # soil_verlet/src/lib.rs:221-233: v = 1.01, x = 0.0101
    b = make_verlet_atom()
    b.phase("initial_integration")
    a = b.atoms()
    assert abs(a["vel"][0, 0] - 1.01) < 1e-10
    assert abs(a["pos"][0, 0] - 0.0101) < 1e-10


def test_final_integration_kat():
    # soil_verlet/src/lib.rs:416-428: v = 1.01, x unchanged
    b = make_verlet_atom()
    b.phase("final_integration")
    a = b.atoms()
    assert abs(a["vel"][0, 0] - 1.01) < 1e-10
    assert a["pos"][0, 0] == 0.0


def _vv_constant_force(dt, nsteps, mass, g, x0, v0):
    b = RefBox(ref_desc([-1e6] * 3, [1e6] * 3, dt=dt))
    b.add_atoms([[x0, 0.0, 0.0]], [mass], [0.001], vel=[[v0, 0.0, 0.0]], force=[[mass * g, 0, 0]])
    for _ in range(nsteps):
        b.phase("initial_integration")
        b.phase("final_integration")
    return b.atoms()


def test_constant_force_parabola():
    # soil_verlet/src/lib.rs:378-414: VV is exact for constant force (1e-10)
    a = _vv_constant_force(1e-4, 10000, 2.0, -9.81, 10.0, 5.0)
    t = 1.0
    assert abs(a["pos"][0, 0] - (10.0 + 5.0 * t + 0.5 * -9.81 * t * t)) < 1e-9
    assert abs(a["vel"][0, 0] - (5.0 + -9.81 * t)) < 1e-10


def test_harmonic_oscillator_second_order():
    # soil_verlet/src/lib.rs:241-298: error ratio at dt, dt/2 is about 4 (O(dt^2)).
    # Drives the oracle's kick/drift with F = -k x recomputed between the phases.
    def run(dt):
        b = RefBox(ref_desc([-100] * 3, [100] * 3, dt=dt))
        b.add_atoms([[1.0, 0.0, 0.0]], [1.0], [0.001], force=[[-1.0, 0, 0]])
        n = int(round(1.0 / dt))
        for _ in range(n):
            b.phase("initial_integration")
            x = b.atoms()["pos"][0, 0]
            b.set_force(0, [-x, 0.0, 0.0])
            b.phase("final_integration")
        return abs(b.atoms()["pos"][0, 0] - math.cos(1.0))
    e1, e2 = run(0.01), run(0.005)
    assert 3.5 < e1 / e2 < 4.5


def test_grid_closed_forms():
    # neighbor.rs:316-389 with the C1 and C2 boxes: 23^3 and 83^3 cells, 13 forward cells
    for L, r, grid, fwd in ((19.112277960442917, 0.5, 23, 13),):
        b = RefBox(ref_desc([0, 0, 0], [L, L, L], skin_fraction=1.1))
        b.add_atoms([[1.0, 1.0, 1.0]], [1.0], [r])
        b.setup()
        g = b.grid()
        assert list(g[:3]) == [grid] * 3
        assert len(b.stencil(True)) == fwd
        # bin = L / floor(L / 1.1); s = ceil(1.2/bin) + ceil(1.1/bin) = 2 + 1
        ni = math.floor(L / 1.1)
        assert g[3] == L / ni
        assert ni + 2 * 3 == grid
        assert abs(g[9] - 1.2) < 1e-15 and abs(g[10] - 1.1) < 1e-15


def test_le_tilt_closed_form():
    # soil_deform/src/lib.rs:576-583: raw = ((rate*dt)*step)*Ly, xy = raw - Lx*round(raw/Lx)
    rate, dt, L = 0.5, 0.01, 10.0
    b = RefBox(ref_desc([0, 0, 0], [L, L, L], dt=dt, skin_fraction=1.1, shear_rate=rate))
    b.add_atoms([[5.0, 5.0, 5.0], [1.0, 2.0, 3.0]], [1, 1], [0.5, 0.5])
    b.setup()
    for step in range(1, 250):
        b.phase("apply_deform")
        g = b.grid()
        raw = rate * dt * step * L
        assert g[12] == raw - L * round(raw / L) or g[12] == raw - L * math.floor(raw / L + 0.5)
        assert g[15] == rate * L
    # one-time streaming profile: v_x += rate (y - y_c) applied once
    a = b.atoms()
    assert abs(a["vel"][0, 0] - rate * (5.0 - 5.0)) < 1e-15
    assert abs(a["vel"][1, 0] - rate * (2.0 - 5.0)) < 1e-15


def test_pair_iter_fixture_shape():
    # neighbor.rs:1405-1413: offsets [0,2,3,3], indices [1,2,2] -> (0,1),(0,2),(1,2).
    # The oracle's CSR for three collinear atoms in a closed box has that shape.
    b = RefBox(ref_desc([0, 0, 0], [10, 10, 10], periodic=(0, 0, 0), bin_min_size=5.0))
    b.add_atoms([[4.0, 5, 5], [4.5, 5, 5], [5.0, 5, 5]], [1, 1, 1], [0.6, 0.6, 0.6])
    b.setup()
    b.step(1)
    off, idx = b.csr()
    pairs = sorted((i, int(j)) for i in range(3) for j in idx[off[i]:off[i + 1]])
    assert pairs == [(0, 1), (0, 2), (1, 2)]
    assert list(off) == [0, 2, 3, 3]


def test_close_pair_fixture():
    # neighbor.rs:1361-1402: atoms (0.5,0.5,0.5) and (1.0,0.5,0.5), r=0.5, box 2^3 -> pair (0,1)
    b = RefBox(ref_desc([0, 0, 0], [2, 2, 2], periodic=(0, 0, 0)))
    b.add_atoms([[0.5, 0.5, 0.5], [1.0, 0.5, 0.5]], [1, 1], [0.5, 0.5])
    b.setup()
    b.step(1)
    off, idx = b.csr()
    assert 1 in list(idx[off[0]:off[1]])


def _reference_all_pairs(pos, rad, sf):
    # neighbor.rs:1433-1454 reference_all_pairs
    n = len(pos)
    out = []
    for i in range(n):
        for j in range(i + 1, n):
            d = pos[j] - pos[i]
            if math.sqrt(d @ d) < (rad[i] + rad[j]) * sf:
                out.append((i, j))
    return out


@pytest.mark.parametrize("tilt", [0.0, 1.0])
def test_fixture_formula_boxes_match_reference_all_pairs(tilt):
    # neighbor.rs:1457-1535 (30 atoms) and 1538-1619 (40 atoms, xy = 1.0)
    n, amp = (30, 1.5) if tilt == 0.0 else (40, 1.4)
    pos = np.array([[math.sin(i * 0.31) * amp + 2.0, math.cos(i * 0.67) * amp + 2.0, math.sin(i * 0.97) * amp + 2.0]
                    for i in range(n)])
    rad = np.full(n, 0.3)
    b = RefBox(ref_desc([0, 0, 0], [4, 4, 4], periodic=(0, 0, 0), bin_min_size=0.5))
    b.add_atoms(pos, np.ones(n), rad)
    b.setup()
    if tilt:
        b.set_tilt(tilt, 0.0)
    # bin_neighbor_list alone, as in the reference test (no pbc / ghosts)
    b.phase("sort_atoms_by_bin")
    b.phase("bin_neighbor_list")
    off, idx = b.csr()
    a = b.atoms()
    tags = a["tag"]
    got = set()
    for i in range(n):
        for j in idx[off[i]:off[i + 1]]:
            ti, tj = int(tags[i]), int(tags[int(j)])
            got.add((min(ti, tj), max(ti, tj)))
    ref = set(_reference_all_pairs(pos, rad, 1.0))
    assert ref <= got  # the reference's assertion
    assert got == ref  # and nothing extra (stricter than the reference test)


def unit_box(dt, hi=(1.0, 1.0, 1.0), atoms=()):
    # soil_deform/src/lib.rs:626-637 make_domain([0,0,0], hi): the deform tests' box
    b = RefBox(ref_desc([0, 0, 0], list(hi), dt=dt, skin_fraction=1.1, bin_min_size=0.2))
    if atoms:
        b.add_atoms(list(atoms), [1.0] * len(atoms), [0.1] * len(atoms))
    return b


def test_erate_shrinks_box_symmetrically():
    # soil_deform/src/lib.rs:650-698: z erate -1, dt 0.001 -> [0.0005, 0.9995], size 0.999
    b = unit_box(0.001)
    b.set_axis_deform(2, "erate", -1.0)
    b.set_stage_steps(1000)
    b.setup()
    b.phase("apply_deform")
    lo, hi, size = b.domain()
    assert abs(lo[2] - 0.0005) < 1e-12 and abs(hi[2] - 0.9995) < 1e-12 and abs(size[2] - 0.999) < 1e-12
    assert abs(lo[0]) < 1e-12 and abs(hi[0] - 1.0) < 1e-12


def test_vel_expands_box():
    # soil_deform/src/lib.rs:700-738: x vel 1, dt 0.01 -> size 1.01 about the fixed center
    b = unit_box(0.01)
    b.set_axis_deform(0, "vel", 1.0)
    b.set_stage_steps(100)
    b.setup()
    b.phase("apply_deform")
    lo, hi, size = b.domain()
    assert abs(size[0] - 1.01) < 1e-12
    assert abs(lo[0] + 0.005) < 1e-12 and abs(hi[0] - 1.005) < 1e-12


def test_final_ramps_to_target():
    # soil_deform/src/lib.rs:740-789: y final [0.1, 0.9] reached after 10 of 10 steps,
    # linear in between: lo = lo0 + (tlo - lo0) step / total
    b = unit_box(0.001)
    b.set_axis_deform(1, "final", target=(0.1, 0.9))
    b.set_stage_steps(10)
    b.setup()
    for step in range(1, 11):
        b.phase("apply_deform")
        lo, hi, _ = b.domain()
        assert abs(lo[1] - 0.1 * step / 10) < 1e-12 and abs(hi[1] - (1.0 - 0.1 * step / 10)) < 1e-12
    assert abs(lo[1] - 0.1) < 1e-10 and abs(hi[1] - 0.9) < 1e-10


def test_remap_scales_atom_positions():
    # soil_deform/src/lib.rs:791-846: the center atom stays at the center, x untouched;
    # the 3/4 atom maps affinely: 0.5 + 0.25 * 0.999
    b = unit_box(0.001, atoms=([0.5, 0.5, 0.5], [0.75, 0.5, 0.75]))
    b.set_axis_deform(2, "erate", -1.0)
    b.set_stage_steps(1000)
    b.setup()
    b.phase("apply_deform")
    a = b.atoms()
    lo, hi, _ = b.domain()
    assert abs(a["pos"][0, 2] - 0.5 * (lo[2] + hi[2])) < 1e-12
    assert abs(a["pos"][0, 0] - 0.5) < 1e-12 and abs(a["pos"][1, 0] - 0.75) < 1e-12
    assert abs(a["pos"][1, 2] - (0.5 + 0.25 * 0.999)) < 1e-12


def test_no_remap_leaves_atoms_in_place():
    # soil_deform/src/lib.rs:848-885
    b = unit_box(0.001, atoms=([0.5, 0.5, 0.5],))
    b.set_axis_deform(2, "erate", -1.0)
    b.set_deform_remap(False)
    b.set_stage_steps(1000)
    b.setup()
    b.phase("apply_deform")
    assert abs(b.atoms()["pos"][0, 2] - 0.5) < 1e-12
    assert abs(b.domain()[2][2] - 0.999) < 1e-12


def test_multiple_axes_deform_independently():
    # soil_deform/src/lib.rs:887-937: x vel 1 -> 1.01, y unchanged, z erate -0.5 -> 0.995
    b = unit_box(0.01)
    b.set_axis_deform(0, "vel", 1.0)
    b.set_axis_deform(2, "erate", -0.5)
    b.set_stage_steps(100)
    b.setup()
    b.phase("apply_deform")
    size = b.domain()[2]
    assert abs(size[0] - 1.01) < 1e-12 and abs(size[1] - 1.0) < 1e-12 and abs(size[2] - 0.995) < 1e-12


def test_volume_follows_the_sizes():
    # soil_deform/src/lib.rs:939-976: box 2x3x4, z erate -1 -> volume = product of sizes
    b = unit_box(0.001, hi=(2.0, 3.0, 4.0))
    b.set_axis_deform(2, "erate", -1.0)
    b.set_stage_steps(1000)
    b.setup()
    b.phase("apply_deform")
    size = b.domain()[2]
    assert abs(size[2] - 4.0 * 0.999) < 1e-12 and abs(size[0] * size[1] * size[2] - 2 * 3 * 4 * 0.999) < 1e-10


def test_collapsed_axis_is_an_error():
    # soil_deform/src/lib.rs:506-515: a size <= 0 ends the run
    b = unit_box(0.1, atoms=([0.5, 0.5, 0.5],))
    b.set_axis_deform(0, "vel", -1.0)
    b.set_stage_steps(100)
    b.setup()
    with pytest.raises(RuntimeError, match="collapsed"):
        b.step(20)


def shrink_box(lo, hi, kinds, padding, atoms):
    # domain.rs:880-996: a domain with the given boundary kinds and padding
    b = RefBox(ref_desc(list(lo), list(hi)))
    for d, k in enumerate(kinds):
        b.set_boundary(d, k)
    b.set_shrink_wrap_padding(padding)
    if atoms:
        b.add_atoms(list(atoms), [1.0] * len(atoms), [0.1] * len(atoms))
    return b


def test_shrink_wrap_expands_to_atom_positions():
    # domain.rs:880-907: z -> [1 - 0.5, 8 + 0.5], x and y untouched
    b = shrink_box([0, 0, 0], [10, 10, 10], "pps", 0.5, ([1.0, 2.0, 3.0], [5.0, 5.0, 8.0], [9.0, 1.0, 1.0]))
    b.phase("shrink_wrap")
    lo, hi, size = b.domain()
    assert abs(lo[2] - 0.5) < 1e-10 and abs(hi[2] - 8.5) < 1e-10 and abs(size[2] - 8.0) < 1e-10
    assert abs(lo[0]) < 1e-10 and abs(hi[0] - 10.0) < 1e-10


def test_shrink_wrap_no_change_within_tolerance():
    # domain.rs:909-933: bounds already at the wrap leave the domain as it is
    b = shrink_box([0, 0, 0.5], [10, 10, 8.5], "pps", 0.5, ([1.0, 2.0, 3.0], [5.0, 5.0, 8.0], [9.0, 1.0, 1.0]))
    b.phase("shrink_wrap")
    lo, hi, _ = b.domain()
    assert lo[2] == 0.5 and hi[2] == 8.5


def test_shrink_wrap_no_atoms_is_noop():
    # domain.rs:935-944
    b = shrink_box([0, 0, 0], [1, 1, 1], "sss", 0.0, ())
    b.phase("shrink_wrap")
    lo, hi, _ = b.domain()
    assert list(lo) == [0, 0, 0] and list(hi) == [1, 1, 1]


def test_shrink_wrap_all_axes():
    # domain.rs:946-975: padding 1 -> [9,51] [19,61] [29,71], sizes 42
    b = shrink_box([0, 0, 0], [100, 100, 100], "sss", 1.0, ([10.0, 20.0, 30.0], [50.0, 60.0, 70.0]))
    b.phase("shrink_wrap")
    lo, hi, size = b.domain()
    assert np.allclose(lo, [9, 19, 29], atol=1e-10) and np.allclose(hi, [51, 61, 71], atol=1e-10)
    assert np.allclose(size, [42, 42, 42], atol=1e-10)


def test_shrink_wrap_uses_ghost_cutoff_fallback():
    # domain.rs:977-994: no padding -> the ghost cutoff (2.0): z -> [3, 7]
    b = shrink_box([0, 0, 0], [10, 10, 10], "pps", 0.0, ([5.0, 5.0, 5.0],))
    b.set_ghost_cutoff(2.0)
    b.phase("shrink_wrap")
    lo, hi, _ = b.domain()
    assert abs(lo[2] - 3.0) < 1e-10 and abs(hi[2] - 7.0) < 1e-10


def test_shrink_wrap_size_fallback():
    # domain.rs:518-528: no padding and no ghost cutoff -> 1% of the largest size
    b = shrink_box([0, 0, 0], [10, 20, 5], "spp", 0.0, ([4.0, 5.0, 1.0], [6.0, 5.0, 1.0]))
    b.phase("shrink_wrap")
    lo, hi, _ = b.domain()
    assert lo[0] == 4.0 - 0.2 and hi[0] == 6.0 + 0.2


def test_classify_boundary_atoms_fixture():
    # neighbor.rs:1344-1354: offsets [0,1,3,3], indices [1,0,3], 3 locals -> [F, T, F]
    from oracle_ref import classify_csr
    assert list(classify_csr([0, 1, 3, 3], [1, 0, 3], 3)) == [False, True, False]


def test_boundary_atoms_lie_near_a_periodic_face():
    # a local with a ghost neighbour is within the pair cutoff of a face whose
    # images it meets; rows without ghosts are interior
    import soil_amd
    d, pos, vel, mass, rad = soil_amd.gen_custom(4500, 3000, 0.4, 0.6, 0.4, (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.0)
    b = RefBox(d)
    b.add_atoms(pos, mass, rad, vel=vel)
    b.setup()
    b.step(3)
    mask = b.classify_boundary_atoms()
    a = b.atoms()
    n = len(mass)
    off, idx = b.csr()
    lo, hi = np.array(d.lo), np.array(d.hi)
    cut = 2 * 0.6 * d.skin_fraction
    p = a["pos"][:n]
    near = np.any((p - lo < cut) | (hi - p <= cut), axis=1)
    assert mask.sum() > 0 and np.all(near[mask])
    for i in np.nonzero(mask)[0][:50]:
        assert np.any(idx[off[i]:off[i + 1]] >= n)
    for i in np.nonzero(~mask)[0][:50]:
        assert np.all(idx[off[i]:off[i + 1]] < n)


def test_compute_ke_closed_form():
    # atom.rs:589-604: 0.5 * sum m |v|^2 over the (masked) locals; exact here
    b = RefBox(ref_desc([0, 0, 0], [10, 10, 10]))
    b.add_atoms([[1, 1, 1], [5, 5, 5], [8, 2, 3]], [2.0, 1.0, 4.0], [0.1] * 3,
                vel=[[1.0, 2.0, 2.0], [0.5, 0.0, -0.5], [0.0, 0.0, 3.0]])
    assert b.compute_ke() == 0.5 * (2.0 * 9.0 + 1.0 * 0.5 + 4.0 * 9.0)
    assert b.compute_ke([1, 0, 1]) == 0.5 * (2.0 * 9.0 + 4.0 * 9.0)
    assert b.compute_ke([0, 0, 0]) == 0.0
