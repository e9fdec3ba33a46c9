"""Writes reference_fixtures.json: the reference's own known-answer fixtures as data.

Every input and expected value below is transcribed from the reference test
it cites (file:line, relative to the reference repository); the only computed
entries are the pair lists of the two formula boxes, which are the
reference's own `reference_all_pairs` brute force (neighbor.rs:1433-1454:
every i < j with |xj - xi| < (ri + rj)·sf) evaluated here on the closed-form
positions of neighbor.rs:1464-1470 and 1549-1555.

Run from the repository root: python tests/golden/make_reference_fixtures.py
"""
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def formula_box(n, amp):
    pos = [[math.sin(i * 0.31) * amp + 2.0, math.cos(i * 0.67) * amp + 2.0, math.sin(i * 0.97) * amp + 2.0]
           for i in range(n)]
    rad = [0.3] * n
    pairs = []
    for i in range(n):
        for j in range(i + 1, n):
            d = [pos[j][k] - pos[i][k] for k in range(3)]
            if math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) < (rad[i] + rad[j]) * 1.0:
                pairs.append([i, j])
    return pos, rad, pairs


def box(lo, hi, periodic=(1, 1, 1), dt=1.0, skin_fraction=1.0, bin_min_size=1.0):
    return {"lo": list(lo), "hi": list(hi), "periodic": list(periodic), "dt": dt, "skin_fraction": skin_fraction,
            "bin_min_size": bin_min_size}


def main():
    fx = []
    # ── domain ──────────────────────────────────────────────────────────────
    fx.append({"name": "triclinic_pbc", "source": "crates/soil_core/src/domain.rs:1031-1058",
               "kind": "pbc", "box": box([0, 0, 0], [4, 3, 2]), "tilt": [1.2, 5.0],
               "atoms": {"pos": [[2.0, 3.03, 1.0]], "vel": [[7.0, 0.0, 0.0]], "mass": [1.0], "radius": [0.1]},
               "expect": {"pos": [[0.8, 0.03, 1.0]], "vel": [[2.0, 0.0, 0.0]], "image": [[0, 1, 0]]},
               "tol": {"pos": 1e-9, "vel": 1e-12}})
    fx.append({"name": "orthogonal_wrap", "source": "crates/soil_core/src/domain.rs:594-605",
               "kind": "pbc", "box": box([0, 0, 0], [10, 10, 10]),
               "atoms": {"pos": [[-0.5, 10.0, 5.0], [9.999, 0.0, 10.25]], "mass": [1.0, 1.0], "radius": [0.1, 0.1]},
               "expect": {"pos": [[9.5, 0.0, 5.0], [9.999, 0.0, 0.25]], "image": [[-1, 1, 0], [0, 0, 1]]},
               "tol": {"pos": 1e-15}})
    # ── velocity Verlet (make_atom: soil_verlet/src/lib.rs:209-219) ────────
    verlet_atoms = {"pos": [[0.0, 0.0, 0.0]], "vel": [[1.0, 0.0, 0.0]], "force": [[2.0, 0.0, 0.0]], "mass": [1.0],
                    "radius": [0.001]}
    fx.append({"name": "initial_integration", "source": "crates/soil_verlet/src/lib.rs:221-233",
               "kind": "phase", "phase": "initial_integration", "box": box([-10] * 3, [10] * 3, dt=0.01),
               "atoms": verlet_atoms,
               "expect": {"pos": [[0.0101, 0.0, 0.0]], "vel": [[1.01, 0.0, 0.0]]}, "tol": {"pos": 1e-10, "vel": 1e-10}})
    fx.append({"name": "final_integration", "source": "crates/soil_verlet/src/lib.rs:416-428",
               "kind": "phase", "phase": "final_integration", "box": box([-10] * 3, [10] * 3, dt=0.01),
               "atoms": verlet_atoms,
               "expect": {"pos": [[0.0, 0.0, 0.0]], "vel": [[1.01, 0.0, 0.0]]}, "tol": {"pos": 0.0, "vel": 1e-10}})
    # ── neighbor lists ──────────────────────────────────────────────────────
    fx.append({"name": "close_pair", "source": "crates/soil_core/src/neighbor.rs:1361-1402",
               "kind": "pairs", "box": box([0, 0, 0], [2, 2, 2], periodic=(0, 0, 0)),
               "atoms": {"pos": [[0.5, 0.5, 0.5], [1.0, 0.5, 0.5]], "mass": [1.0, 1.0], "radius": [0.5, 0.5]},
               "expect": {"pairs": [[0, 1]]}})
    fx.append({"name": "pair_iter_shape", "source": "crates/soil_core/src/neighbor.rs:1405-1413",
               "note": "three collinear atoms whose half list has the fixture's shape: offsets [0,2,3,3], "
                       "indices [1,2,2] -> pairs (0,1),(0,2),(1,2)",
               "kind": "pairs", "box": box([0, 0, 0], [10, 10, 10], periodic=(0, 0, 0), bin_min_size=5.0),
               "atoms": {"pos": [[4.0, 5.0, 5.0], [4.5, 5.0, 5.0], [5.0, 5.0, 5.0]], "mass": [1.0] * 3,
                         "radius": [0.6] * 3},
               "expect": {"pairs": [[0, 1], [0, 2], [1, 2]], "offsets": [0, 2, 3, 3]}})
    for n, amp, tilt, src in ((30, 1.5, 0.0, "crates/soil_core/src/neighbor.rs:1457-1535"),
                              (40, 1.4, 1.0, "crates/soil_core/src/neighbor.rs:1538-1619")):
        pos, rad, pairs = formula_box(n, amp)
        fx.append({"name": f"formula_box_{n}", "source": src, "kind": "build_only",
                   "box": box([0, 0, 0], [4, 4, 4], periodic=(0, 0, 0), bin_min_size=0.5), "tilt": [tilt, 0.0],
                   "atoms": {"pos": pos, "mass": [1.0] * n, "radius": rad},
                   "expect": {"pairs": pairs}})
    fx.append({"name": "classify_boundary_atoms", "source": "crates/soil_core/src/neighbor.rs:1344-1354",
               "kind": "classify_csr", "csr": {"offsets": [0, 1, 3, 3], "indices": [1, 0, 3], "nlocal": 3},
               "expect": {"mask": [False, True, False]}})
    # ── axis deformation (make_domain: soil_deform/src/lib.rs:626-637) ─────
    def deform(name, src, dt, hi, axes, stage, nsteps, expect, atoms=None, remap=True, tol=1e-12):
        fx.append({"name": name, "source": src, "kind": "deform",
                   "box": box([0, 0, 0], hi, dt=dt, skin_fraction=1.1, bin_min_size=0.2), "axes": axes,
                   "remap": remap, "stage_steps": stage, "steps": nsteps,
                   "atoms": atoms or {"pos": [], "mass": [], "radius": []}, "expect": expect, "tol": tol})
    deform("erate_shrinks", "crates/soil_deform/src/lib.rs:650-698", 0.001, [1, 1, 1],
           [[2, "erate", -1.0, [0, 0]]], 1000, 1,
           {"lo": [0.0, None, 0.0005], "hi": [1.0, None, 0.9995], "size": [None, None, 0.999]})
    deform("vel_expands", "crates/soil_deform/src/lib.rs:700-738", 0.01, [1, 1, 1],
           [[0, "vel", 1.0, [0, 0]]], 100, 1,
           {"lo": [-0.005, None, None], "hi": [1.005, None, None], "size": [1.01, None, None]})
    deform("final_ramps", "crates/soil_deform/src/lib.rs:740-789", 0.001, [1, 1, 1],
           [[1, "final", 0.0, [0.1, 0.9]]], 10, 10,
           {"lo": [None, 0.1, None], "hi": [None, 0.9, None]}, tol=1e-10)
    deform("remap_scales_atoms", "crates/soil_deform/src/lib.rs:791-846", 0.001, [1, 1, 1],
           [[2, "erate", -1.0, [0, 0]]], 1000, 1,
           {"pos": [[0.5, 0.5, 0.5], [0.75, 0.5, 0.5 + 0.25 * 0.999]]},
           atoms={"pos": [[0.5, 0.5, 0.5], [0.75, 0.5, 0.75]], "mass": [1.0, 1.0], "radius": [0.1, 0.1]})
    deform("no_remap", "crates/soil_deform/src/lib.rs:848-885", 0.001, [1, 1, 1],
           [[2, "erate", -1.0, [0, 0]]], 1000, 1,
           {"pos": [[0.5, 0.5, 0.5]], "size": [None, None, 0.999]},
           atoms={"pos": [[0.5, 0.5, 0.5]], "mass": [1.0], "radius": [0.1]}, remap=False)
    deform("multiple_axes", "crates/soil_deform/src/lib.rs:887-937", 0.01, [1, 1, 1],
           [[0, "vel", 1.0, [0, 0]], [2, "erate", -0.5, [0, 0]]], 100, 1,
           {"size": [1.01, 1.0, 0.995]})
    deform("volume", "crates/soil_deform/src/lib.rs:939-976", 0.001, [2, 3, 4],
           [[2, "erate", -1.0, [0, 0]]], 1000, 1,
           {"size": [2.0, 3.0, 4.0 * 0.999]})
    # ── shrink-wrap (domain.rs:880-996) ────────────────────────────────────
    def shrink(name, src, lo, hi, kinds, padding, atoms, expect, ghost_cutoff=None):
        fx.append({"name": name, "source": src, "kind": "shrink_wrap", "box": box(lo, hi), "boundary": kinds,
                   "padding": padding, "ghost_cutoff": ghost_cutoff,
                   "atoms": {"pos": atoms, "mass": [1.0] * len(atoms), "radius": [0.1] * len(atoms)},
                   "expect": expect, "tol": 1e-10})
    three = [[1.0, 2.0, 3.0], [5.0, 5.0, 8.0], [9.0, 1.0, 1.0]]
    shrink("shrink_expands", "crates/soil_core/src/domain.rs:880-907", [0, 0, 0], [10, 10, 10], "pps", 0.5, three,
           {"lo": [0.0, None, 0.5], "hi": [10.0, None, 8.5], "size": [None, None, 8.0]})
    shrink("shrink_within_tolerance", "crates/soil_core/src/domain.rs:909-933", [0, 0, 0.5], [10, 10, 8.5], "pps",
           0.5, three, {"lo": [None, None, 0.5], "hi": [None, None, 8.5]})
    shrink("shrink_no_atoms", "crates/soil_core/src/domain.rs:935-944", [0, 0, 0], [1, 1, 1], "sss", 0.0, [],
           {"lo": [0.0, 0.0, 0.0], "hi": [1.0, 1.0, 1.0]})
    shrink("shrink_all_axes", "crates/soil_core/src/domain.rs:946-975", [0, 0, 0], [100, 100, 100], "sss", 1.0,
           [[10.0, 20.0, 30.0], [50.0, 60.0, 70.0]],
           {"lo": [9.0, 19.0, 29.0], "hi": [51.0, 61.0, 71.0], "size": [42.0, 42.0, 42.0]})
    shrink("shrink_ghost_cutoff_fallback", "crates/soil_core/src/domain.rs:977-994", [0, 0, 0], [10, 10, 10], "pps",
           0.0, [[5.0, 5.0, 5.0]], {"lo": [None, None, 3.0], "hi": [None, None, 7.0]}, ghost_cutoff=2.0)
    shrink("shrink_size_fallback", "crates/soil_core/src/domain.rs:518-528", [0, 0, 0], [10, 20, 5], "spp", 0.0,
           [[4.0, 5.0, 1.0], [6.0, 5.0, 1.0]], {"lo": [3.8, None, None], "hi": [6.2, None, None]})
    # ── compute_ke (atom.rs:589-604) ────────────────────────────────────────
    fx.append({"name": "compute_ke", "source": "crates/soil_core/src/atom.rs:589-604", "kind": "ke",
               "box": box([0, 0, 0], [10, 10, 10]),
               "atoms": {"pos": [[1, 1, 1], [5, 5, 5], [8, 2, 3]], "mass": [2.0, 1.0, 4.0], "radius": [0.1] * 3,
                         "vel": [[1.0, 2.0, 2.0], [0.5, 0.0, -0.5], [0.0, 0.0, 3.0]]},
               "masks": [None, [1, 0, 1], [0, 0, 0]],
               "expect": {"ke": [0.5 * (2.0 * 9.0 + 1.0 * 0.5 + 4.0 * 9.0), 0.5 * (2.0 * 9.0 + 4.0 * 9.0), 0.0]}})
    with open(os.path.join(HERE, "reference_fixtures.json"), "w") as f:
        json.dump({"about": "known-answer fixtures of the reference's own tests (see make_reference_fixtures.py)",
                   "fixtures": fx}, f, indent=1)
    print(f"{len(fx)} fixtures")


if __name__ == "__main__":
    main()
