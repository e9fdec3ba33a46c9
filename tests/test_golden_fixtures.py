"""Replays the reference's known-answer fixtures (tests/golden/reference_fixtures.json).

Each fixture is data transcribed from the reference test it cites. The CPU
tests run every fixture on the oracle; the GPU tests run, through the C ABI,
every fixture whose inputs the device entry points can take (a fixture that
sets a ghost cutoff by hand, hands over a CSR of its own, or has no atoms at
all is oracle-only) and require the device to meet the fixture's expected
values within the reference test's tolerance AND to equal the oracle's
answer on the same fixture bit for bit.

Deform fixtures that hold no atoms check only the domain; the device replay
puts one atom at the box centre there, since a box on the device needs one.
"""
import json
import os

import numpy as np
import pytest

from oracle_ref import RefBox, classify_csr, ref_desc

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = json.load(open(os.path.join(HERE, "golden", "reference_fixtures.json")))["fixtures"]
BY_NAME = {f["name"]: f for f in FIXTURES}

DEVICE_KINDS = {"pbc", "phase", "pairs", "build_only", "deform", "shrink_wrap", "ke"}
ORACLE_ONLY = {"shrink_no_atoms", "shrink_ghost_cutoff_fallback", "shrink_size_fallback"}
DEVICE = [f["name"] for f in FIXTURES if f["kind"] in DEVICE_KINDS and f["name"] not in ORACLE_ONLY]


def _desc_args(fx):
    b = fx["box"]
    return dict(lo=b["lo"], hi=b["hi"], periodic=b["periodic"], dt=b["dt"], skin_fraction=b["skin_fraction"],
                bin_min_size=b["bin_min_size"])


def _atoms(fx, device):
    a = fx["atoms"]
    pos = np.array(a["pos"], dtype=np.float64).reshape(-1, 3)
    if device and fx["kind"] == "deform" and len(pos) == 0:
        b = fx["box"]
        pos = 0.5 * (np.array(b["lo"]) + np.array(b["hi"]))[None, :]
        return pos, np.ones(1), np.full(1, 0.1), None, None
    vel = np.array(a["vel"], dtype=np.float64) if "vel" in a else None
    force = np.array(a["force"], dtype=np.float64) if "force" in a else None
    return pos, np.array(a["mass"], dtype=np.float64), np.array(a["radius"], dtype=np.float64), vel, force


class _Oracle:
    """The oracle behind the device's method names."""

    def __init__(self, fx):
        self.b = RefBox(ref_desc(**_desc_args(fx)))
        self.ghost_cutoff = fx.get("ghost_cutoff")

    def upload(self, pos, mass, rad, vel=None, force=None):
        if len(pos):
            self.b.add_atoms(pos, mass, rad, vel=vel, force=force)

    def __getattr__(self, name):
        return getattr(self.b, name)

    def run(self, name):
        self.b.phase(name)


class _Device:
    def __init__(self, ctx, fx):
        import soil_amd
        self.b = soil_amd.Box(ctx, soil_amd.make_desc(**_desc_args(fx)))

    def upload(self, pos, mass, rad, vel=None, force=None):
        self.b.upload(pos, mass, rad, vel=vel, force=force)

    def __getattr__(self, name):
        return getattr(self.b, name)

    def run(self, name):
        getattr(self.b, name)()


def replay(fx, box):
    """Runs one fixture on `box` (oracle or device); returns what it checks."""
    kind = fx["kind"]
    device = isinstance(box, _Device)
    if kind == "shrink_wrap":
        for d, k in enumerate(fx["boundary"]):
            box.set_boundary(d, k)
        box.set_shrink_wrap_padding(fx["padding"])
        if fx.get("ghost_cutoff") is not None:
            box.set_ghost_cutoff(fx["ghost_cutoff"])
    if kind == "deform":
        for dim, style, value, target in fx["axes"]:
            box.set_axis_deform(dim, style, value, tuple(target))
        box.set_deform_remap(fx["remap"])
        box.set_stage_steps(fx["stage_steps"])
    pos, mass, rad, vel, force = _atoms(fx, device)
    box.upload(pos, mass, rad, vel=vel, force=force)
    out = {}
    if kind == "pbc":
        if device:
            box.setup()
        if "tilt" in fx:
            box.set_tilt(*fx["tilt"])
        box.run("pbc")
    elif kind == "phase":
        if device:
            box.setup()
        box.run(fx["phase"])
    elif kind == "pairs":
        box.setup()
        box.step(1)
    elif kind == "build_only":
        # bin_neighbor_list alone, as in the reference test: no pbc, no ghosts
        box.setup()
        if fx["tilt"][0]:
            box.set_tilt(*fx["tilt"])
        box.run("sort_atoms_by_bin")
        box.run("bin_neighbor_list")
    elif kind == "deform":
        box.setup()
        for _ in range(fx["steps"]):
            box.run("apply_deform")
    elif kind == "shrink_wrap":
        if device:
            box.setup()
        box.run("shrink_wrap")
    elif kind == "ke":
        if device:
            box.setup()
        out["ke"] = [box.compute_ke(None if m is None else np.array(m, dtype=bool)) for m in fx["masks"]]
        return out
    a = box.atoms()
    n = len(pos)
    out["pos"], out["vel"], out["image"] = a["pos"][:n].copy(), a["vel"][:n].copy(), a["image"][:n].copy()
    out["lo"], out["hi"], out["size"] = box.domain()
    if kind in ("pairs", "build_only"):
        off, idx = box.csr()
        tags = a["tag"].astype(np.int64)
        got = set()
        for i in range(n):
            for j in idx[off[i]:off[i + 1]]:
                ti, tj = int(tags[i]), int(tags[int(j)])
                got.add((min(ti, tj), max(ti, tj)))
        out["pairs"], out["offsets"] = got, np.asarray(off).astype(np.int64)
    return out


def check_expect(fx, out):
    exp, tol = fx["expect"], fx.get("tol", 0.0)
    tol_of = (lambda k: tol.get(k, 0.0)) if isinstance(tol, dict) else (lambda k: tol)
    for key in ("lo", "hi", "size"):
        for d, v in enumerate(exp.get(key, [])):
            if v is not None:
                assert abs(out[key][d] - v) <= tol_of(key), f"{fx['name']}: {key}[{d}] {out[key][d]} != {v}"
    for key in ("pos", "vel"):
        if key in exp:
            want = np.array(exp[key])
            assert np.all(np.abs(out[key] - want) <= tol_of(key)), f"{fx['name']}: {key} {out[key]} != {want}"
    if "image" in exp:
        assert np.array_equal(out["image"], np.array(exp["image"])), fx["name"]
    if "pairs" in exp:
        want = {tuple(p) for p in exp["pairs"]}
        assert want <= out["pairs"], f"{fx['name']}: missing {sorted(want - out['pairs'])}"
        assert out["pairs"] == want, f"{fx['name']}: extra {sorted(out['pairs'] - want)}"
    if "offsets" in exp:
        assert list(out["offsets"]) == exp["offsets"], fx["name"]
    if "ke" in exp:
        assert out["ke"] == exp["ke"], fx["name"]


def test_fixture_file_cites_the_reference():
    assert len(FIXTURES) >= 20
    for fx in FIXTURES:
        assert fx["source"].startswith("crates/") and ":" in fx["source"], fx["name"]


@pytest.mark.parametrize("name", [f["name"] for f in FIXTURES])
def test_oracle_meets_reference_fixture(name):
    fx = BY_NAME[name]
    if fx["kind"] == "classify_csr":
        c = fx["csr"]
        assert list(classify_csr(c["offsets"], c["indices"], c["nlocal"])) == fx["expect"]["mask"]
        return
    check_expect(fx, replay(fx, _Oracle(fx)))


@pytest.fixture(scope="module")
def ctx():
    import soil_amd
    return soil_amd.Context(0)


@pytest.mark.gpu
@pytest.mark.parametrize("name", DEVICE)
def test_device_meets_reference_fixture(ctx, name):
    fx = BY_NAME[name]
    print(f"case: golden fixture {name} ({fx['source']})", flush=True)
    dev = _Device(ctx, fx)
    try:
        got = replay(fx, dev)
    finally:
        dev.b.close()
    check_expect(fx, got)
    ref = replay(fx, _Oracle(fx))
    if not fx.get("atoms", {}).get("pos", [None]):
        ref = {k: v for k, v in ref.items() if k in ("lo", "hi", "size")}  # the device's centre atom aside
    for key, v in ref.items():
        if isinstance(v, np.ndarray):
            assert np.array_equal(got[key], v), f"{name}: {key} differs from the oracle"
        else:
            assert got[key] == v, f"{name}: {key} differs from the oracle"
