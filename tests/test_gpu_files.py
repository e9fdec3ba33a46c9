"""Restart and dump files written from, and read into, device boxes.

A job checkpointed to a RestartData file and resumed in a new box must step
exactly like the oracle resumed from the same file (read_restart semantics,
crates/soil_print/src/lib.rs:1527-1620: atoms, forces and AtomData columns
restored, image flags zero, a fresh neighbor state). Dump frames must carry the
box's atoms exactly, in the reference's text formatting (Rust `{}` for f64:
the shortest round-trip decimal in positional notation, which numpy's
format_float_positional(unique=True, trim='-') also produces).
"""
import struct

import numpy as np
import pytest

import soil_amd
from soil_amd import _ffi, files
from oracle_ref import RefBox
from paritylib import assert_same

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return soil_amd.Context(0)


COLS = ((2, _ffi.COL_FORWARD), (1, _ffi.COL_REVERSE | _ffi.COL_ZERO))


def make_box(ctx, d, pos, vel, mass, rad, types, col_data):
    b = soil_amd.Box(ctx, d)
    for w, f in COLS:
        b.register_column(w, f)
    b.upload(pos, mass, rad, vel=vel, type_=types)
    for c, data in enumerate(col_data):
        b.upload_column(c, data)
    b.set_consumer(1000.0)
    b.setup()
    return b


@pytest.mark.parametrize("fmt", [files.BINCODE, files.JSON])
def test_restart_resume_matches_oracle(ctx, tmp_path, fmt):
    print(f"case: restart resume fmt={fmt}", flush=True)
    d, pos, vel, mass, rad = soil_amd.gen_custom(1201, 3000, 0.4, 0.6, 0.4, (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.0)
    n = len(mass)
    rng = np.random.default_rng(12)
    types = rng.integers(0, 4, size=n).astype(np.uint32)
    col_data = (rng.normal(size=2 * n), np.zeros(n))
    a = make_box(ctx, d, pos, vel, mass, rad, types, col_data)
    a.step(37)
    path = tmp_path / f"restart_37_rank0.{'json' if fmt == files.JSON else 'bin'}"
    files.box_write_restart(a, path, fmt, 37)
    r = files.read_restart(path, fmt)
    at = a.atoms()
    assert r["step"] == 37 and r["natoms"] == n and r["dt"] == d.dt
    assert np.array_equal(r["tag"], at["tag"][:n])
    for key in ("pos", "vel", "force", "mass", "radius"):
        assert np.array_equal(r[key], at[key][:n]), key
    for c, (w, _) in enumerate(COLS):
        assert np.array_equal(r["buffers"][c], a.column(c)[: w * n])

    b = soil_amd.Box(ctx, d)
    for w, f in COLS:
        b.register_column(w, f)
    assert files.box_read_restart(b, path, fmt) == 37
    b.set_consumer(1000.0)
    b.setup()
    o = RefBox(d)
    for w, f in COLS:
        o.register_column(w, f)
    o.add_atoms(r["pos"], r["mass"], r["radius"], vel=r["vel"], force=r["force"], tag=r["tag"], type_=r["type"])
    for c, (w, _) in enumerate(COLS):
        o.set_column(c, r["buffers"][c], w)
    o.set_consumer(1000.0)
    o.setup()
    for k in range(5):
        b.step(5)
        o.step(5)
        assert_same(b, o, what=f"resumed step {5 * (k + 1)}")
    for c in range(len(COLS)):
        assert np.array_equal(b.column(c), o.column(c))
    a.close()
    b.close()


def rust_f64(v):
    return np.format_float_positional(v, unique=True, trim="-")


@pytest.mark.parametrize("ghost", [False, True])
def test_dump_frames_hold_the_box_atoms(ctx, tmp_path, ghost):
    print(f"case: dumps ghost={ghost}", flush=True)
    d, pos, vel, mass, rad = soil_amd.gen_custom(1301, 1500, 0.4, 0.6, 0.35, (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.0)
    n = len(mass)
    types = (np.arange(n) % 3).astype(np.uint32)
    b = soil_amd.Box(ctx, d)
    b.upload(pos, mass, rad, vel=vel, type_=types)
    b.set_consumer(1000.0)
    b.setup()
    b.step(12)
    at = b.atoms()
    m = at["nlocal"] + (at["nghost"] if ghost else 0)
    typ = np.zeros(at["nlocal"] + at["nghost"], dtype=np.uint32)
    _ffi.check(_ffi.lib().soil_hip_get_types(b.h, typ.ctypes.data_as(_ffi.U32P), typ.size))
    rows = []
    for i in range(m):
        vals = [*at["pos"][i], *at["vel"][i], *at["force"][i], at["radius"][i]]
        rows.append((int(at["tag"][i]), int(typ[i]), vals, int(i >= at["nlocal"])))

    stem = tmp_path / "dump_12"
    files.box_write_dump(b, stem, files.TEXT, 12, ghost)
    lines = (tmp_path / "dump_12.csv").read_text().splitlines()
    assert lines[0] == "tag,type,x,y,z,vx,vy,vz,fx,fy,fz,radius" + (",IsGhost" if ghost else "")
    assert len(lines) == m + 1
    for line, (tag, ty, vals, g) in zip(lines[1:], rows):
        want = [str(tag), str(ty)] + [rust_f64(v) for v in vals] + ([str(g)] if ghost else [])
        assert line.split(",") == want

    files.box_write_dump(b, stem, files.BINARY, 12, ghost)
    raw = (tmp_path / "dump_12.bin").read_bytes()
    rec = "<II10d" + ("B" if ghost else "")
    size = struct.calcsize(rec)
    assert struct.unpack_from("<I", raw)[0] == m and len(raw) == 4 + m * size
    for k, (tag, ty, vals, g) in enumerate(rows):
        got = struct.unpack_from(rec, raw, 4 + k * size)
        assert got[0] == tag and got[1] == ty and list(got[2:12]) == vals
        if ghost:
            assert got[12] == g

    files.box_write_dump(b, stem, files.LAMMPS, 12, ghost)
    lines = (tmp_path / "dump_12.lammpstrj").read_text().splitlines()
    head = ["ITEM: TIMESTEP", "12", "ITEM: NUMBER OF ATOMS", str(m), "ITEM: BOX BOUNDS pp pp pp"]
    head += [f"{rust_f64(d.lo[k])} {rust_f64(d.hi[k])}" for k in range(3)]
    head += ["ITEM: ATOMS id type x y z vx vy vz fx fy fz radius" + (" IsGhost" if ghost else "")]
    assert lines[:9] == head
    for line, (tag, ty, vals, g) in zip(lines[9:], rows):
        assert line.split(" ") == [str(tag), str(ty)] + [rust_f64(v) for v in vals] + ([str(g)] if ghost else [])
    b.close()
