"""Restart files (soil_print RestartData) on the host: byte layout, JSON, round trips.

The reference writes RestartData with bincode 1 (crates/soil_print/src/lib.rs:
1441) or serde_json (:1436); neither library is in the reference tree and the
reference has no restart test or fixture, so the layout is pinned by an
independent decoder written from bincode 1's published format (fixint,
little-endian, u64 sequence lengths, struct fields in declaration order, every
field written) and by Python's json module.
"""
import json
import struct

import numpy as np
import pytest

from soil_amd import files

F64_FIELDS = ["pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "force_x", "force_y", "force_z", "mass",
              "cutoff_radius"]
LEGACY = ["omega_x", "omega_y", "omega_z", "torque_x", "torque_y", "torque_z", "ang_mom_x", "ang_mom_y",
          "ang_mom_z"]


def sample(n=5, nbuf=2, seed=3):
    rng = np.random.default_rng(seed)
    special = np.array([0.1, -0.0, 5e-324, 1.7976931348623157e308, 1e-7, 1e21, 2.5, -123.456])
    vals = np.concatenate([special, rng.normal(size=16 * n) * 10.0 ** rng.integers(-12, 12, size=16 * n)])
    take = iter(vals)
    pos = np.array([next(take) for _ in range(3 * n)]).reshape(n, 3)
    vel = np.array([next(take) for _ in range(3 * n)]).reshape(n, 3)
    force = np.array([next(take) for _ in range(3 * n)]).reshape(n, 3)
    return dict(tag=rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32),
                type_=rng.integers(0, 7, size=n).astype(np.uint32), pos=pos, vel=vel, force=force,
                mass=rng.uniform(0.5, 2.0, size=n), radius=rng.uniform(0.4, 0.6, size=n),
                buffers=[rng.normal(size=n * (k + 1)) for k in range(nbuf)], natoms=n + 11, step=4242, dt=2.5e-4)


def bincode_layout(s):
    """RestartData as bincode 1 encodes it, built field by field with struct."""
    n = len(s["tag"])
    out = struct.pack("<QQd", s["natoms"], s["step"], s["dt"])
    out += struct.pack("<Q", n) + struct.pack(f"<{n}I", *s["tag"])
    out += struct.pack("<Q", n) + struct.pack(f"<{n}I", *s["type_"])
    cols = [s["pos"][:, 0], s["pos"][:, 1], s["pos"][:, 2], s["vel"][:, 0], s["vel"][:, 1], s["vel"][:, 2],
            s["force"][:, 0], s["force"][:, 1], s["force"][:, 2], s["mass"], s["radius"]]
    for c in cols:
        out += struct.pack("<Q", n) + struct.pack(f"<{n}d", *c)
    out += struct.pack("<Q", len(s["buffers"]))
    for b in s["buffers"]:
        out += struct.pack("<Q", len(b)) + struct.pack(f"<{len(b)}d", *b)
    out += struct.pack("<Q", 0) * 9  # legacy rotational vectors, empty
    out += struct.pack("<Q", 0)      # quaternion, empty
    return out


def same_bits(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def check_equal(s, r):
    assert r["natoms"] == s["natoms"] and r["step"] == s["step"] and r["dt"] == s["dt"]
    assert np.array_equal(r["tag"], s["tag"]) and np.array_equal(r["type"], s["type_"])
    for k in ("pos", "vel", "force", "mass", "radius"):
        assert same_bits(r[k], s[k]), k
    assert len(r["buffers"]) == len(s["buffers"])
    for x, y in zip(r["buffers"], s["buffers"]):
        assert same_bits(x, y)


def test_bincode_bytes_follow_the_published_layout(tmp_path):
    s = sample()
    p = tmp_path / "restart_4242_rank0.bin"
    files.write_restart(p, files.BINCODE, **s)
    assert p.read_bytes() == bincode_layout(s)


@pytest.mark.parametrize("fmt", [files.BINCODE, files.JSON])
def test_restart_round_trip_is_exact(tmp_path, fmt):
    s = sample(n=40, nbuf=3, seed=9)
    p = tmp_path / "r"
    files.write_restart(p, fmt, **s)
    check_equal(s, files.read_restart(p, fmt))


def test_json_is_the_serde_object(tmp_path):
    s = sample()
    p = tmp_path / "restart_4242_rank0.json"
    files.write_restart(p, files.JSON, **s)
    d = json.loads(p.read_text())
    assert d["natoms"] == s["natoms"] and d["total_cycle"] == s["step"] and d["dt"] == s["dt"]
    assert d["tag"] == [int(t) for t in s["tag"]] and d["atom_type"] == [int(t) for t in s["type_"]]
    cols = [s["pos"][:, 0], s["pos"][:, 1], s["pos"][:, 2], s["vel"][:, 0], s["vel"][:, 1], s["vel"][:, 2],
            s["force"][:, 0], s["force"][:, 1], s["force"][:, 2], s["mass"], s["radius"]]
    for name, c in zip(F64_FIELDS, cols):
        assert same_bits(d[name], c), name
    assert len(d["atom_data_buffers"]) == len(s["buffers"])
    for x, y in zip(d["atom_data_buffers"], s["buffers"]):
        assert same_bits(x, y)
    for k in LEGACY + ["quaternion"]:
        assert d[k] == []


def test_reads_files_with_legacy_fields_and_other_spellings(tmp_path):
    # older files carry rotational vectors; serde_json may space, reorder and
    # spell numbers differently (integers for floats, exponents)
    s = sample(n=2, nbuf=1)
    raw = bytearray(bincode_layout(s))
    tail = struct.pack("<Q", 0) * 10
    assert raw.endswith(tail)
    legacy = b"".join(struct.pack("<Q", 2) + struct.pack("<2d", 1.5, -2.5) for _ in range(9))
    legacy += struct.pack("<Q", 1) + struct.pack("<4d", 1.0, 0.0, 0.0, 0.0)
    raw = raw[: -len(tail)] + legacy
    p = tmp_path / "old.bin"
    p.write_bytes(bytes(raw))
    check_equal(s, files.read_restart(p, files.BINCODE))

    doc = {"quaternion": [[1, 0, 0, 0]], "omega_x": [1e-7, 2], "dt": 1, "natoms": 2, "total_cycle": 7,
           "tag": [5, 6], "atom_type": [0, 1], "atom_data_buffers": [[1e21, -0.0]]}
    for name in F64_FIELDS:
        doc[name] = [1, 2.5e-3]
    p = tmp_path / "other.json"
    p.write_text(json.dumps(doc, indent=2))
    r = files.read_restart(p, files.JSON)
    assert r["dt"] == 1.0 and r["step"] == 7 and list(r["tag"]) == [5, 6]
    assert same_bits(r["mass"], [1.0, 2.5e-3]) and same_bits(r["buffers"][0], [1e21, -0.0])


def test_truncated_or_incomplete_files_are_errors(tmp_path):
    s = sample()
    p = tmp_path / "t.bin"
    files.write_restart(p, files.BINCODE, **s)
    p.write_bytes(p.read_bytes()[:-9])
    with pytest.raises(RuntimeError):
        files.read_restart(p, files.BINCODE)
    q = tmp_path / "t.json"
    q.write_text(json.dumps({"natoms": 1, "dt": 0.1}))
    with pytest.raises(RuntimeError, match="missing field"):
        files.read_restart(q, files.JSON)
