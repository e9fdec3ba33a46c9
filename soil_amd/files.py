"""Restart and dump files of soil_print through the C ABI.

RestartData (crates/soil_print/src/lib.rs:317-398) in bincode 1 or JSON, and
dump frames (:1272-1380) as CSV text, little-endian binary or LAMMPS text.
"""
import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import check, lib

BINCODE, JSON = _ffi.RESTART_BINCODE, _ffi.RESTART_JSON
TEXT, BINARY, LAMMPS = _ffi.DUMP_TEXT, _ffi.DUMP_BINARY, _ffi.DUMP_LAMMPS


def _dp(a):
    return a.ctypes.data_as(_ffi.DP)


def _u32p(a):
    return a.ctypes.data_as(_ffi.U32P)


def write_restart(path, fmt, *, tag, type_, pos, vel, force, mass, radius, buffers=(), natoms=None, step=0,
                  dt=1e-3):
    """Writes one RestartData file (n atoms, AoS-of-3 pos/vel/force)."""
    tag = np.ascontiguousarray(tag, dtype=np.uint32)
    n = len(tag)
    arrs = {
        "type": np.ascontiguousarray(type_, dtype=np.uint32),
        "pos": np.ascontiguousarray(pos, dtype=np.float64).reshape(n, 3),
        "vel": np.ascontiguousarray(vel, dtype=np.float64).reshape(n, 3),
        "force": np.ascontiguousarray(force, dtype=np.float64).reshape(n, 3),
        "mass": np.ascontiguousarray(mass, dtype=np.float64),
        "radius": np.ascontiguousarray(radius, dtype=np.float64),
    }
    bufs = [np.ascontiguousarray(b, dtype=np.float64) for b in buffers]
    blen = (C.c_size_t * max(len(bufs), 1))(*[b.size for b in bufs])
    bptr = (_ffi.DP * max(len(bufs), 1))(*[_dp(b) for b in bufs])
    r = _ffi.Restart(natoms if natoms is not None else n, step, dt, n, _u32p(tag), _u32p(arrs["type"]),
                     _dp(arrs["pos"]), _dp(arrs["vel"]), _dp(arrs["force"]), _dp(arrs["mass"]),
                     _dp(arrs["radius"]), len(bufs), blen, bptr)
    check(lib().soil_hip_restart_write(str(path).encode(), fmt, C.byref(r)))


def read_restart(path, fmt):
    """Reads one RestartData file into numpy arrays."""
    q = _ffi.Restart()
    check(lib().soil_hip_restart_read(str(path).encode(), fmt, C.byref(q)))
    n, nb = q.n, q.nbuf
    lens = (C.c_size_t * max(nb, 1))()
    q2 = _ffi.Restart()
    q2.nbuf = nb
    q2.buf_len = lens
    check(lib().soil_hip_restart_read(str(path).encode(), fmt, C.byref(q2)))
    out = {
        "tag": np.zeros(n, dtype=np.uint32), "type": np.zeros(n, dtype=np.uint32),
        "pos": np.zeros((n, 3)), "vel": np.zeros((n, 3)), "force": np.zeros((n, 3)),
        "mass": np.zeros(n), "radius": np.zeros(n),
    }
    bufs = [np.zeros(lens[k]) for k in range(nb)]
    bptr = (_ffi.DP * max(nb, 1))(*[_dp(b) for b in bufs])
    r = _ffi.Restart(0, 0, 0.0, n, _u32p(out["tag"]), _u32p(out["type"]), _dp(out["pos"]), _dp(out["vel"]),
                     _dp(out["force"]), _dp(out["mass"]), _dp(out["radius"]), nb, lens, bptr)
    check(lib().soil_hip_restart_read(str(path).encode(), fmt, C.byref(r)))
    out.update({"natoms": r.natoms, "step": r.total_cycle, "dt": r.dt, "buffers": bufs})
    return out


def box_write_restart(box, path, fmt, step):
    check(lib().soil_hip_box_write_restart(box.h, str(path).encode(), fmt, int(step)))


def box_read_restart(box, path, fmt):
    """Loads a restart into a new box (in place of upload); returns total_cycle."""
    step = C.c_uint64(0)
    check(lib().soil_hip_box_read_restart(box.h, str(path).encode(), fmt, C.byref(step)))
    return step.value


def box_write_dump(box, path_stem, fmt, step, ghost=False):
    check(lib().soil_hip_box_write_dump(box.h, str(path_stem).encode(), fmt, int(step), 1 if ghost else 0))
