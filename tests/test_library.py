"""CPU checks of the product library: it loads, exports every header symbol,
and its host-side generator is deterministic. No GPU call is made here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

import soil_amd
from soil_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "soil_hip.h")).read()
    return sorted(set(re.findall(r"\b(soil_hip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    lib = _ffi.lib()
    syms = header_symbols()
    assert len(syms) >= 40
    for s in syms:
        assert hasattr(lib, s), f"missing export {s}"


def test_ffi_signatures_cover_header():
    assert set(header_symbols()) <= set(_ffi._SIGS)


def test_generator_is_deterministic_and_matches_closed_form():
    d1, p1, v1, m1, r1 = soil_amd.gen_box(1)
    d2, p2, v2, m2, r2 = soil_amd.gen_box(1)
    assert np.array_equal(p1, p2) and np.array_equal(v1, v2)
    # C1: 4000 spheres d=1 at phi=0.30 -> L = cbrt(4000 * pi/6 / 0.3)
    L = (4000 * np.pi / 6 / 0.3) ** (1 / 3)
    assert abs(d1.hi[0] - L) < 1e-9
    assert np.all((p1 >= 0) & (p1 < L))
    assert abs(v1.std() - 0.1) < 0.005
    assert np.all(r1 == 0.5) and np.all(m1 == 1.0)
    assert d1.skin_fraction == 1.1 and d1.dt == 1e-3 and d1.has_shear == 0


def test_ensemble_parameters_in_range():
    ns = []
    for k in range(32):
        d, p, v, m, r = soil_amd.gen_box(4, k)
        ns.append(len(p))
        assert 1000 <= len(p) <= 64000
        assert 0.0 < d.shear_rate <= 0.2 and d.has_shear == 1
        phi = np.sum(4 / 3 * np.pi * r ** 3) / np.prod(np.array(d.hi) - np.array(d.lo))
        assert 0.2 - 1e-9 <= phi <= 0.55 + 1e-9
    assert len(set(ns)) > 20


def test_missing_device_is_an_error_not_a_fallback():
    if soil_amd.device_count() > 0:
        return
    try:
        soil_amd.Context(0)
    except _ffi.SoilHipError as e:
        assert e.code == _ffi.EDEVICE
    else:
        raise AssertionError("context creation must fail without a device")


def _grid_for(desc, rad, xy):
    import ctypes as C
    out = np.zeros(21)
    st = np.zeros(4096, dtype=np.int32)
    n = C.c_size_t(0)
    rad = np.ascontiguousarray(rad, dtype=np.float64)
    _ffi.check(_ffi.lib().soil_hip_grid_for(C.byref(desc), rad.ctypes.data_as(_ffi.DP), rad.size, xy,
                                            out.ctypes.data_as(_ffi.DP), st.ctypes.data_as(_ffi.I32P), st.size,
                                            C.byref(n)))
    return out, st[: n.value]


def test_host_geometry_model_matches_oracle_grids():
    # the product's host geometry (grid, ghost cut, stencil) against the oracle's, orthogonal and tilted
    from oracle_ref import RefBox
    for cfg, tilts in ((1, [0.0]), (4, [0.0, 0.3, -2.0, 5.5])):
        for box_id in range(3):
            d, pos, vel, mass, rad = soil_amd.gen_box(cfg, box_id) if cfg == 4 else soil_amd.gen_box(cfg)
            for xy in tilts:
                d.has_shear = 0
                r = RefBox(d)
                r.add_atoms(pos, mass, rad)
                r.setup()
                if xy != 0.0:
                    r.set_tilt(xy, 0.0)
                    r.phase("sort_atoms_by_bin")
                g_ref = r.grid()
                s_ref = r.stencil(True)
                g, s = _grid_for(d, rad, xy)
                assert np.array_equal(g[:12], g_ref[:12]), (cfg, box_id, xy)
                assert np.array_equal(s, s_ref), (cfg, box_id, xy)


def test_device_view_mirror_matches_the_header(tmp_path):
    # the ctypes mirror of soil_hip_device_view has the C struct's size and
    # field offsets (a C program built against the header prints them)
    fields = [f[0] for f in _ffi.DeviceView._fields_]
    body = "".join(f'printf(" %zu", offsetof(soil_hip_device_view, {f}));' for f in fields)
    src = tmp_path / "view.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "soil_hip.h"\n'
                   'int main(void) { printf("%zu", sizeof(soil_hip_device_view)); ' + body + " return 0; }\n")
    exe = tmp_path / "view"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = [int(t) for t in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == C.sizeof(_ffi.DeviceView)
    for name, off in zip(fields, out[1:]):
        assert getattr(_ffi.DeviceView, name).offset == off, name


def test_server_stat_keys_match_the_header():
    from soil_amd import serve
    text = open(os.path.join(ROOT, "include", "soil_hip.h")).read()
    n = int(re.search(r"#define SOIL_HIP_SERVER_NSTATS (\d+)", text).group(1))
    assert len(serve.STAT_KEYS) == n == serve.NSTATS
