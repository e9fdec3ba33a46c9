"""ctypes binding of libsoil_hip.so (the C ABI declared in include/soil_hip.h).

The library is built in-tree by `python -m soil_amd.build`. Loading fails
loudly when it is missing: there is no CPU fallback for the device path.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libsoil_hip.so")

OK = 0
EINVAL = -1
EDEVICE = -2
ECAPACITY = -3
ESTATE = -4
EUNSUPPORTED = -5
EBUDGET = -6

COL_FORWARD = 1
COL_REVERSE = 2
COL_ZERO = 4


class BoxDesc(C.Structure):
    """soil_hip_box_desc (include/soil_hip.h)."""

    _fields_ = [
        ("lo", C.c_double * 3),
        ("hi", C.c_double * 3),
        ("periodic", C.c_int32 * 3),
        ("newton", C.c_int32),
        ("skin_fraction", C.c_double),
        ("bin_min_size", C.c_double),
        ("every", C.c_uint64),
        ("check", C.c_int32),
        ("pad0", C.c_int32),
        ("sort_every", C.c_uint64),
        ("dt", C.c_double),
        ("shear_rate", C.c_double),
        ("has_shear", C.c_int32),
        ("pad1", C.c_int32),
    ]


class SoilHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"soil_hip error {code}: {msg}")
        self.code = code


_lib = None

P = C.c_void_p
DP = C.POINTER(C.c_double)
U32P = C.POINTER(C.c_uint32)
I32P = C.POINTER(C.c_int32)
U64P = C.POINTER(C.c_uint64)
SZ = C.c_size_t
SZP = C.POINTER(C.c_size_t)


class Restart(C.Structure):
    """soil_hip_restart (host view of one RestartData)."""
    _fields_ = [("natoms", C.c_uint64), ("total_cycle", C.c_uint64), ("dt", C.c_double), ("n", C.c_size_t),
                ("tag", U32P), ("type", U32P), ("pos", DP), ("vel", DP), ("force", DP), ("mass", DP),
                ("radius", DP), ("nbuf", C.c_size_t), ("buf_len", SZP), ("buf", C.POINTER(DP))]


MAX_COLS = 8


class DeviceView(C.Structure):
    """soil_hip_device_view (a force hook's view of a box; device pointers)."""
    _fields_ = ([(n, C.c_uint32) for n in ("nlocal", "nghost", "npairs", "rebuilt", "cap_local", "cap_atoms",
                                           "cap_pairs", "ncols")]
                + [(n, C.c_void_p) for n in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "xr", "mass",
                                             "inv_mass", "radius", "tag", "type", "origin", "offsets", "indices",
                                             "tstart", "tsrc")]
                + [("cols", C.c_void_p * MAX_COLS), ("col_width", C.c_int32 * MAX_COLS), ("step", C.c_uint64)])


HOOK_TRANSPOSE = 1
RESTART_BINCODE, RESTART_JSON = 0, 1
DUMP_TEXT, DUMP_BINARY, DUMP_LAMMPS = 0, 1, 2

_SIGS = {
    "soil_hip_last_error": ([], C.c_char_p),
    "soil_hip_device_count": ([C.POINTER(C.c_int)], C.c_int),
    "soil_hip_ctx_create": ([C.c_int, C.POINTER(P)], C.c_int),
    "soil_hip_ctx_destroy": ([P], C.c_int),
    "soil_hip_box_create": ([P, C.POINTER(BoxDesc), C.POINTER(P)], C.c_int),
    "soil_hip_box_destroy": ([P], C.c_int),
    "soil_hip_register_column": ([P, C.c_int, C.c_uint32, C.POINTER(C.c_int)], C.c_int),
    "soil_hip_upload": ([P, SZ, DP, DP, DP, DP, DP, U32P, U32P, I32P], C.c_int),
    "soil_hip_upload_column": ([P, C.c_int, DP, SZ], C.c_int),
    "soil_hip_setup": ([P], C.c_int),
    "soil_hip_set_consumer": ([P, C.c_int, C.c_double], C.c_int),
    "soil_hip_set_force_hook": ([P, C.c_void_p, C.c_void_p, C.c_int], C.c_int),
    "soil_hip_get_device_view": ([P, C.POINTER(DeviceView)], C.c_int),
    "soil_hip_run_force_hook": ([P], C.c_int),
    "soil_hip_classify_boundary_atoms": ([P, C.POINTER(C.c_uint8), SZ, C.POINTER(SZ)], C.c_int),
    "soil_hip_box_device_bytes": ([P, SZP, SZP], C.c_int),
    "soil_hip_compute_ke": ([P, C.POINTER(C.c_uint8), C.POINTER(C.c_double)], C.c_int),
    "soil_hip_compute_ke_noalloc": ([P, C.POINTER(C.c_uint8), C.POINTER(C.c_double)], C.c_int),
    "soil_hip_set_axis_deform": ([P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double], C.c_int),
    "soil_hip_set_deform_remap": ([P, C.c_int], C.c_int),
    "soil_hip_set_stage_steps": ([P, C.c_uint64], C.c_int),
    "soil_hip_get_domain": ([P, DP], C.c_int),
    "soil_hip_apply_deform": ([P], C.c_int),
    "soil_hip_initial_integration": ([P], C.c_int),
    "soil_hip_decide_rebuild": ([P, C.POINTER(C.c_int)], C.c_int),
    "soil_hip_remove_ghost_atoms": ([P], C.c_int),
    "soil_hip_shrink_wrap": ([P], C.c_int),
    "soil_hip_set_boundary": ([P, C.c_int, C.c_int], C.c_int),
    "soil_hip_set_shrink_wrap_padding": ([P, C.c_double], C.c_int),
    "soil_hip_zero_all_forces": ([P], C.c_int),
    "soil_hip_pbc": ([P], C.c_int),
    "soil_hip_sort_atoms_by_bin": ([P], C.c_int),
    "soil_hip_borders": ([P], C.c_int),
    "soil_hip_forward_comm_borders": ([P], C.c_int),
    "soil_hip_bin_neighbor_list": ([P], C.c_int),
    "soil_hip_harmonic_force": ([P], C.c_int),
    "soil_hip_reverse_send_force": ([P], C.c_int),
    "soil_hip_final_integration": ([P], C.c_int),
    "soil_hip_step": ([P, C.c_int], C.c_int),
    "soil_hip_sync": ([P], C.c_int),
    "soil_hip_set_tilt": ([P, C.c_double, C.c_double], C.c_int),
    "soil_hip_get_counts": ([P, U64P], C.c_int),
    "soil_hip_get_csr": ([P, U32P, SZ, U32P, SZ, SZP], C.c_int),
    "soil_hip_get_atoms": ([P, SZ, DP, DP, DP, DP, DP, DP, U32P, I32P, I32P], C.c_int),
    "soil_hip_get_column": ([P, C.c_int, DP, SZ], C.c_int),
    "soil_hip_get_types": ([P, U32P, SZ], C.c_int),
    "soil_hip_restart_write": ([C.c_char_p, C.c_int, C.POINTER(Restart)], C.c_int),
    "soil_hip_restart_read": ([C.c_char_p, C.c_int, C.POINTER(Restart)], C.c_int),
    "soil_hip_box_write_restart": ([P, C.c_char_p, C.c_int, C.c_uint64], C.c_int),
    "soil_hip_box_read_restart": ([P, C.c_char_p, C.c_int, U64P], C.c_int),
    "soil_hip_box_write_dump": ([P, C.c_char_p, C.c_int, C.c_uint64, C.c_int], C.c_int),
    "soil_hip_get_bins": ([P, U32P, SZ, U32P, U32P, U32P, SZ], C.c_int),
    "soil_hip_get_swaps": ([P, DP, SZ, SZP], C.c_int),
    "soil_hip_get_grid": ([P, DP], C.c_int),
    "soil_hip_get_stencil": ([P, I32P, SZ, SZP], C.c_int),
    "soil_hip_grid_for": ([C.POINTER(BoxDesc), DP, SZ, C.c_double, DP, I32P, SZ, SZP], C.c_int),
    "soil_hip_batch_create": ([P, C.POINTER(P), C.c_int, C.POINTER(P)], C.c_int),
    "soil_hip_batch_destroy": ([P], C.c_int),
    "soil_hip_batch_autotune": ([P], C.c_int),
    "soil_hip_batch_get_tuning": ([P, DP, SZ, SZP], C.c_int),
    "soil_hip_batch_step": ([P, C.c_int, C.c_int], C.c_int),
    "soil_hip_batch_sync": ([P], C.c_int),
    "soil_hip_set_timing": ([P, C.c_int], C.c_int),
    "soil_hip_get_timing": ([P, DP], C.c_int),
    "soil_hip_batch_set_timing": ([P, C.c_int], C.c_int),
    "soil_hip_batch_get_timing": ([P, DP], C.c_int),
    "soil_hip_gen_box": ([C.c_int, C.c_int, C.c_uint64, C.POINTER(BoxDesc), SZP, DP, DP, DP, DP], C.c_int),
    "soil_hip_gen_custom": ([C.c_uint64, SZ, C.c_double, C.c_double, C.c_double, DP, C.c_int, C.c_double,
                             C.c_double, C.c_double, C.c_double, C.POINTER(BoxDesc), DP, DP, DP, DP], C.c_int),
    "soil_hip_gen_arrivals": ([C.c_uint64, C.c_int, C.c_double, DP, I32P], C.c_int),
    "soil_hip_probe_copy": ([P, SZ, C.c_int, DP], C.c_int),
    "soil_hip_server_create": ([P, C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(P)], C.c_int),
    "soil_hip_server_start": ([P], C.c_int),
    "soil_hip_server_submit": ([P, C.POINTER(BoxDesc), SZ, DP, DP, DP, DP, C.c_int, C.c_double, C.c_int,
                                C.POINTER(C.c_int)], C.c_int),
    "soil_hip_server_submit_restart": ([P, C.POINTER(BoxDesc), C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_int,
                                        C.POINTER(C.c_int)], C.c_int),
    "soil_hip_server_set_output": ([P, C.c_char_p, C.c_int], C.c_int),
    "soil_hip_server_evict": ([P, C.c_int, C.c_char_p, C.c_int], C.c_int),
    "soil_hip_server_drain": ([P], C.c_int),
    "soil_hip_server_stats": ([P, DP], C.c_int),
    "soil_hip_server_result": ([P, C.c_int, DP, DP, SZ, SZP], C.c_int),
    "soil_hip_server_destroy": ([P], C.c_int),
    "soil_hip_server_bench_stream": ([P, C.c_int, C.c_double, C.c_uint64, C.c_double, C.c_int, C.c_int, C.c_int,
                                      DP], C.c_int),
}



def lib():
    """Loads libsoil_hip.so once. Raises if the in-tree build is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m soil_amd.build` "
                          "(the device path has no CPU fallback)")
    h = C.CDLL(LIB_PATH)
    for name, (args, res) in _SIGS.items():
        fn = getattr(h, name)
        fn.argtypes = args
        fn.restype = res
    _lib = h
    return h


def check(rc):
    if rc != OK:
        msg = lib().soil_hip_last_error()
        raise SoilHipError(rc, msg.decode() if msg else "")
    return rc
