"""ctypes binding of the CPU oracle (oracle/libsoil_ref.so), for tests only."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# SOIL_REF_LIB points the tests at another build of the same source (for
# example the sanitizer build, `make -C oracle asan`); it is used as built
LIB_PATH = os.environ.get("SOIL_REF_LIB") or os.path.join(ORACLE_DIR, "libsoil_ref.so")


class RefDesc(C.Structure):
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


_lib = None
DP = C.POINTER(C.c_double)
U32P = C.POINTER(C.c_uint32)
I32P = C.POINTER(C.c_int32)
U64P = C.POINTER(C.c_uint64)
P = C.c_void_p


def build_oracle():
    if os.environ.get("SOIL_REF_LIB"):
        return
    src = os.path.join(ORACLE_DIR, "soil_ref.cpp")
    if not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)


def lib():
    global _lib
    if _lib is None:
        build_oracle()
        h = C.CDLL(LIB_PATH)
        sigs = {
            "soil_ref_create": ([C.POINTER(RefDesc)], P),
            "soil_ref_destroy": ([P], None),
            "soil_ref_add_atoms": ([P, C.c_size_t, DP, DP, DP, DP, DP, U32P, U32P, I32P], C.c_int),
            "soil_ref_register_column": ([P, C.c_int, C.c_uint], C.c_int),
            "soil_ref_set_column": ([P, C.c_int, DP, C.c_size_t], C.c_int),
            "soil_ref_get_column": ([P, C.c_int, DP, C.c_size_t], C.c_size_t),
            "soil_ref_set_consumer": ([P, C.c_int, C.c_double], None),
            "soil_ref_set_axis_deform": ([P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double], C.c_int),
            "soil_ref_set_deform_remap": ([P, C.c_int], None),
            "soil_ref_set_stage_steps": ([P, C.c_uint64], None),
            "soil_ref_get_domain": ([P, DP], None),
            "soil_ref_classify_csr": ([U32P, U32P, C.c_size_t, C.POINTER(C.c_uint8)], None),
            "soil_ref_classify_boundary_atoms": ([P, C.POINTER(C.c_uint8)], None),
            "soil_ref_compute_ke": ([P, C.POINTER(C.c_uint8)], C.c_double),
            "soil_ref_set_boundary": ([P, C.c_int, C.c_int], C.c_int),
            "soil_ref_set_shrink_wrap_padding": ([P, C.c_double], None),
            "soil_ref_setup": ([P], C.c_int),
            "soil_ref_step": ([P, C.c_int], C.c_int),
            "soil_ref_phase": ([P, C.c_int], C.c_int),
            "soil_ref_get_state_flag": ([P], C.c_int),
            "soil_ref_set_state_flag": ([P, C.c_int], None),
            "soil_ref_counts": ([P, U64P], None),
            "soil_ref_grid": ([P, DP], None),
            "soil_ref_stencil": ([P, C.c_int, I32P, C.c_size_t], C.c_size_t),
            "soil_ref_csr": ([P, U32P, U32P, C.c_size_t], C.c_size_t),
            "soil_ref_bins": ([P, U32P, U32P, U32P, U32P], None),
            "soil_ref_bins_size": ([P], C.c_size_t),
            "soil_ref_sort_perm": ([P, U64P], C.c_size_t),
            "soil_ref_atoms": ([P, DP, DP, DP, DP, DP, DP, U32P, I32P, I32P], None),
            "soil_ref_swaps": ([P, DP], C.c_size_t),
            "soil_ref_swap_send": ([P, C.c_int, U64P], C.c_size_t),
            "soil_ref_set_tilt": ([P, C.c_double, C.c_double], None),
            "soil_ref_set_ghost_cutoff": ([P, C.c_double], None),
            "soil_ref_x2lamda": ([P, DP, DP], None),
            "soil_ref_lamda2x": ([P, DP, DP], None),
            "soil_ref_set_vel": ([P, C.c_size_t, DP], None),
            "soil_ref_set_force": ([P, C.c_size_t, DP], None),
            "soil_ref_set_pos": ([P, C.c_size_t, DP], None),
            "soil_ref_time_steps": ([P, C.c_int], C.c_double),
        }
        for name, (args, res) in sigs.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = res
        _lib = h
    return _lib


# phase ids of soil_ref_phase
PH = dict(apply_deform=0, initial_integration=1, decide_rebuild=2, remove_ghost_atoms=3, zero_all_forces=4,
          pbc=5, sort_atoms_by_bin=6, borders=7, forward_comm_borders=8, bin_neighbor_list=9, harmonic_force=10,
          reverse_send_force=11, final_integration=12, shrink_wrap=13)
BOUNDARY = {"p": 0, "f": 1, "s": 2}


def _dp(a):
    return a.ctypes.data_as(DP) if a is not None else None


def desc_from(d):
    """RefDesc from a soil_amd BoxDesc or a dict with the same fields."""
    r = RefDesc()
    get = (lambda k: getattr(d, k)) if not isinstance(d, dict) else (lambda k: d[k])
    for k in range(3):
        r.lo[k] = get("lo")[k]
        r.hi[k] = get("hi")[k]
        r.periodic[k] = get("periodic")[k]
    for f in ("newton", "skin_fraction", "bin_min_size", "every", "check", "sort_every", "dt", "shear_rate",
              "has_shear"):
        setattr(r, f, get(f))
    return r


DEFORM_STYLES = {None: -1, "none": -1, "erate": 0, "vel": 1, "final": 2}


class RefBox:
    def __init__(self, desc):
        self.desc = desc_from(desc)
        self.h = lib().soil_ref_create(C.byref(self.desc))

    def __del__(self):
        try:
            lib().soil_ref_destroy(self.h)
        except Exception:
            pass

    def add_atoms(self, pos, mass, radius, vel=None, force=None, tag=None, image=None, type_=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        radius = np.ascontiguousarray(radius, dtype=np.float64)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        force = None if force is None else np.ascontiguousarray(force, dtype=np.float64).reshape(-1, 3)
        tag = None if tag is None else np.ascontiguousarray(tag, dtype=np.uint32)
        image = None if image is None else np.ascontiguousarray(image, dtype=np.int32).reshape(-1, 3)
        type_ = None if type_ is None else np.ascontiguousarray(type_, dtype=np.uint32)
        rc = lib().soil_ref_add_atoms(self.h, n, _dp(pos), _dp(vel), _dp(force), _dp(mass), _dp(radius),
                                      tag.ctypes.data_as(U32P) if tag is not None else None,
                                      type_.ctypes.data_as(U32P) if type_ is not None else None,
                                      image.ctypes.data_as(I32P) if image is not None else None)
        assert rc == 0

    def register_column(self, width, flags):
        return lib().soil_ref_register_column(self.h, width, flags)

    def set_column(self, col, data, width):
        data = np.ascontiguousarray(data, dtype=np.float64)
        assert lib().soil_ref_set_column(self.h, col, _dp(data), data.size // width) == 0

    def column(self, col):
        n = lib().soil_ref_get_column(self.h, col, None, 0)
        out = np.zeros(n)
        lib().soil_ref_get_column(self.h, col, _dp(out), n)
        return out

    def set_consumer(self, k, enabled=True):
        lib().soil_ref_set_consumer(self.h, 1 if enabled else 0, float(k))

    def set_axis_deform(self, dim, style, value=0.0, target=(0.0, 0.0)):
        """[deform] x/y/z: style "erate" | "vel" | "final" (soil_deform/src/lib.rs:98-126)."""
        assert lib().soil_ref_set_axis_deform(self.h, int(dim), DEFORM_STYLES[style], float(value),
                                              float(target[0]), float(target[1])) == 0

    def set_deform_remap(self, remap):
        lib().soil_ref_set_deform_remap(self.h, 1 if remap else 0)

    def set_stage_steps(self, n):
        lib().soil_ref_set_stage_steps(self.h, int(n))

    def compute_ke(self, mask=None):
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        return lib().soil_ref_compute_ke(self.h, None if m is None else m.ctypes.data_as(C.POINTER(C.c_uint8)))

    def classify_boundary_atoms(self):
        """neighbor.rs:1321-1336 on the current list."""
        out = np.zeros(self.counts()["nlocal"], dtype=np.uint8)
        lib().soil_ref_classify_boundary_atoms(self.h, out.ctypes.data_as(C.POINTER(C.c_uint8)))
        return out.astype(bool)

    def set_boundary(self, dim, kind):
        """[domain] boundary of one axis: "p" periodic, "f" fixed, "s" shrink-wrap."""
        assert lib().soil_ref_set_boundary(self.h, int(dim), BOUNDARY[kind]) == 0

    def set_shrink_wrap_padding(self, padding):
        lib().soil_ref_set_shrink_wrap_padding(self.h, float(padding))

    def domain(self):
        out = np.zeros(9)
        lib().soil_ref_get_domain(self.h, _dp(out))
        return out[0:3].copy(), out[3:6].copy(), out[6:9].copy()

    def setup(self):
        assert lib().soil_ref_setup(self.h) == 0

    def step(self, n=1):
        """Steps n times; raises when an axis collapses (soil_deform/src/lib.rs:506-515)."""
        if lib().soil_ref_step(self.h, int(n)) != 0:
            raise RuntimeError("[deform] an axis collapsed to zero size")

    def phase(self, name):
        assert lib().soil_ref_phase(self.h, PH[name]) == 0

    def time_steps(self, n):
        return lib().soil_ref_time_steps(self.h, int(n))

    @property
    def state(self):
        return lib().soil_ref_get_state_flag(self.h)

    @state.setter
    def state(self, v):
        lib().soil_ref_set_state_flag(self.h, int(v))

    def counts(self):
        out = np.zeros(8, dtype=np.uint64)
        lib().soil_ref_counts(self.h, out.ctypes.data_as(U64P))
        keys = ["nlocal", "nghost", "npairs", "ncells", "rebuilds", "steps", "nstencil_fwd", "nswaps"]
        return {k: int(v) for k, v in zip(keys, out)}

    def grid(self):
        out = np.zeros(21)
        lib().soil_ref_grid(self.h, _dp(out))
        return out

    def stencil(self, forward=True):
        n = lib().soil_ref_stencil(self.h, 1 if forward else 0, None, 0)
        out = np.zeros(max(n, 1), dtype=np.int32)
        lib().soil_ref_stencil(self.h, 1 if forward else 0, out.ctypes.data_as(I32P), n)
        return out[:n]

    def csr(self):
        c = self.counts()
        off = np.zeros(c["nlocal"] + 1, dtype=np.uint32)
        idx = np.zeros(max(c["npairs"], 1), dtype=np.uint32)
        n = lib().soil_ref_csr(self.h, off.ctypes.data_as(U32P), idx.ctypes.data_as(U32P), idx.size)
        return off, idx[:n]

    def bins(self):
        c = self.counts()
        n = c["nlocal"] + c["nghost"]
        nb = lib().soil_ref_bins_size(self.h)
        bs = np.zeros(max(nb, 1), dtype=np.uint32)
        sa = np.zeros(max(n, 1), dtype=np.uint32)
        si = np.zeros(max(n, 1), dtype=np.uint32)
        ac = np.zeros(max(n, 1), dtype=np.uint32)
        lib().soil_ref_bins(self.h, bs.ctypes.data_as(U32P), sa.ctypes.data_as(U32P), si.ctypes.data_as(U32P),
                            ac.ctypes.data_as(U32P))
        return bs[:nb], sa[:n], si[:n], ac[:n]

    def sort_perm(self):
        n = lib().soil_ref_sort_perm(self.h, None)
        out = np.zeros(max(n, 1), dtype=np.uint64)
        n = lib().soil_ref_sort_perm(self.h, out.ctypes.data_as(U64P))
        return out[:n] if n else None

    def atoms(self):
        c = self.counts()
        n = c["nlocal"] + c["nghost"]
        out = {
            "pos": np.zeros((n, 3)), "vel": np.zeros((n, 3)), "force": np.zeros((n, 3)),
            "radius": np.zeros(n), "mass": np.zeros(n), "inv_mass": np.zeros(n),
            "tag": np.zeros(n, dtype=np.uint32), "origin": np.zeros(n, dtype=np.int32),
            "image": np.zeros((n, 3), dtype=np.int32),
        }
        lib().soil_ref_atoms(self.h, _dp(out["pos"]), _dp(out["vel"]), _dp(out["force"]), _dp(out["radius"]),
                             _dp(out["mass"]), _dp(out["inv_mass"]), out["tag"].ctypes.data_as(U32P),
                             out["origin"].ctypes.data_as(I32P), out["image"].ctypes.data_as(I32P))
        out["nlocal"] = c["nlocal"]
        out["nghost"] = c["nghost"]
        return out

    def swaps(self):
        out = np.zeros((16, 11))
        n = lib().soil_ref_swaps(self.h, _dp(out))
        return out[:n]

    def swap_send(self, s):
        n = int(self.swaps()[s][2])
        out = np.zeros(max(n, 1), dtype=np.uint64)
        lib().soil_ref_swap_send(self.h, s, out.ctypes.data_as(U64P))
        return out[:n]

    def set_tilt(self, xy, bvx=0.0):
        lib().soil_ref_set_tilt(self.h, float(xy), float(bvx))

    def set_ghost_cutoff(self, gc):
        lib().soil_ref_set_ghost_cutoff(self.h, float(gc))

    def x2lamda(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        out = np.zeros(3)
        lib().soil_ref_x2lamda(self.h, _dp(r), _dp(out))
        return out

    def lamda2x(self, lam):
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        out = np.zeros(3)
        lib().soil_ref_lamda2x(self.h, _dp(lam), _dp(out))
        return out

    def set_vel(self, i, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        lib().soil_ref_set_vel(self.h, i, _dp(v))

    def set_force(self, i, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        lib().soil_ref_set_force(self.h, i, _dp(f))

    def set_pos(self, i, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        lib().soil_ref_set_pos(self.h, i, _dp(p))


def ref_desc(lo, hi, periodic=(1, 1, 1), skin_fraction=1.0, bin_min_size=1.0, every=0, check=1, sort_every=1000,
             newton=1, dt=1.0, shear_rate=0.0):
    return dict(lo=list(lo), hi=list(hi), periodic=list(periodic), newton=newton, skin_fraction=skin_fraction,
                bin_min_size=bin_min_size, every=every, check=check, sort_every=sort_every, dt=dt,
                shear_rate=shear_rate, has_shear=1 if shear_rate != 0.0 else 0)


def classify_csr(offsets, indices, nlocal):
    """neighbor.rs:1321-1336 on a given CSR."""
    off = np.ascontiguousarray(offsets, dtype=np.uint32)
    idx = np.ascontiguousarray(indices, dtype=np.uint32)
    out = np.zeros(nlocal, dtype=np.uint8)
    lib().soil_ref_classify_csr(off.ctypes.data_as(U32P), idx.ctypes.data_as(U32P), nlocal,
                                out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out.astype(bool)
