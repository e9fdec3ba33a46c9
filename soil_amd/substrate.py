"""Host-side wrapper of the device substrate, shaped like SOIL's systems.

One `Box` stands in for the resources one SOIL App holds for the substrate
(`Atom`, `Domain`, `Neighbor`, `CommTopology`, `CurrentState<CommState>`,
`DeformState`); its methods are the systems the reference plugins register
(soil_core/src/neighbor.rs:541-566, comm.rs:177-199, domain.rs:441-449,
atom.rs:613-619, soil_verlet/src/lib.rs:103-118, soil_deform/src/lib.rs:249-269).
Everything runs on the device through libsoil_hip.so; there is no CPU path.
"""
import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import BoxDesc, check, lib

FULL_REBUILD = 0
COMMUNICATE_ONLY = 1


def _dp(a):
    return a.ctypes.data_as(_ffi.DP) if a is not None else None


def _u32p(a):
    return a.ctypes.data_as(_ffi.U32P) if a is not None else None


def _i32p(a):
    return a.ctypes.data_as(_ffi.I32P) if a is not None else None


def make_desc(lo, hi, periodic=(1, 1, 1), skin_fraction=1.0, bin_min_size=1.0, every=0, check_=1,
              sort_every=1000, newton=1, dt=1.0, shear_rate=0.0):
    """Box description with the reference defaults of [neighbor] (neighbor.rs:68-79)."""
    d = BoxDesc()
    for k in range(3):
        d.lo[k] = float(lo[k])
        d.hi[k] = float(hi[k])
        d.periodic[k] = int(periodic[k])
    d.skin_fraction = skin_fraction
    d.bin_min_size = bin_min_size
    d.every = every
    d.check = check_
    d.sort_every = sort_every
    d.newton = newton
    d.dt = dt
    d.shear_rate = shear_rate
    d.has_shear = 1 if shear_rate != 0.0 else 0
    return d


def desc_to_dict(d):
    return {
        "lo": list(d.lo), "hi": list(d.hi), "periodic": list(d.periodic), "newton": d.newton,
        "skin_fraction": d.skin_fraction, "bin_min_size": d.bin_min_size, "every": d.every, "check": d.check,
        "sort_every": d.sort_every, "dt": d.dt, "shear_rate": d.shear_rate, "has_shear": d.has_shear,
    }


def device_count():
    n = C.c_int(0)
    lib().soil_hip_device_count(C.byref(n))
    return n.value


class Context:
    def __init__(self, device=0):
        self.h = C.c_void_p()
        check(lib().soil_hip_ctx_create(device, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().soil_hip_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


DEFORM_STYLES = {None: -1, "none": -1, "erate": 0, "vel": 1, "final": 2}
BOUNDARY = {"p": 0, "f": 1, "s": 2}


class Box:
    """One simulation box on the device."""

    def __init__(self, ctx, desc):
        self.ctx = ctx
        self.desc = desc
        self.h = C.c_void_p()
        check(lib().soil_hip_box_create(ctx.h, C.byref(desc), C.byref(self.h)))
        self.columns = []

    def close(self):
        if self.h:
            lib().soil_hip_box_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ── setup ──
    def register_column(self, width, flags):
        cid = C.c_int(-1)
        check(lib().soil_hip_register_column(self.h, width, flags, C.byref(cid)))
        self.columns.append((width, flags))
        return cid.value

    def upload(self, pos, mass, radius, vel=None, force=None, tag=None, type_=None, image=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        radius = np.ascontiguousarray(radius, dtype=np.float64)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        force = None if force is None else np.ascontiguousarray(force, dtype=np.float64).reshape(-1, 3)
        tag = None if tag is None else np.ascontiguousarray(tag, dtype=np.uint32)
        type_ = None if type_ is None else np.ascontiguousarray(type_, dtype=np.uint32)
        image = None if image is None else np.ascontiguousarray(image, dtype=np.int32).reshape(-1, 3)
        check(lib().soil_hip_upload(self.h, n, _dp(pos), _dp(vel), _dp(force), _dp(mass), _dp(radius),
                                    _u32p(tag), _u32p(type_), _i32p(image)))

    def upload_column(self, col, data):
        data = np.ascontiguousarray(data, dtype=np.float64)
        w = self.columns[col][0]
        check(lib().soil_hip_upload_column(self.h, col, _dp(data), data.size // w))

    def setup(self):
        check(lib().soil_hip_setup(self.h))

    def set_consumer(self, k, enabled=True):
        check(lib().soil_hip_set_consumer(self.h, 1 if enabled else 0, float(k)))

    def set_force_hook(self, fn, user=None, transpose=False):
        """Registers a force hook (soil_hip_force_hook; `fn` and `user` are raw
        addresses, e.g. from a ctypes-loaded library) run in the Force phase;
        fn=None removes it. transpose: also build the incoming-pair transpose."""
        flags = _ffi.HOOK_TRANSPOSE if transpose else 0
        check(lib().soil_hip_set_force_hook(self.h, fn, user, flags))

    def device_view(self):
        """The hook view as it stands (counts, capacities, device addresses)."""
        v = _ffi.DeviceView()
        check(lib().soil_hip_get_device_view(self.h, C.byref(v)))
        return v

    def device_bytes(self):
        """(estimate, allocated): device bytes set-up allocates (the server's
        admission figure) and what the arenas hold now."""
        e, a = C.c_size_t(0), C.c_size_t(0)
        check(lib().soil_hip_box_device_bytes(self.h, C.byref(e), C.byref(a)))
        return e.value, a.value

    def compute_ke(self, mask=None):
        """atom.rs:589-604: kinetic energy of the locals (mask: bool per local)."""
        out = C.c_double(0.0)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        check(lib().soil_hip_compute_ke(self.h, None if m is None else m.ctypes.data_as(C.POINTER(C.c_uint8)),
                                        C.byref(out)))
        return out.value

    def classify_boundary_atoms(self):
        """neighbor.rs:1321-1336: True for local rows with a ghost neighbour."""
        n = C.c_size_t(0)
        check(lib().soil_hip_classify_boundary_atoms(self.h, None, 0, C.byref(n)))
        out = np.zeros(n.value, dtype=np.uint8)
        check(lib().soil_hip_classify_boundary_atoms(self.h, out.ctypes.data_as(C.POINTER(C.c_uint8)), out.size,
                                                     C.byref(n)))
        return out.astype(bool)

    def run_force_hook(self):
        check(lib().soil_hip_run_force_hook(self.h))

    def set_boundary(self, dim, kind):
        """[domain] boundary of one axis (domain.rs:27-34): "p" periodic, "f"
        fixed, "s" shrink-wrap. Before setup."""
        check(lib().soil_hip_set_boundary(self.h, int(dim), BOUNDARY[kind]))

    def set_shrink_wrap_padding(self, padding):
        check(lib().soil_hip_set_shrink_wrap_padding(self.h, float(padding)))

    def set_axis_deform(self, dim, style, value=0.0, target=(0.0, 0.0)):
        """[deform] x/y/z (soil_deform/src/lib.rs:98-126): style "erate" (value =
        strain rate), "vel" (value = boundary velocity) or "final" (target =
        (lo, hi) reached at the stage's last step); None removes it. Before setup."""
        check(lib().soil_hip_set_axis_deform(self.h, int(dim), DEFORM_STYLES[style], float(value),
                                             float(target[0]), float(target[1])))

    def set_deform_remap(self, remap):
        check(lib().soil_hip_set_deform_remap(self.h, 1 if remap else 0))

    def set_stage_steps(self, n):
        """Steps in the current stage: the horizon of the "final" style."""
        check(lib().soil_hip_set_stage_steps(self.h, int(n)))

    def domain(self):
        """(lo, hi, size) of the current domain."""
        out = np.zeros(9)
        check(lib().soil_hip_get_domain(self.h, _dp(out)))
        return out[0:3].copy(), out[3:6].copy(), out[6:9].copy()

    # ── systems ──
    def apply_deform(self):
        check(lib().soil_hip_apply_deform(self.h))

    def initial_integration(self):
        check(lib().soil_hip_initial_integration(self.h))

    def decide_rebuild(self):
        fr = C.c_int(0)
        check(lib().soil_hip_decide_rebuild(self.h, C.byref(fr)))
        return bool(fr.value)

    def remove_ghost_atoms(self):
        check(lib().soil_hip_remove_ghost_atoms(self.h))

    def shrink_wrap(self):
        check(lib().soil_hip_shrink_wrap(self.h))

    def zero_all_forces(self):
        check(lib().soil_hip_zero_all_forces(self.h))

    def pbc(self):
        check(lib().soil_hip_pbc(self.h))

    def sort_atoms_by_bin(self):
        check(lib().soil_hip_sort_atoms_by_bin(self.h))

    def borders(self):
        check(lib().soil_hip_borders(self.h))

    def forward_comm_borders(self):
        check(lib().soil_hip_forward_comm_borders(self.h))

    def bin_neighbor_list(self):
        check(lib().soil_hip_bin_neighbor_list(self.h))

    def harmonic_force(self):
        check(lib().soil_hip_harmonic_force(self.h))

    def reverse_send_force(self):
        check(lib().soil_hip_reverse_send_force(self.h))

    def final_integration(self):
        check(lib().soil_hip_final_integration(self.h))

    def step(self, n=1):
        check(lib().soil_hip_step(self.h, int(n)))

    def sync(self):
        check(lib().soil_hip_sync(self.h))

    def set_tilt(self, xy, boundary_vel_x=0.0):
        check(lib().soil_hip_set_tilt(self.h, float(xy), float(boundary_vel_x)))

    # ── readback ──
    def counts(self):
        out = np.zeros(8, dtype=np.uint64)
        check(lib().soil_hip_get_counts(self.h, out.ctypes.data_as(_ffi.U64P)))
        keys = ["nlocal", "nghost", "npairs", "ncells", "rebuilds", "steps", "nstencil_fwd", "state"]
        return {k: int(v) for k, v in zip(keys, out)}

    def csr(self):
        c = self.counts()
        off = np.zeros(c["nlocal"] + 1, dtype=np.uint32)
        idx = np.zeros(max(c["npairs"], 1), dtype=np.uint32)
        n = C.c_size_t(0)
        check(lib().soil_hip_get_csr(self.h, _u32p(off), off.size, _u32p(idx), idx.size, C.byref(n)))
        return off, idx[: n.value]

    def atoms(self):
        c = self.counts()
        n = c["nlocal"] + c["nghost"]
        out = {
            "pos": np.zeros((n, 3)), "vel": np.zeros((n, 3)), "force": np.zeros((n, 3)),
            "radius": np.zeros(n), "mass": np.zeros(n), "inv_mass": np.zeros(n),
            "tag": np.zeros(n, dtype=np.uint32), "origin": np.zeros(n, dtype=np.int32),
            "image": np.zeros((n, 3), dtype=np.int32),
        }
        check(lib().soil_hip_get_atoms(self.h, n, _dp(out["pos"]), _dp(out["vel"]), _dp(out["force"]),
                                       _dp(out["radius"]), _dp(out["mass"]), _dp(out["inv_mass"]),
                                       _u32p(out["tag"]), _i32p(out["origin"]), _i32p(out["image"])))
        out["nlocal"] = c["nlocal"]
        out["nghost"] = c["nghost"]
        return out

    def column(self, col):
        c = self.counts()
        n = c["nlocal"] + c["nghost"]
        w = self.columns[col][0]
        out = np.zeros(n * w)
        check(lib().soil_hip_get_column(self.h, col, _dp(out), n))
        return out

    def bins(self):
        c = self.counts()
        n = c["nlocal"] + c["nghost"]
        bs = np.zeros(c["ncells"] + 1, dtype=np.uint32)
        sa = np.zeros(n, dtype=np.uint32)
        si = np.zeros(n, dtype=np.uint32)
        ac = np.zeros(n, dtype=np.uint32)
        check(lib().soil_hip_get_bins(self.h, _u32p(bs), bs.size, _u32p(sa), _u32p(si), _u32p(ac), n))
        return bs, sa, si, ac

    def swaps(self):
        out = np.zeros((6, 11))
        n = C.c_size_t(0)
        check(lib().soil_hip_get_swaps(self.h, _dp(out), 6, C.byref(n)))
        return out[: n.value]

    def grid(self):
        out = np.zeros(21)
        check(lib().soil_hip_get_grid(self.h, _dp(out)))
        return out

    def stencil(self):
        n = C.c_size_t(0)
        check(lib().soil_hip_get_stencil(self.h, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.int32)
        check(lib().soil_hip_get_stencil(self.h, _i32p(out), out.size, C.byref(n)))
        return out[: n.value]

    def set_timing(self, period):
        """period: False/0 off, True/1 every rebuild, n every n-th rebuild."""
        check(lib().soil_hip_set_timing(self.h, int(period)))

    def timing(self):
        out = np.zeros(5)
        check(lib().soil_hip_get_timing(self.h, _dp(out)))
        return {"builds": out[0], "csr_ms": out[1], "csr_bytes": out[2], "fill_ms": out[3], "count_ms": out[4]}


class Batch:
    """Many boxes stepping together: one launch per phase, one captured graph per step."""

    def __init__(self, ctx, boxes):
        self.ctx = ctx
        self.boxes = list(boxes)
        arr = (C.c_void_p * len(self.boxes))(*[b.h.value for b in self.boxes])
        self.h = C.c_void_p()
        check(lib().soil_hip_batch_create(ctx.h, arr, len(self.boxes), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().soil_hip_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1, graph=True):
        check(lib().soil_hip_batch_step(self.h, int(n), 1 if graph else 0))

    def sync(self):
        check(lib().soil_hip_batch_sync(self.h))

    def autotune(self):
        """Picks the CSR strategy and blocks per launch; returns what it chose and measured."""
        check(lib().soil_hip_batch_autotune(self.h))
        return self.tuning()

    def tuning(self):
        n = C.c_size_t(0)
        check(lib().soil_hip_batch_get_tuning(self.h, None, 0, C.byref(n)))
        out = np.zeros(n.value)
        check(lib().soil_hip_batch_get_tuning(self.h, _dp(out), out.size, C.byref(n)))
        trials = [{"csr_mode": int(out[4 + 3 * t]), "blocks_per_launch": int(out[5 + 3 * t]),
                   "ms_per_step": float(out[6 + 3 * t])} for t in range(int(out[3]))]
        return {"blocks_per_launch": int(out[0]), "csr_mode": int(out[1]), "ms_per_step": float(out[2]),
                "trials": trials}

    def set_timing(self, enabled):
        check(lib().soil_hip_batch_set_timing(self.h, 1 if enabled else 0))

    def timing(self):
        out = np.zeros(4)
        check(lib().soil_hip_batch_get_timing(self.h, _dp(out)))
        return {"builds": out[0], "ms": out[1], "csr_bytes": out[2]}


# ── synthetic inputs (host) ──
def gen_box(config, box_id=0, seed=0x50494C):
    """Box description and atoms of workload `config` (1..5)."""
    d = BoxDesc()
    n = C.c_size_t(0)
    check(lib().soil_hip_gen_box(config, box_id, seed, C.byref(d), C.byref(n), None, None, None, None))
    pos = np.zeros((n.value, 3))
    vel = np.zeros((n.value, 3))
    mass = np.zeros(n.value)
    rad = np.zeros(n.value)
    check(lib().soil_hip_gen_box(config, box_id, seed, C.byref(d), C.byref(n), _dp(pos), _dp(vel), _dp(mass),
                                 _dp(rad)))
    return d, pos, vel, mass, rad


def gen_custom(seed, n, rmin, rmax, phi, aspect=(1.0, 1.0, 1.0), nblobs=0, blob_frac=0.0, sigma=0.0,
               vel_sigma=0.1, shear_rate=0.0):
    d = BoxDesc()
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    mass = np.zeros(n)
    rad = np.zeros(n)
    asp = np.array(aspect, dtype=np.float64)
    check(lib().soil_hip_gen_custom(seed, n, rmin, rmax, phi, _dp(asp), nblobs, blob_frac, sigma, vel_sigma,
                                    shear_rate, C.byref(d), _dp(pos), _dp(vel), _dp(mass), _dp(rad)))
    return d, pos, vel, mass, rad


def gen_arrivals(seed, njobs, rate):
    t = np.zeros(njobs)
    s = np.zeros(njobs, dtype=np.int32)
    check(lib().soil_hip_gen_arrivals(seed, njobs, rate, _dp(t), _i32p(s)))
    return t, s
