"""Algorithmic HBM bytes of the substrate's phases (the measurement model).

N locals, G ghosts, C cells, P half pairs; `poly` is 1 for polydisperse radii
(the candidate record then carries a radius the uniform test does not read).
Each phase counts the bytes its inputs and outputs must cross HBM once; the
kernels' measured traffic (rocprofv3 FETCH_SIZE / WRITE_SIZE) is compared
against these figures.
"""

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def csr_bytes(N, G, C, P, poly=True):
    """The CSR build: each sorted candidate record (x, y, z [, r]) and index
    read once, bin_start read once, the offsets and the row entries written
    once (a count-then-fill build's second pass is not counted)."""
    return (24 + 8 * int(poly) + 4) * (N + G) + 4 * (C + 1) + 4 * (N + 1) + 4 * P


def step_bytes(N, G, P):
    """A step without a rebuild: velocity Verlet (128 N + 80 N), the
    displacement check (48 N), force zeroing (24 (N + G)), forward comm
    (100 G) and reverse comm (76 G)."""
    return 280 * N + 200 * G


def rebuild_bytes(N, G, C, P, poly=True, lees_edwards=False):
    """What a rebuild adds: pbc (72 N, + 48 N of velocities under
    Lees-Edwards), the cell keys and the permutation of ~141 B of columns per
    atom read and written plus last_build_pos (366 N), the border scans (24 N)
    and ghost writes (282 G), binning and the sorted gather (56 (N + G)), the
    CSR build, and save_build_positions (48 N)."""
    pbc = 72 * N + (48 * N if lees_edwards else 0)
    return pbc + 366 * N + 24 * N + 282 * G + 56 * (N + G) + csr_bytes(N, G, C, P, poly) + 48 * N


def consumer_bytes(N, P):
    """The harmonic test consumer: 84 B per atom and per pair."""
    return 84 * (N + P)


def run_bytes(N, G, C, P, steps, rebuilds, poly=True, lees_edwards=False, consumer=True):
    """Bytes of `steps` steps of which `rebuilds` rebuild."""
    b = steps * step_bytes(N, G, P) + rebuilds * rebuild_bytes(N, G, C, P, poly, lees_edwards)
    if consumer:
        b += steps * consumer_bytes(N, P)
    return b


def roofline(bytes_, seconds):
    gbs = bytes_ / seconds / 1e9
    return {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}
