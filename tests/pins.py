"""Invariants that the reference's text fixes, checked on any box's outputs.

They pin what bit-for-bit agreement between the device and the oracle
cannot: a misreading shared by both would pass a device-vs-oracle test, but
not these. Each states the reference's rule directly:

* forward comm (crates/soil_core/src/comm.rs:400-432, 496-592): on a
  CommunicateOnly step every ghost is its source atom plus the swap's offset
  FIXED at the last build, pos and vel, applied swap by swap, so a corner
  ghost adds onto its already refreshed source ghost;
* the AtomData column contract (docs/SOIL_ATOMDATA_CONTRACT.md:87-143,
  soil_derive/src/lib.rs:392-500): #[forward] columns of a ghost are its
  source's values, #[reverse] ghost values are added onto their owners (a
  ghost of a ghost first onto its source), #[zero] columns are zero after
  zero_all_forces, and every column follows its atom through the sorts;
* the rebuild schedule (crates/soil_core/src/neighbor.rs:762-784, 842-848,
  989-996): with `every` = E and `check` off, rebuilds land on steps
  1, E+2, 2E+3, ...; a periodic sort at `sort_every` forces a rebuild
  exactly when the stable cell order of the locals is not the identity;
* the pair cutoff's two forms (crates/soil_core/src/neighbor.rs:679-688,
  :1175-1304): when the locals' radii span less than 1e-15 the build tests
  r2 < c*c with c = 2.0 * min_r * sf, otherwise r2 < (ri+rj)*(ri+rj)*(sf*sf);
  the two round differently, so a pair can sit between them;
* the squared distance of the build's test (neighbor.rs:1194-1197 and the
  same lines of the other three loops): dx.mul_add(dx, dy.mul_add(dy, dz*dz)),
  two fused multiply-adds nested in that order, which rounds differently
  from a plain sum or the reverse nesting near the cutoff;
* the displacement trigger (crates/soil_core/src/neighbor.rs:693-758): the
  threshold is t*t with t = (sf - 1) * min_r; a fully periodic box compares
  dx.mul_add(dx, dy.mul_add(dy, dz*dz)), a box with a non-periodic axis the
  plain dx*dx + dy*dy + dz*dz;
* the periodic wrap (crates/soil_core/src/domain.rs:593-604): the upper
  bound is low + size computed at the wrap, which need not equal the box's
  high bound (size = high - low rounds);
* the ghost selection faces (crates/soil_core/src/comm.rs:336-347): an atom
  is sent by the low swap when pos < lo + gc and by the high swap when
  pos >= hi - gc, the thresholds rounded once, which a form like
  pos - lo < gc does not reproduce at every position;
* the Lees-Edwards tilt (crates/soil_deform/src/lib.rs:579-580):
  raw = rate * dt * step * ly, multiplied left to right, then
  xy = raw - lx * round(raw / lx);
* the triclinic wrap (crates/soil_core/src/domain.rs:211-229, 686-712):
  every local goes through x2lamda, the lamda wrap and lamda2x, wrapped or
  not (the round trip can move it by an ulp), with x2lamda subtracting left
  to right, x - x0 - xy*ly - xz*lz;
* reverse comm's order (crates/soil_core/src/comm.rs:962-1021): rounds
  (dims) from the last to the first, in each the low swap and then the high
  swap adding their ghosts onto their sources in receive order, so a ghost
  that is the source of later ghosts has their values before its own is
  added on; with inexact values the order shows;
* the one-time streaming profile (crates/soil_deform/src/lib.rs:568-574):
  v_x += rate * (y - 0.5 * (ylo + yhi)), the difference taken first;
* the triclinic ghost faces (crates/soil_core/src/comm.rs:305-320): in
  lamda, the low x face's threshold is lo + (gc / Lx) * (1 + |xy| / Ly),
  the quotient taken before the inflation;
* the cell of a position (crates/soil_core/src/neighbor.rs:886-888,
  915-922): floor((p - origin) * inv) with inv = 1.0 / bin_size computed
  once, which differs from dividing by bin_size for positions at some cell
  faces;
* velocity Verlet's kick and drift (crates/soil_verlet/src/lib.rs:143-160,
  183-192): h = 0.5 * dt * inv_mass left to right, v0 = v + h*f and
  x = x + v0*dt, no fused multiply-add anywhere;
* a periodic sort on a CommunicateOnly step (neighbor.rs:845-848, 989-996;
  the run_if guards of pbc and remove_ghost_atoms, domain.rs:441-449,
  atom.rs:613-619): the sort sets FullRebuild after pbc was skipped, so the
  step rebuilds from unwrapped positions and an atom past a periodic face
  stays there, its image unchanged, until a later FullRebuild's pbc;
* the CSR row order (crates/soil_core/src/neighbor.rs:1175-1304, stencil
  generation :371-389 and :461-492): row i lists its own cell first (the
  half list only the records after i's own, the full list all but i), then
  the stencil's cells in generation order, the dx, dy, dz loops nested
  outermost to innermost, each cell's records in bin order. The pair SET is
  pinned by brute force elsewhere; this fixes where each pair sits.
"""
import numpy as np


def ghost_swaps(swaps, nlocal, nghost):
    """Swap index of every ghost (rows of the swap table: recv_start, recv_count, ...)."""
    sw = np.full(nghost, -1, dtype=np.int64)
    for s, row in enumerate(swaps):
        start, count = int(row[0]), int(row[1])
        if count:
            assert nlocal <= start and start + count <= nlocal + nghost
            sw[start - nlocal:start - nlocal + count] = s
    assert (sw >= 0).all(), "a ghost belongs to no swap"
    return sw


def chain_depth(origin, nlocal):
    """Swaps between each ghost and its local root (locals: 0)."""
    n = len(origin)
    depth = np.zeros(n, dtype=np.int64)
    for g in range(nlocal, n):  # a ghost's source precedes it, so its depth is known
        src = int(origin[g])
        assert 0 <= src < g
        depth[g] = depth[src] + 1
    return depth


def chain_roots(origin, nlocal):
    n = len(origin)
    root = np.arange(n)
    for g in range(nlocal, n):
        root[g] = root[int(origin[g])]
    return root


def check_forward(a, swaps):
    """Every ghost = source + its swap's fixed offset (pos and vel), exactly.
    Returns the chain depth of every atom."""
    nl, ng = a["nlocal"], a["nghost"]
    sw = ghost_swaps(swaps, nl, ng)
    off = swaps[:, 5:8]
    voff = swaps[:, 8:11]
    src = a["origin"][nl:].astype(np.int64)
    want_p = a["pos"][src] + off[sw]
    want_v = a["vel"][src] + voff[sw]
    bad = np.nonzero((a["pos"][nl:] != want_p).any(axis=1) | (a["vel"][nl:] != want_v).any(axis=1))[0]
    assert bad.size == 0, f"ghosts {bad[:5] + nl} are not source + fixed offset"
    # the same, unrolled from the root: root + offsets of the chain, added root first
    depth = chain_depth(a["origin"], nl)
    for g in np.nonzero(depth >= 2)[0][:200]:
        chain = []
        x = int(g)
        while x >= nl:
            chain.append(x)
            x = int(a["origin"][x])
        p = a["pos"][x].copy()
        for y in reversed(chain):
            p = p + off[sw[y - nl]]
        assert np.array_equal(p, a["pos"][g]), f"ghost {g}: chained offsets"
    return depth


def rebuild_steps_every(every, nsteps):
    """Steps (1-based) that rebuild with `every` = E, check off and no other trigger."""
    return [s for s in range(1, nsteps + 1) if (s - 1) % (every + 1) == 0]


def cells(pos, grid):
    """The cell of each position under a grid record (bin_count, bin_size,
    bin_origin, ...): floor((p - o) / size) clamped into the grid."""
    n = grid[0:3].astype(np.int64)
    inv = 1.0 / grid[3:6]
    c = np.floor((pos - grid[6:9]) * inv)
    c = np.clip(c, 0, n - 1).astype(np.int64)
    return c[:, 0] * n[1] * n[2] + c[:, 1] * n[2] + c[:, 2]


def sort_is_identity(pos, grid):
    cl = cells(pos, grid)
    return bool(np.array_equal(np.argsort(cl, kind="stable"), np.arange(len(cl))))


def removal_order(removed):
    """Index order left by Atom::swap_remove over the removed locals in
    descending index order (domain.rs:633-661, atom.rs:454-458), by the
    closed form the device uses: the hole at a removed index p below n - m
    (rank k = removals at or above p) takes position c = n - k, followed
    through c = n - rank(c) while c is itself removed."""
    removed = np.asarray(removed, dtype=bool)
    n, m = len(removed), int(removed.sum())
    below = np.concatenate([[0], np.cumsum(removed)])  # removals below each index
    order = np.arange(n - m)
    for p in np.nonzero(removed[: n - m])[0]:
        c = n - (m - below[p])
        while removed[c]:
            c = n - (m - below[c])
        order[p] = c
    return order


def removal_order_sequential(removed):
    """The same by the reference's own loop, on an index array."""
    a = list(range(len(removed)))
    for i in range(len(removed) - 1, -1, -1):
        if removed[i]:
            a[i] = a[-1]
            a.pop()
    return np.array(a, dtype=np.int64)


def stencil_digits(offsets, nbin):
    """(dx, dy, dz) of each cell offset dx*ny*nz + dy*nz + dz. The grid holds
    the stencil's reach of ghost bins on both sides of at least one bin, so
    |dz| <= (nz - 1) // 2 and |dy| <= (ny - 1) // 2, and the digits are unique."""
    ny, nz = int(nbin[1]), int(nbin[2])
    hz, hy = (nz - 1) // 2, (ny - 1) // 2
    off = np.asarray(offsets, dtype=np.int64)
    dz = (off + hz) % nz - hz
    rest = (off - dz) // nz
    dy = (rest + hy) % ny - hy
    dx = (rest - dy) // ny
    assert np.array_equal(dx * ny * nz + dy * nz + dz, off)
    return np.stack([dx, dy, dz], axis=1)


def check_stencil_order(stencil, nbin, newton):
    """The stencil a row walks after its own cell: generation order is
    lexicographic in (dx, dy, dz); the half list keeps the offsets > 0, the
    full list every offset but 0."""
    st = np.asarray(stencil, dtype=np.int64)
    d = stencil_digits(st, nbin)
    for k in range(1, len(st)):
        assert tuple(d[k - 1]) < tuple(d[k]), f"stencil entries {k - 1}, {k} out of generation order"
    assert (st > 0).all() if newton else (st != 0).all()


def check_csr_order(off, idx, bins, stencil, nbin, newton):
    """Every row of the CSR in the reference's loop order (see the module
    docstring). `stencil`: the offsets the rows walk after their own cell."""
    check_stencil_order(stencil, nbin, newton)
    bin_start, sorted_atoms, sorted_idx, atom_cell = (np.asarray(a, dtype=np.int64) for a in bins)
    seg_of = {int(o): k + 1 for k, o in enumerate(stencil)}
    off = np.asarray(off, dtype=np.int64)
    idx = np.asarray(idx, dtype=np.int64)
    nl = len(off) - 1
    rows = np.repeat(np.arange(nl), np.diff(off))
    dcell = atom_cell[idx] - atom_cell[rows]
    seg = np.array([0 if d == 0 else seg_of.get(int(d), -1) for d in dcell], dtype=np.int64)
    assert (seg >= 0).all(), "an entry lies in a cell outside the row's stencil"
    pos = sorted_idx[idx]  # each entry's place in bin order
    assert np.array_equal(sorted_atoms[pos], idx)
    own = seg == 0
    if newton:
        # the half list's own cell: only the records after the row's own
        assert (pos[own] > sorted_idx[rows[own]]).all(), "a self-cell entry at or before its row's record"
    else:
        assert (idx[own] != rows[own]).all(), "a row lists itself"
    # within each row the key (segment, place in bin order) strictly increases
    same_row = rows[1:] == rows[:-1]
    key_up = (seg[1:] > seg[:-1]) | ((seg[1:] == seg[:-1]) & (pos[1:] > pos[:-1]))
    bad = np.nonzero(same_row & ~key_up)[0]
    assert bad.size == 0, f"row {rows[bad[0]] if bad.size else -1}: entries out of the reference's loop order"


# Two atoms 0.96 apart along x (0.99125 - 0.03125, exact), y = z = 10, in a
# periodic 20-cube, with sf = 1.2 and radius 0.4: r2 = fl(0.96 * 0.96) =
# 0.9216 equals the uniform form's c*c = 0.9216 (not below it) and lies below
# the per-pair form's 0.8*0.8*1.44 = 0.9216000000000002.
CUT_SF = 1.2
CUT_R = 0.4
CUT_POS = np.array([[0.03125, 10.0, 10.0], [0.99125, 10.0, 10.0], [10.0, 10.0, 10.0]])


def cutoff_sq_rule(r_i, r_j, radii, sf):
    """The squared cutoff the build applies to the pair (r_i, r_j) of a box
    whose locals have `radii` (neighbor.rs:679-688 save_build_positions, then
    the uniform or per-pair test of bin_neighbor_list)."""
    mn, mx = min(radii), max(radii)
    if abs(mx - mn) < 1e-15:
        c = 2.0 * mn * sf
        return c * c
    s = r_i + r_j
    return s * s * (sf * sf)


def cutoff_cases():
    """(third atom's radius, whether the build lists the pair 0-1): the
    third atom is far from both and only decides which form applies."""
    x = CUT_POS[1, 0] - CUT_POS[0, 0]
    r2 = x * x
    out = []
    for r3 in (CUT_R, float(np.nextafter(CUT_R, 1.0)), CUT_R + 2e-15, CUT_R + 0.05):
        t = cutoff_sq_rule(CUT_R, CUT_R, [CUT_R, CUT_R, r3], CUT_SF)
        out.append((r3, r2 < t))
    return out


def fma_exact(a, b, c):
    """a*b + c rounded once (Rust's mul_add), by exact rational arithmetic."""
    from fractions import Fraction
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def r2_reference(dx, dy, dz):
    return fma_exact(dx, dx, fma_exact(dy, dy, dz * dz))


def r2_alternatives(dx, dy, dz):
    """Plausible misreadings: a plain sum, and the nesting reversed."""
    return dx * dx + dy * dy + dz * dz, fma_exact(dz, dz, fma_exact(dy, dy, dx * dx))


# Two pairs in a periodic 20-cube with radius 0.4 everywhere and sf = 1.2, so
# the cutoff is c*c = 0.9216: atom 0 at (0.25, 0.25, 0.25), atom 1 at one of
# these positions (differences exact). The reference's nesting puts the first
# at 0.9216 (not listed) and the second at 0.9215999999999999 (listed); the
# plain sum and the reversed nesting give the opposite verdict for each.
FMA_R, FMA_SF = 0.4, 1.2
FMA_PI = (0.25, 0.25, 0.25)
FMA_PJ = [(0.9818395051144911, 0.7291151231438562, 0.6455497914657826),
          (0.9118833742240997, 0.6027101868239697, 0.8492544726877895)]


# One atom whose first step separates the Verlet expression from three
# misreadings (a fused kick, a fused drift, h from dividing by the mass):
# each changes at least one of its velocity or position components.
VERLET_DT = 1e-3
VERLET_ATOM = dict(
    m=0.663704541435512,
    v=(0.13442849658938183, -0.5672019555172545, 1.5736560485162756),
    f=(278.13949508158737, -403.1574471379613, 317.914802580167),
    x=(11.923914911422258, 14.98587806422198, 10.067779050664738))


def verlet_reference(x, v, f, m, dt):
    """initial_integration (soil_verlet/src/lib.rs:143-160) for one component."""
    im = 1.0 / m
    h = 0.5 * dt * im
    v0 = v + h * f
    return v0, x + v0 * dt


def verlet_alternatives(x, v, f, m, dt):
    im = 1.0 / m
    h = 0.5 * dt * im
    fused_kick = fma_exact(h, f, v)
    fused_drift = v + h * f
    h_div = 0.5 * dt / m
    v_div = v + h_div * f
    return [(fused_kick, x + fused_kick * dt), (fused_drift, fma_exact(fused_drift, dt, x)), (v_div, x + v_div * dt)]


# Positions where the reference's product by the inverse bin size and a
# division by the bin size give different cells, in a periodic 20-cube of
# radius-0.5 atoms with sf = 1.1 (24 bins of 20/18 per axis from -10/3).
CELL_POS = np.array([[7.777777777777778, 11.11111111111111, 17.77777777777778],
                     [2.2222222222222223, 12.222222222222221, 18.88888888888889],
                     [15.0, 5.0, 9.0]])


def cells_by_division(pos, grid):
    """The misreading: floor((p - o) / size)."""
    n = grid[0:3].astype(np.int64)
    c = np.floor((pos - grid[6:9]) / grid[3:6])
    c = np.clip(c, 0, n - 1).astype(np.int64)
    return c[:, 0] * n[1] * n[2] + c[:, 1] * n[2] + c[:, 2]


# One atom of radius 0.5 (sf = 1.1, dt = 1e-3, no force) per case, moving
# about the displacement threshold in one step: after the build of step 1,
# step 2's decision depends on whether the squared displacement is fused
# (fully periodic boxes) or a plain sum (boxes with a non-periodic axis).
DISP_SF, DISP_R, DISP_DT = 1.1, 0.5, 1e-3
DISP_CASES = [
    ((0.4211645829016162, 0.42858815844774134, 0.4700675058101765),
     (49.81585995812322, -3.8877342541753896, 1.8071023771698262)),
    ((0.5597565455328373, 0.4586275074444136, 0.5460616877157702),
     (-18.44897763477604, -37.71252681317814, -27.15512006228029)),
]


def disp_decisions(x0, v):
    """(rebuild at step 2 in a fully periodic box, in a box with a fixed axis)."""
    t = (DISP_SF - 1.0) * DISP_R
    thr = t * t
    x1 = [x0[k] + v[k] * DISP_DT for k in range(3)]
    x2 = [x1[k] + v[k] * DISP_DT for k in range(3)]
    d = [x2[k] - x1[k] for k in range(3)]
    fused = fma_exact(d[0], d[0], fma_exact(d[1], d[1], d[2] * d[2]))
    plain = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
    return fused > thr, plain > thr


# A periodic box whose x bounds give lo + (hi - lo) > hi and whose y bounds
# give lo + (hi - lo) < hi. Atom 0 sits at x = hi (below lo + size: not
# wrapped), atom 1 at y = lo + size (below hi, but wrapped).
WRAP_LO = (1.423, 4.391, 0.0)
WRAP_HI = (18.282, 23.203, 20.0)


def wrap_case():
    """Positions, and the (position, image) the wrap leaves, per atom."""
    size = [WRAP_HI[k] - WRAP_LO[k] for k in range(3)]
    top = [WRAP_LO[k] + size[k] for k in range(3)]
    assert top[0] > WRAP_HI[0] and top[1] < WRAP_HI[1]
    pos = np.array([[WRAP_HI[0], 10.0, 10.0], [10.0, top[1], 10.0]])
    want_pos = pos.copy()
    want_img = np.zeros((2, 3), dtype=np.int32)
    want_pos[1, 1] = top[1] - size[1]
    want_img[1, 1] = 1
    return pos, want_pos, want_img


# x bounds whose face thresholds round: with radius 0.5 and sf = 1.1 the
# ghost cutoff is 1.2000000000000002; atom 0 at x just below lo + gc (where
# x - lo < gc is false) and atom 1 at x = hi - gc (where hi - x <= gc is
# false) are both sent.
FACE_LO = (-1.5206, 0.0, 0.0)
FACE_HI = (15.0177, 20.0, 20.0)
FACE_X = (-0.32059999999999983, 13.817699999999999)


def face_sends(gc):
    """(sent low, sent high) by the reference's thresholds, and by the
    difference form, for the two atoms."""
    lo, hi = FACE_LO[0], FACE_HI[0]
    ref = (FACE_X[0] < lo + gc, FACE_X[1] >= hi - gc)
    alt = (FACE_X[0] - lo < gc, hi - FACE_X[1] <= gc)
    return ref, alt


def sent_by_swap(a, swaps):
    """Tags of the locals each swap's ghosts come from."""
    nl, ng = a["nlocal"], a["nghost"]
    sw = ghost_swaps(swaps, nl, ng)
    out = {}
    for g in range(nl, nl + ng):
        out.setdefault(int(sw[g - nl]), set()).add(int(a["tag"][int(a["origin"][g])]))
    return out


# A shear rate whose tilt after 9 steps (dt 1e-3, a 20-cube) the reference's
# left-to-right product rounds differently from every regrouping.
TILT_RATE, TILT_STEPS = 0.112785, 9


def tilt_reference(rate, dt, step, lx, ly):
    raw = rate * dt * float(step) * ly
    return raw - lx * round(raw / lx)


def tilt_regroupings(rate, dt, step, ly):
    s = float(step)
    return [rate * (dt * s) * ly, (rate * ly) * (dt * s), rate * ly * dt * s, (rate * dt) * (s * ly)]


# A box tilted by xy = 1.7 (boundary velocity 0.3) with non-zero x and y
# lower bounds. Atom A is inside and its lamda round trip moves every
# coordinate by an ulp; atom B sits on the sheared x face where the
# reference's x2lamda keeps lx < 1 and x - (x0 + xy*ly) would wrap it.
TRI_LO = (-1.5206, 0.3, 0.0)
TRI_HI = (15.0177, 20.3, 20.0)
TRI_XY, TRI_BVEL = 1.7, 0.3
TRI_POS = np.array([[3.7638479564302947, 13.354988781294496, 12.359699890685233],
                    [15.986470841224186, 11.69730401440221, 8.081364575891442]])


def tri_pbc(r, regrouped=False, keep=False):
    """(position, image delta) after pbc_triclinic for one atom of the
    TRI box; `regrouped` and `keep` are the misreadings."""
    size = [TRI_HI[k] - TRI_LO[k] for k in range(3)]
    xy, xz, yz = TRI_XY, 0.0, 0.0
    lz = (r[2] - TRI_LO[2]) / size[2]
    ly = (r[1] - TRI_LO[1] - yz * lz) / size[1]
    if regrouped:
        lx = (r[0] - (TRI_LO[0] + xy * ly + xz * lz)) / size[0]
    else:
        lx = (r[0] - TRI_LO[0] - xy * ly - xz * lz) / size[0]
    lam, delta = [lx, ly, lz], [0, 0, 0]
    for d in range(3):
        if lam[d] < 0.0:
            lam[d] += 1.0
            delta[d] = -1
        elif lam[d] >= 1.0:
            lam[d] -= 1.0
            delta[d] = 1
    if keep and delta == [0, 0, 0]:
        return list(r), delta
    return [TRI_LO[0] + size[0] * lam[0] + xy * lam[1] + xz * lam[2],
            TRI_LO[1] + size[1] * lam[1] + yz * lam[2],
            TRI_LO[2] + size[2] * lam[2]], delta


def reverse_order(nswaps):
    """Swap order of reverse_send_force: rounds of two from the last, low swap first."""
    seq, end = [], nswaps
    while end > 0:
        start = max(end - 2, 0)
        seq += list(range(start, end))
        end = start
    return seq


def reverse_sequential(vals, swaps, origin, order=None):
    """reverse_send_force's sums as the reference orders them (reverse_order,
    or `order`), ghost by ghost in receive order, source += ghost. A round
    packs both swaps' ghost values before adding either; no ghost of a round
    is a source within it (its candidates precede the round's ghosts), so
    adding as it goes is the same. `vals`: one row per atom (copied)."""
    out = np.array(vals, dtype=np.float64, copy=True)
    seq = reverse_order(len(swaps)) if order is None else order
    for s in seq:
        start, count = int(swaps[s][0]), int(swaps[s][1])
        for g in range(start, start + count):
            out[int(origin[g])] += out[g]
    return out


# Two atoms in a periodic 20-cube (radius 0.5, sf 1.1, sort_every 3): A moves
# +x at 10 from x = 19.975 and crosses the face on step 3 by less than the
# displacement threshold; B rests in A's cell column at a larger y, so A's
# move into the next cell reverses their cell order and step 3's periodic
# sort rebuilds.
PSORT_EVERY = 3
PSORT_POS = np.array([[19.975, 1.0, 10.0], [19.5, 15.0, 10.0]])
PSORT_VEL = np.array([[10.0, 0.0, 0.0], [0.0, 0.0, 0.0]])


def psort_x_after(steps, dt=1e-3):
    x = PSORT_POS[0, 0]
    for _ in range(steps):
        x = x + PSORT_VEL[0, 0] * dt
    return x


# An atom at rest in a Lees-Edwards 20-cube (rate TILT_RATE) at a height
# where rate * (y - yc) and rate * y - rate * yc round differently.
PROFILE_Y = 4.984913705011256


def profile_vx(rate, y, ylo=0.0, yhi=20.0):
    return 0.0 + rate * (y - 0.5 * (ylo + yhi))


# One atom in a periodic 20-cube tilted by xy = 3.86 whose lamda x lies
# between the reference's low-face threshold (gc / Lx) * inflate and the
# regrouped gc * inflate / Lx: the reference sends it, the regrouping would not.
TRIF_XY = 3.86
TRIF_POS = np.array([[2.640418637, 6.263309, 10.0]])


def trif_lx(p, xy=TRIF_XY, L=20.0):
    lz = (p[2] - 0.0) / L
    ly = (p[1] - 0.0 - 0.0 * lz) / L
    return (p[0] - 0.0 - xy * ly - 0.0 * lz) / L


def trif_thresholds(gc, xy=TRIF_XY, L=20.0):
    inflate = 1.0 + abs(xy) / L
    return 0.0 + (gc / L) * inflate, gc * inflate / L
