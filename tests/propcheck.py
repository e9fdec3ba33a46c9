"""Size-independent checks of a box's state at the end of a rebuild step,
for boxes too large to step beside the oracle:

* the half list equals the set of image pairs within (ri+rj)·sf that a k-d
  tree finds among the locals and their ghosts in real space (each unordered
  pair of atoms once, ghosts resolved to their local), and the CSR is well
  formed;
* the harmonic consumer's 2^-32 fixed-point forces sum to exactly zero over
  the locals after reverse comm (Newton's third law, no rounding).
"""
import numpy as np
from scipy.spatial import cKDTree


def roots(origin, n):
    """Local atom each atom images (ghosts of ghosts resolved to the local)."""
    root = np.arange(len(origin), dtype=np.int64)
    for g in range(n, len(origin)):  # a ghost's source precedes it
        root[g] = root[int(origin[g])]
    return root


def check_csr_well_formed(off, idx, counts):
    n, ng = counts["nlocal"], counts["nghost"]
    assert off[0] == 0 and off[-1] == counts["npairs"] == len(idx)
    assert np.all(np.diff(off.astype(np.int64)) >= 0)
    if len(idx):
        assert idx.max() < n + ng
    rows = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    assert not np.any(idx == rows), "self pair"


def check_image_pair_set(a, off, idx, sf):
    n = a["nlocal"]
    nall = n + a["nghost"]
    pos, rad = a["pos"][:nall], a["radius"][:nall]
    root = roots(a["origin"][:nall].astype(np.int64), n)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(off.astype(np.int64)))
    tj = root[idx.astype(np.int64)]
    key_gpu = np.minimum(rows, tj) * n + np.maximum(rows, tj)
    assert len(np.unique(key_gpu)) == len(key_gpu), "a pair is listed twice"
    tree = cKDTree(pos)
    cand = tree.query_pairs(2.0 * rad.max() * sf, output_type="ndarray")
    cand = cand[(cand[:, 0] < n) | (cand[:, 1] < n)]  # a row is a local
    d = pos[cand[:, 1]] - pos[cand[:, 0]]
    r2 = np.einsum("ij,ij->i", d, d)
    s = rad[cand[:, 0]] + rad[cand[:, 1]]
    thr = (s * s) * (sf * sf)
    inside = r2 < thr
    ambiguous = np.abs(r2 - thr) <= 1e-9 * thr
    ri, rj = root[cand[:, 0]], root[cand[:, 1]]
    key_ref = np.minimum(ri, rj) * n + np.maximum(ri, rj)
    want = set(key_ref[inside & ~ambiguous].tolist())
    allowed = want | set(key_ref[ambiguous].tolist())
    got = set(key_gpu.tolist())
    assert want <= got, f"{len(want - got)} pairs missing"
    assert got <= allowed, f"{len(got - allowed)} extra pairs"
    return len(got)


def check_force_balance(a):
    f = a["force"][:a["nlocal"]]
    q = f * 4294967296.0
    assert np.array_equal(q, np.round(q)), "forces are not multiples of 2^-32"
    total = q.astype(np.int64).sum(axis=0)
    assert np.all(total == 0), f"local forces sum to {total.tolist()} (2^-32 units)"
