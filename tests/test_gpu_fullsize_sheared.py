"""Full-size checks on the 4,000,000-sphere sheared, clustered box.

The box is 3:1:1, Lees-Edwards sheared (strain rate 0.1, so the xy tilt grows
every step and every step rebuilds), with half the spheres in 64 Gaussian
blobs: the triclinic path at its full size. As for the 1M box, the oracle is
too slow to step beside it, so the tests check properties that fix the
answer independently of size, at the end of a step (each step rebuilt):
  * the CSR is well formed;
  * the half list equals the set of pairs within (ri+rj)·sf found by a k-d
    tree over the locals and their ghost images (the ghosts carry the tilted
    periodic images, so no periodic metric is needed), each unordered pair
    exactly once (the box is over 100 cutoffs wide, so a pair has one image
    within the cutoff);
  * every ghost equals its source plus its swap's shift (the y swaps carry
    the tilt), the swaps following one another in scan order;
  * the consumer's forces, 2^-32 fixed point, sum to exactly zero.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import soil_amd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3():
    print("case: 4M-sphere sheared clustered box, 6 steps", flush=True)
    ctx = soil_amd.Context(0)
    d, pos, vel, mass, rad = soil_amd.gen_box(3)
    box = soil_amd.Box(ctx, d)
    box.upload(pos, mass, rad, vel=vel)
    box.set_consumer(1000.0)
    box.setup()
    before = box.counts()["rebuilds"]
    box.step(6)
    c = box.counts()
    assert c["rebuilds"] == before + 6, "a sheared box rebuilds every step"
    a = box.atoms()
    off, idx = box.csr()
    out = {"atoms": a, "off": off, "idx": idx, "counts": c, "swaps": box.swaps(), "sf": d.skin_fraction,
           "tilt": box.domain()}
    yield out
    box.close()


def roots(origin, n):
    """Local atom each atom images (ghosts of ghosts resolved to the local)."""
    root = np.arange(len(origin), dtype=np.int64)
    root[n:] = origin[n:]
    for _ in range(4):
        g = root >= n
        if not g.any():
            break
        root[g] = root[root[g]]
    assert np.all(root < n)
    return root


def test_sheared_csr_is_well_formed(c3):
    off, idx, c = c3["off"], c3["idx"], c3["counts"]
    n, ng = c["nlocal"], c["nghost"]
    assert n == 4_000_000
    assert off[0] == 0 and off[-1] == c["npairs"] == len(idx)
    assert np.all(np.diff(off.astype(np.int64)) >= 0)
    assert idx.max() < n + ng
    rows = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    assert not np.any(idx == rows), "self pair"
    print(f"npairs {len(idx)} nghost {ng}", flush=True)


def test_sheared_half_list_equals_image_pair_set(c3):
    a, off, idx, sf = c3["atoms"], c3["off"], c3["idx"], c3["sf"]
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
    print(f"pairs gpu {len(got)} kd-tree {len(want)} ambiguous {int(ambiguous.sum())}", flush=True)
    assert want <= got, f"{len(want - got)} pairs missing"
    assert got <= allowed, f"{len(got - allowed)} extra pairs"


def test_sheared_ghosts_are_shifted_sources(c3):
    a, sw = c3["atoms"], c3["swaps"]
    n = a["nlocal"]
    pos, origin = a["pos"], a["origin"]
    lo, hi, size = c3["tilt"]
    scanned = n
    for s in range(len(sw)):
        start, count = int(sw[s, 0]), int(sw[s, 1])
        assert start == scanned, f"swap {s} does not follow the previous one"
        src = origin[start:start + count].astype(np.int64)
        assert np.all(src < start), f"swap {s}: a source after its ghost"
        shift = sw[s, 5:8]
        assert np.array_equal(pos[start:start + count], pos[src] + shift), f"swap {s}: positions"
        scanned = start + count
    assert scanned == n + a["nghost"]
    # the y swaps carry the tilt in x, the other swaps only their own axis
    assert sw[2, 5] != 0.0 and sw[2, 6] == size[1] and sw[3, 6] == -size[1]
    assert sw[0, 5] == size[0] and sw[0, 6] == 0.0 and sw[4, 7] == size[2]


def test_sheared_consumer_forces_sum_to_zero_exactly(c3):
    a = c3["atoms"]
    f = a["force"][:a["nlocal"]]
    q = f * 4294967296.0
    assert np.array_equal(q, np.round(q)), "forces are not multiples of 2^-32"
    total = q.astype(np.int64).sum(axis=0)
    print("sum of local forces (2^-32 units)", total.tolist(), flush=True)
    assert np.all(total == 0)
    assert np.abs(f).max() > 0
