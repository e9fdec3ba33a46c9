"""Brute-force checks (numpy): minimum-image neighbor sets and ghost images.

Independent of any cell list: every pair of locals and all 27 lattice images.
"""
import itertools

import numpy as np


def lattice(box_len, xy=0.0):
    a = np.array([box_len[0], 0.0, 0.0])
    b = np.array([xy, box_len[1], 0.0])
    c = np.array([0.0, 0.0, box_len[2]])
    return a, b, c


def brute_pairs(pos, rad, tags, box_len, sf, xy=0.0, amb=1e-9):
    """Set of (tag_lo, tag_hi, nx, ny, nz) physical pairs within (ri+rj)*sf, plus ambiguous keys."""
    a, b, c = lattice(box_len, xy)
    n = len(pos)
    out, ambiguous = set(), set()
    for nv in itertools.product((-1, 0, 1), repeat=3):
        shift = nv[0] * a + nv[1] * b + nv[2] * c
        d = (pos[None, :, :] + shift[None, None, :]) - pos[:, None, :]
        r2 = np.einsum("ijk,ijk->ij", d, d)
        s = rad[:, None] + rad[None, :]
        cut = s * s * (sf * sf)
        hit = r2 < cut
        amb_m = np.abs(r2 - cut) < amb
        ii, jj = np.nonzero(hit | amb_m)
        for i, j in zip(ii, jj):
            if i == j and nv == (0, 0, 0):
                continue
            ti, tj = int(tags[i]), int(tags[j])
            key = (ti, tj, *nv) if ti < tj else (tj, ti, *(-np.array(nv)))
            if ti == tj:
                key = (ti, tj, *max(nv, tuple(-np.array(nv))))
            key = tuple(int(v) for v in key)
            if amb_m[i, j]:
                ambiguous.add(key)
            elif hit[i, j]:
                out.add(key)
    return out, ambiguous


def image_of(ghost_pos, root_pos, box_len, xy=0.0):
    """Integer lattice vector n with ghost = root + n.(a,b,c) (rounded)."""
    d = ghost_pos - root_pos
    nz = d[2] / box_len[2]
    ny = (d[1]) / box_len[1]
    nx = (d[0] - xy * np.round(ny)) / box_len[0]
    return np.round(np.array([nx, ny, nz])).astype(int)


def csr_pairs(off, idx, atoms, box_len, xy=0.0):
    """Physical pair keys of a CSR half list (ghost columns mapped to their local root)."""
    nl = atoms["nlocal"]
    origin = atoms["origin"]
    pos = atoms["pos"]
    tags = atoms["tag"]

    def root(j):
        while j >= nl:
            j = int(origin[j])
        return j

    keys = []
    for i in range(nl):
        for j in idx[off[i]:off[i + 1]]:
            j = int(j)
            r = root(j)
            nv = image_of(pos[j], pos[r], box_len, xy) if j >= nl else np.zeros(3, dtype=int)
            ti, tj = int(tags[i]), int(tags[r])
            if ti < tj:
                key = (ti, tj, *nv)
            elif ti > tj:
                key = (tj, ti, *(-nv))
            else:
                key = (ti, tj, *max(tuple(nv), tuple(-nv)))
            keys.append(tuple(int(v) for v in key))
    return keys


def brute_ghosts(pos, box_lo, box_hi, gc):
    """Orthogonal box: set of (local index, nx, ny, nz) ghost images the swap rules create."""
    out = set()
    for i, p in enumerate(pos):
        opts = []
        for d in range(3):
            o = [0]
            if p[d] < box_lo[d] + gc:
                o.append(1)
            if p[d] >= box_hi[d] - gc:
                o.append(-1)
            opts.append(o)
        for nv in itertools.product(*opts):
            if nv != (0, 0, 0):
                out.add((i, *nv))
    return out
