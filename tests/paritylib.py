"""Helpers shared by the GPU parity tests: build matching GPU/oracle boxes and compare."""
import numpy as np

import soil_amd
from oracle_ref import RefBox


def make_pair(ctx, desc, pos, vel, mass, rad, k=None, columns=(), col_data=(), pre_setup=None):
    g = soil_amd.Box(ctx, desc)
    r = RefBox(desc)
    for (w, f) in columns:
        g.register_column(w, f)
        r.register_column(w, f)
    g.upload(pos, mass, rad, vel=vel)
    r.add_atoms(pos, mass, rad, vel=vel)
    for c, data in enumerate(col_data):
        g.upload_column(c, data)
        r.set_column(c, data, columns[c][0])
    if k is not None:
        g.set_consumer(k)
        r.set_consumer(k)
    if pre_setup is not None:
        pre_setup(g)
        pre_setup(r)
    g.setup()
    r.setup()
    return g, r


def assert_same(g, r, bins=True, exact_state=True, rtol=1e-12, what=""):
    cg, cr = g.counts(), r.counts()
    for key in ("nlocal", "nghost", "npairs", "rebuilds", "steps"):
        assert cg[key] == cr[key], f"{what}: {key} gpu={cg[key]} ref={cr[key]}"
    ag, ar = g.atoms(), r.atoms()
    for key in ("tag", "origin", "image", "radius", "mass", "inv_mass"):
        assert np.array_equal(ag[key], ar[key]), f"{what}: {key} differs"
    for key in ("pos", "vel", "force"):
        if exact_state:
            bad = np.nonzero(ag[key] != ar[key])
            assert bad[0].size == 0, f"{what}: {key} differs at {bad[0][:5]} {ag[key][bad][:5]} {ar[key][bad][:5]}"
        else:
            scale = np.maximum(np.abs(ar[key]), 1.0)
            err = np.max(np.abs(ag[key] - ar[key]) / scale) if ar[key].size else 0.0
            assert err <= rtol, f"{what}: {key} rel err {err}"
    og, ig = g.csr()
    orr, ir = r.csr()
    assert np.array_equal(og, orr), f"{what}: CSR offsets differ"
    assert np.array_equal(ig, ir), f"{what}: CSR indices differ"
    if bins:
        bg, br = g.bins(), r.bins()
        for name, x, y in zip(("bin_start", "sorted", "sorted_idx", "cell"), bg, br):
            assert np.array_equal(x, y), f"{what}: bins {name} differ"
    sg, sr = g.swaps(), r.swaps()
    assert np.array_equal(sg[:, :3], sr[:, :3]), f"{what}: swap table differs"
    assert np.array_equal(sg[:, 5:], sr[:, 5:]), f"{what}: swap offsets differ"
