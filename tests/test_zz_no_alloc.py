"""soil_hip_compute_ke_noalloc against soil_hip_compute_ke, through the C ABI.

The allocation-free variant keeps its mask and per-tile partial sums in the
library's fixed per-GPU workspace instead of the box's scratch; it must give
the same bits as the allocating entry point, call after call, also when a
second call covers fewer locals than the first (the workspace keeps the
first call's partials, which the second must not read).
"""
import ctypes as C

import numpy as np
import pytest

import soil_amd
from soil_amd import _ffi
from soil_amd._ffi import lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return soil_amd.Context(0)


def _box(ctx, seed, n, steps):
    desc, pos, vel, mass, rad = soil_amd.gen_custom(seed, n, 0.4, 0.6, 0.45)
    b = soil_amd.Box(ctx, desc)
    b.upload(pos, mass, rad, vel=vel)
    b.set_consumer(1000.0)
    b.setup()
    b.step(steps)
    return b


def _ke(fn, box, mask):
    out = C.c_double(-1.0)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    mp = None if m is None else m.ctypes.data_as(C.POINTER(C.c_uint8))
    rc = fn(box.h, mp, C.byref(out))
    assert rc == 0, _ffi.lib().soil_hip_last_error()
    return out.value


def test_noalloc_ke_equals_allocating_ke_on_two_calls(ctx):
    big = _box(ctx, 9101, 20000, 15)
    small = _box(ctx, 9102, 3000, 10)
    for box in (big, small):  # the second call covers fewer locals than the first
        n = box.counts()["nlocal"]
        print(f"case: compute_ke_noalloc nlocal={n}", flush=True)
        masks = (None, (np.arange(n) % 3 == 0), np.zeros(n, dtype=bool), np.ones(n, dtype=bool))
        for mask in masks:
            k_alloc = _ke(lib().soil_hip_compute_ke, box, mask)
            k_free = _ke(lib().soil_hip_compute_ke_noalloc, box, mask)
            assert k_free == k_alloc, (n, k_free, k_alloc)
        assert _ke(lib().soil_hip_compute_ke_noalloc, box, np.zeros(n, dtype=bool)) == 0.0
        assert _ke(lib().soil_hip_compute_ke_noalloc, box, None) > 0.0


def test_noalloc_ke_after_a_larger_masked_call(ctx):
    # a masked call on a large box leaves mask bytes and partials beyond a
    # smaller box's range; the smaller box's unmasked and masked sums ignore them
    big = _box(ctx, 9103, 30000, 5)
    small = _box(ctx, 9104, 2000, 5)
    nb, ns = big.counts()["nlocal"], small.counts()["nlocal"]
    _ke(lib().soil_hip_compute_ke_noalloc, big, np.ones(nb, dtype=bool))
    mask = np.arange(ns) % 2 == 1
    assert _ke(lib().soil_hip_compute_ke_noalloc, small, mask) == _ke(lib().soil_hip_compute_ke, small, mask)
    assert _ke(lib().soil_hip_compute_ke_noalloc, small, None) == _ke(lib().soil_hip_compute_ke, small, None)
