"""Force hooks (the GpuForce role, soil_gpu/src/gpu_state.rs:45-55) on the GPU.

The test hooks (tests/hooks/fixed_point_hook.hip) restate the harmonic test
consumer as user code over the device view: a half-list scatter with int64
atomics, and a gather over the incoming-pair transpose. Both sum the same
2^-32 fixed-point pair forces exactly, so a hooked box must step bit for bit
like the oracle running the consumer, in the plain step and recorded into a
batch's step graph.
"""
import ctypes as C
import os

import numpy as np
import pytest

import soil_amd
from paritylib import assert_same, make_pair
from oracle_ref import RefBox

pytestmark = pytest.mark.gpu

HOOK_SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hooks", "libsoil_test_hook.so")


@pytest.fixture(scope="module")
def ctx():
    return soil_amd.Context(0)


@pytest.fixture(scope="module")
def hooks():
    if not os.path.exists(HOOK_SO):
        raise RuntimeError(f"{HOOK_SO} is missing: run tests/hooks/build_hooks.py")
    lib = C.CDLL(HOOK_SO)
    lib.soil_test_hook_create.argtypes = [C.c_size_t, C.c_double]
    lib.soil_test_hook_create.restype = C.c_void_p
    lib.soil_test_hook_destroy.argtypes = [C.c_void_p]
    lib.soil_test_hook_calls.argtypes = [C.c_void_p]
    lib.soil_test_hook_calls.restype = C.c_uint64
    return lib


def fn_addr(lib, name):
    return C.cast(getattr(lib, name), C.c_void_p).value


def hooked_box(ctx, hooks, gen, kind, k=1000.0):
    d, pos, vel, mass, rad = gen
    b = soil_amd.Box(ctx, d)
    b.upload(pos, mass, rad, vel=vel)
    state = hooks.soil_test_hook_create(8 * len(mass) + 65536, k)
    assert state
    b.set_force_hook(fn_addr(hooks, f"soil_test_hook_{kind}"), state, transpose=(kind == "gather"))
    b.setup()
    return b, state


def oracle_box(gen, k=1000.0):
    d, pos, vel, mass, rad = gen
    r = RefBox(d)
    r.add_atoms(pos, mass, rad, vel=vel)
    r.set_consumer(k)
    r.setup()
    return r


@pytest.mark.parametrize("kind", ["scatter", "gather"])
def test_hook_steps_like_the_consumer(ctx, hooks, kind):
    print(f"case: hook {kind}", flush=True)
    gen = soil_amd.gen_custom(6100, 5000, 0.4, 0.6, 0.5, (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.0)
    b, state = hooked_box(ctx, hooks, gen, kind)
    r = oracle_box(gen)
    for s in range(3):
        b.step(20)
        r.step(20)
        assert_same(b, r, what=f"{kind} step {20 * (s + 1)}")
    assert hooks.soil_test_hook_calls(state) == 60
    v = b.device_view()
    c = b.counts()
    assert (v.nlocal, v.nghost, v.npairs) == (c["nlocal"], c["nghost"], c["npairs"])
    assert v.x and v.fx and v.offsets and v.indices and v.xr
    assert bool(v.tstart) == (kind == "gather")
    b.close()
    hooks.soil_test_hook_destroy(state)


def test_hook_recorded_into_a_batch_graph(ctx, hooks):
    # hooks read counts and buffers from their device views at run time, so the
    # captured graph stays right across sorts (column buffers swap) and rebuilds
    print("case: hooks in a graph batch", flush=True)
    rng = np.random.default_rng(61)
    gens = [soil_amd.gen_custom(6200 + k, int(rng.integers(500, 4000)), 0.4, 0.6, float(rng.uniform(0.25, 0.5)),
                                (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.1 if k % 2 else 0.0) for k in range(5)]
    boxes, states = [], []
    for k, gen in enumerate(gens):
        b, st = hooked_box(ctx, hooks, gen, "scatter" if k % 2 == 0 else "gather")
        boxes.append(b)
        states.append(st)
    oracles = [oracle_box(g) for g in gens]
    batch = soil_amd.Batch(ctx, boxes)
    for _ in range(3):
        batch.step(10, graph=True)
    batch.sync()
    batch.close()
    for k, (b, r) in enumerate(zip(boxes, oracles)):
        r.step(30)
        assert_same(b, r, what=f"batch box {k}")
    for b, st in zip(boxes, states):
        b.close()
        hooks.soil_test_hook_destroy(st)


def test_hook_adds_to_the_consumer(ctx, hooks):
    # a hook adds to what the consumer assigned: consumer k plus a hook of the
    # same k give exactly twice the consumer's forces (both are the same exact
    # integer sum times 2^-32, and doubling commutes with every f64 addition)
    print("case: hook + consumer", flush=True)
    d, pos, vel, mass, rad = soil_amd.gen_custom(6300, 3000, 0.4, 0.6, 0.5, (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.0)
    n = len(mass)
    b = soil_amd.Box(ctx, d)
    b.upload(pos, mass, rad, vel=vel)
    b.set_consumer(500.0)
    state = hooks.soil_test_hook_create(8 * n + 65536, 500.0)
    b.set_force_hook(fn_addr(hooks, "soil_test_hook_scatter"), state)
    b.setup()
    c = soil_amd.Box(ctx, d)
    c.upload(pos, mass, rad, vel=vel)
    c.set_consumer(500.0)
    c.setup()
    b.step(1)
    c.step(1)
    fb, fc = b.atoms()["force"][:n], c.atoms()["force"][:n]
    assert np.any(fc != 0.0)
    assert np.array_equal(fb, 2.0 * fc)
    b.set_force_hook(None)
    b.close()
    c.close()
    hooks.soil_test_hook_destroy(state)


def test_hook_as_a_per_phase_system(ctx, hooks):
    # soil_hip_run_force_hook in the Force phase of the per-system sequence
    print("case: hook per phase", flush=True)
    gen = soil_amd.gen_custom(6400, 2500, 0.4, 0.6, 0.45, (1, 1, 1), 0, 0.0, 0.0, 0.1, 0.0)
    g, state = hooked_box(ctx, hooks, gen, "scatter")
    r = oracle_box(gen)
    for step in range(3):
        g.apply_deform(); r.phase("apply_deform")
        g.initial_integration(); r.phase("initial_integration")
        if step > 0:
            g.decide_rebuild(); r.phase("decide_rebuild")
        if r.state == 0:
            g.remove_ghost_atoms(); r.phase("remove_ghost_atoms")
        g.zero_all_forces(); r.phase("zero_all_forces")
        if r.state == 0:
            g.pbc(); r.phase("pbc")
        g.sort_atoms_by_bin(); r.phase("sort_atoms_by_bin")
        if r.state == 0:
            g.borders(); r.phase("borders")
            g.bin_neighbor_list(); r.phase("bin_neighbor_list")
        else:
            g.forward_comm_borders(); r.phase("forward_comm_borders")
        g.run_force_hook(); r.phase("harmonic_force")
        g.reverse_send_force(); r.phase("reverse_send_force")
        g.final_integration(); r.phase("final_integration")
        ag, ar = g.atoms(), r.atoms()
        for key in ("pos", "vel", "force"):
            assert np.array_equal(ag[key], ar[key]), f"step {step}: {key} differs"
    assert hooks.soil_test_hook_calls(state) == 3
    g.close()
    hooks.soil_test_hook_destroy(state)
