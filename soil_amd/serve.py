"""Served stream: continuous batching of simulation jobs on one device.

`Server` wraps the C ABI's soil_hip_server_*: jobs (boxes) are submitted at
any time, admitted against an HBM budget and a slot limit, kept resident and
stepped together in the batched form; `bench_served` drives the served-stream
workload (Poisson arrivals of 2k-200k sphere boxes, half Lees-Edwards, 500-5000
steps) and reports throughput with p50/p99 step and completion latency.
"""
import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import check, lib

STAT_KEYS = ["jobs_done", "jobs_failed", "particle_steps", "busy_s", "wall_s", "step_p50_ms", "step_p99_ms",
             "completion_p50_s", "completion_p99_s", "wait_p50_s", "wait_p99_s", "mean_active_jobs", "reconfigs",
             "reconfig_s", "budget_gb", "chunks", "step_mean_ms", "completion_max_s", "wait_max_s", "server_steps",
             "jobs_evicted", "gpu_step_p50_ms", "gpu_step_p99_ms", "update_gap_p99_ms"]
NSTATS = len(STAT_KEYS)  # SOIL_HIP_SERVER_NSTATS


def _dp(a):
    return a.ctypes.data_as(_ffi.DP) if a is not None else None


class Server:
    def __init__(self, ctx, budget_gb=200.0, max_slots=256, chunk_steps=16, graph=True):
        self.ctx = ctx
        self.h = C.c_void_p()
        check(lib().soil_hip_server_create(ctx.h, budget_gb, max_slots, chunk_steps, 1 if graph else 0,
                                           C.byref(self.h)))

    def start(self):
        check(lib().soil_hip_server_start(self.h))

    def submit(self, desc, pos, vel, mass, radius, nsteps, k=1000.0, keep_result=False):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        radius = np.ascontiguousarray(radius, dtype=np.float64)
        jid = C.c_int(-1)
        check(lib().soil_hip_server_submit(self.h, C.byref(desc), len(mass), _dp(pos), _dp(vel), _dp(mass),
                                           _dp(radius), int(nsteps), float(k), 1 if keep_result else 0,
                                           C.byref(jid)))
        return jid.value

    def submit_restart(self, desc, path, fmt, nsteps, k=1000.0, keep_result=False):
        """Queues a job whose atoms come from a RestartData file; it runs `nsteps` more steps."""
        jid = C.c_int(-1)
        check(lib().soil_hip_server_submit_restart(self.h, C.byref(desc), str(path).encode(), fmt, int(nsteps),
                                                   float(k), 1 if keep_result else 0, C.byref(jid)))
        return jid.value

    def set_output(self, directory, fmt):
        """Finished jobs write restart_{step}_job{id}.bin/.json into `directory` (None: off)."""
        check(lib().soil_hip_server_set_output(self.h, None if directory is None else str(directory).encode(), fmt))

    def evict(self, job_id, path, fmt):
        """Preempts a job; its state is written to `path` at the next chunk boundary."""
        check(lib().soil_hip_server_evict(self.h, job_id, str(path).encode(), fmt))

    def drain(self):
        check(lib().soil_hip_server_drain(self.h))

    def stats(self):
        out = np.zeros(NSTATS)
        check(lib().soil_hip_server_stats(self.h, _dp(out)))
        return dict(zip(STAT_KEYS, out.tolist()))

    def result(self, job_id, n):
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        got = C.c_size_t(0)
        check(lib().soil_hip_server_result(self.h, job_id, _dp(pos), _dp(vel), n, C.byref(got)))
        return pos, vel

    def close(self):
        if self.h:
            lib().soil_hip_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bench_stream(ctx, njobs, rate, seed=0x50494C, budget_gb=200.0, max_slots=256, chunk_steps=16, graph=True):
    out = np.zeros(NSTATS)
    check(lib().soil_hip_server_bench_stream(ctx.h, njobs, rate, seed, budget_gb, max_slots, chunk_steps,
                                             1 if graph else 0, _dp(out)))
    return dict(zip(STAT_KEYS, out.tolist()))


def bench_served(ctx, njobs=120, rate=6.0, chunk_steps=16):
    st = bench_stream(ctx, njobs, rate, chunk_steps=chunk_steps)
    st["throughput_particle_steps_per_s"] = st["particle_steps"] / st["wall_s"] if st["wall_s"] else 0.0
    st["workload"] = (f"{njobs} jobs, Poisson arrivals at {rate} jobs/s, N log-uniform in [2k,200k], "
                      "phi U[0.2,0.55], half periodic / half Lees-Edwards (strain rate U[0.01,0.2]), "
                      "steps U{500..5000}, harmonic consumer, 200 GB HBM budget, 256 slots")
    return st
