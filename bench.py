"""Benchmark: particle-steps/s including rebuilds, batched vs plain, on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--parts plain,batched,served,cpu]

Prints ONE JSON line. `value` is the plain single-box path on the 1,000,000
sphere polydisperse periodic box (phi 0.50, skin fraction 1.1, harmonic
consumer k=1000, dt 1e-3, trigger-driven rebuilds), inputs resident in HBM
before the timed region. Extra objects carry the batched ensemble (256 boxes,
one launch per phase in a captured graph) against the plain path stepping the
same ensemble one box per launch sequence, the served stream's latencies, the
CSR-build roofline and the single-core CPU oracle baseline.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HIP events around every 8th rebuild's CSR build: each event costs a few
# microseconds of the step, so sampling keeps them out of most timed steps
TIMING_PERIOD = 8


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_plain_1M.json")


# the count-pass kernel of each CSR strategy (SOIL_HIP_CSR_MODE; 3 is the default)
COUNT_KERNEL = {0: "k_csr_count", 1: "k_csr_count_g", 2: "k_csr_count_g", 3: "k_csr_count_w"}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (copy-probe calibrated), or None."""
    try:
        with open(TRAFFIC_FILE) as fh:
            data = json.load(fh)
    except (OSError, ValueError):
        return None, None
    for name, v in data.get("kernels", {}).items():
        if kernel in name:
            return v["read_bytes"] + v["write_bytes"], os.path.relpath(TRAFFIC_FILE, ROOT)
    return None, None


def run_plain(ctx, steps, warmup):
    import soil_amd
    d, pos, vel, mass, rad = soil_amd.gen_box(2)
    n = len(pos)
    box = soil_amd.Box(ctx, d)
    box.upload(pos, mass, rad, vel=vel)
    box.set_consumer(1000.0)
    box.setup()
    if warmup:
        box.step(warmup)
    box.sync()
    c0 = box.counts()
    box.set_timing(TIMING_PERIOD)
    t0 = time.perf_counter()
    box.step(steps)
    box.sync()
    t1 = time.perf_counter()
    tm = box.timing()
    c1 = box.counts()
    el = t1 - t0
    res = {
        "n": n,
        "seconds": el,
        "particle_steps_per_s": n * steps / el,
        "ms_per_step": 1e3 * el / steps,
        "rebuilds": c1["rebuilds"] - c0["rebuilds"],
        "nghost": c1["nghost"],
        "npairs": c1["npairs"],
        "csr": tm,
    }
    box.close()
    return res, d


def ensemble(count=256, config=4):
    import soil_amd
    return [soil_amd.gen_box(config, k) for k in range(count)]


def make_boxes(ctx, gens):
    import soil_amd
    boxes = []
    for d, pos, vel, mass, rad in gens:
        b = soil_amd.Box(ctx, d)
        b.upload(pos, mass, rad, vel=vel)
        b.set_consumer(1000.0)
        b.setup()
        boxes.append(b)
    return boxes


def run_batched(ctx, steps_batched, steps_plain, warmup):
    import soil_amd
    gens = ensemble()
    ntot = sum(len(g[1]) for g in gens)
    # plain: one box per launch sequence, boxes in turn
    boxes = make_boxes(ctx, gens)
    for b in boxes:
        b.step(1)
        b.sync()
    t0 = time.perf_counter()
    for _ in range(steps_plain):
        for b in boxes:
            b.step(1)
    for b in boxes:
        b.sync()
    t_plain = time.perf_counter() - t0
    for b in boxes:
        b.close()
    del boxes
    # batched: one launch per phase for all boxes, graph-captured step
    boxes = make_boxes(ctx, gens)
    batch = soil_amd.Batch(ctx, boxes)
    batch.step(1, graph=True)
    batch.sync()
    batch.autotune()
    batch.step(max(warmup, 1), graph=True)
    batch.sync()
    t0 = time.perf_counter()
    batch.step(steps_batched, graph=True)
    batch.sync()
    t_batch = time.perf_counter() - t0
    batch.close()
    for b in boxes:
        b.close()
    plain = ntot * steps_plain / t_plain
    batched = ntot * steps_batched / t_batch
    return {
        "workload": "256 periodic Lees-Edwards boxes, N log-uniform in [1k,64k], phi U[0.20,0.55], "
                    "r U[0.4,0.6], strain rate U[0,0.2], harmonic consumer, every step rebuilds",
        "particles": ntot,
        "batched_particle_steps_per_s": batched,
        "plain_particle_steps_per_s": plain,
        "speedup": batched / plain,
        "batched_steps": steps_batched,
        "plain_steps": steps_plain,
        "batched_ms_per_step": 1e3 * t_batch / steps_batched,
        "plain_ms_per_step": 1e3 * t_plain / steps_plain,
    }


def run_cpu_baseline(target_s=12.0):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import soil_amd
    from oracle_ref import RefBox
    d, pos, vel, mass, rad = soil_amd.gen_box(2)
    n = len(pos)
    r = RefBox(d)
    r.add_atoms(pos, mass, rad, vel=vel)
    r.set_consumer(1000.0)
    r.setup()
    total, steps = 0.0, 0
    while total < target_s and steps < 200:
        total += r.time_steps(1)
        steps += 1
    return {
        "value": n * steps / total,
        "unit": "particle-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"same 1,000,000-sphere box, first {steps} steps from the initial state "
                  f"({total:.1f} s, single thread, {cpu_model()})",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--parts", default="plain,batched,served,cpu")
    ap.add_argument("--batched-steps", type=int, default=200)
    ap.add_argument("--plain-ensemble-steps", type=int, default=10)
    ap.add_argument("--probe-copy-mb", type=int, default=0,
                    help="also run the calibration copy kernel over this many MB (traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (each box or job shards "
              "to one device); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    parts = set(args.parts.split(","))
    import soil_amd
    if soil_amd.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    ctx = soil_amd.Context(0)
    out = {
        "metric": "particle-steps/s incl. rebuilds, batched vs plain; p50/p99 request latency",
        "value": None,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 inputs)",
        "config": {
            "workload": "single box, plain form: 1,000,000 polydisperse spheres r~U[0.4,0.6], periodic cube "
                        "phi 0.50, skin fraction 1.1, half CSR list, harmonic consumer k=1000, dt 1e-3, "
                        "trigger-driven rebuilds",
            "particles": 1000000,
            "parallelism": "1 GPU",
        },
    }
    if args.probe_copy_mb:
        import ctypes as C
        ms = C.c_double(0.0)
        soil_amd._ffi.check(soil_amd._ffi.lib().soil_hip_probe_copy(ctx.h, args.probe_copy_mb * 1000000, 20,
                                                                   C.byref(ms)))
        out["probe_copy"] = {"bytes": args.probe_copy_mb * 1000000, "ms": ms.value,
                             "GBps": 2 * args.probe_copy_mb * 1e6 / (ms.value * 1e-3) / 1e9}
    if "plain" in parts:
        log("plain: 1M box")
        res, d = run_plain(ctx, args.steps, args.warmup)
        out["value"] = res["particle_steps_per_s"]
        out["ms_per_step"] = res["ms_per_step"]
        out["config"]["rebuilds_in_timed_steps"] = res["rebuilds"]
        out["config"]["nghost"] = res["nghost"]
        out["config"]["npairs"] = res["npairs"]
        tm = res["csr"]
        if tm["builds"] > 0 and tm["count_ms"] > 0:
            per_launch_bytes = tm["csr_bytes"] / tm["builds"]
            avg_count_ms = tm["count_ms"] / tm["builds"]
            ach = per_launch_bytes / (avg_count_ms * 1e-3) / 1e9
            kname = COUNT_KERNEL[int(os.environ.get("SOIL_HIP_CSR_MODE", "3") or 3)]
            traffic, traffic_src = measured_traffic(kname)
            out["roofline"] = {
                "kernel": f"{kname} (CSR count pass: candidate tests, row counts and row entries)",
                "bound": "hbm",
                "achieved": ach,
                "peak": HBM_PEAK_GBS,
                "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS,
                "traffic": traffic,
                "traffic_source": traffic_src,
                "algorithmic_bytes_per_launch": per_launch_bytes,
                "launches_timed": tm["builds"],
                "timing_period": TIMING_PERIOD,
                "avg_launch_ms": avg_count_ms,
                "build_avg_ms": tm["csr_ms"] / tm["builds"],
                "build_note": "whole CSR build = count + offset scan + fill; the fill also takes the consumer's "
                              "transpose slots",
                "fill_avg_ms": tm["fill_ms"] / tm["builds"],
            }
        log(f"plain: {out['value']:.3e} particle-steps/s, {res['ms_per_step']:.3f} ms/step")
    if "batched" in parts:
        log("batched: 256-box ensemble")
        out["batched"] = run_batched(ctx, args.batched_steps, args.plain_ensemble_steps, args.warmup)
        log(f"batched: {out['batched']['speedup']:.2f}x over plain")
    if "served" in parts:
        from soil_amd import serve
        out["served"] = serve.bench_served(ctx)
    if "cpu" in parts:
        log("cpu baseline")
        out["cpu_baseline"] = run_cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
