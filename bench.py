"""Benchmark: particle-steps/s including rebuilds, batched vs plain, on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--parts plain,sheared,batched,served,cpu]

Prints ONE JSON line. `value` is the plain single-box path on the 1,000,000
sphere polydisperse periodic box (phi 0.50, skin fraction 1.1, harmonic
consumer k=1000, dt 1e-3, trigger-driven rebuilds), inputs resident in HBM
before the timed region. `roofline` is the whole CSR build (count, offset
scan, fill) on that box, timed with HIP events on its stream; `step_roofline`
prices every timed step with the algorithmic byte model (soil_amd/roofline.py).
Extra objects: the 4,000,000-sphere sheared, clustered box (every step
rebuilds); the 256-box ensemble batched (one launch per phase in a captured
graph) against the plain path stepping the same boxes (one step per call in
turn, and n steps per call); the served stream at three arrival rates; and
the CPU oracle (the restatement of the reference's algorithm) timed on this
host: one core on the single boxes, K cores on independent boxes.
"""
import argparse
import json
import os
import platform
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from soil_amd import roofline as rl  # noqa: E402

# The plain leg's throughput is timed with no HIP events in its steps (each
# event costs a few microseconds of the step); the CSR build is timed in a
# second run of the same box of at least ROOFLINE_STEPS steps with events
# around every build.
ROOFLINE_STEPS = 100
LATENCY_BOUND_MS = 5.0   # stated p99 step-latency bound of the served stream
SERVED_RATES = (6.0, 24.0, 96.0)


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


def host_workers():
    # worker threads for the K-core CPU lines: the host's cores, at most 16
    return max(1, min(16, os.cpu_count() or 1))


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_plain_1M.json")

# the count-pass kernel of each CSR strategy (SOIL_HIP_CSR_MODE; 5, one pass, is the default)
COUNT_KERNEL = {0: "k_csr_count", 1: "k_csr_count_g", 2: "k_csr_count_g", 3: "k_csr_count_w",
                4: "k_csr_count_flat", 5: "k_csr_tcount"}
FILL_KERNEL = {0: "k_csr_fill", 1: "k_csr_fill", 2: "k_csr_fill", 3: "k_csr_fill", 4: "k_csr_fill_flat",
               5: "k_csr_tplace"}


def csr_mode():
    return int(os.environ.get("SOIL_HIP_CSR_MODE", "5") or 5)


def csr_kernel_text(mode):
    if mode == 5:  # no scan launch: the placement sums the tile totals itself (scan_ms is the events' gap)
        return ("CSR build: k_csr_tcount (candidate tests, row counts and rows kept in LDS, staged densely per "
                "tile; tile totals and group sums) + k_csr_tplace (tile bases from the group sums, placement, with "
                "the consumer's transpose slots)")
    staged = {4: "dense wave regions"}.get(mode, "row slots")
    scan = {4: "scan of wave totals"}.get(mode, "offset scan")
    return (f"CSR build: {COUNT_KERNEL[mode]} (candidate tests, row counts and {staged}) + {scan} + "
            f"{FILL_KERNEL[mode]} (placement, with the consumer's transpose slots)")


def measured_traffic(kernels):
    """HBM bytes per launch of `kernels` (summed) from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of the plain 10^6-sphere box (copy-probe
    calibrated), or None when it holds none of them."""
    try:
        with open(TRAFFIC_FILE) as fh:
            data = json.load(fh)
    except (OSError, ValueError):
        return None, None
    total, found = 0.0, 0
    for k in kernels:
        # a kernel's entries by its unqualified name (template arguments
        # aside); of several instantiations, the one dispatched most
        hits = [v for name, v in data.get("kernels", {}).items() if kernel_base_name(name) == k]
        if hits:
            v = max(hits, key=lambda h: h.get("dispatches", 0))
            total += v["read_bytes"] + v["write_bytes"]
            found += 1
    if found != len(kernels):
        return None, None
    return total, os.path.relpath(TRAFFIC_FILE, ROOT)


def kernel_base_name(name):
    """'void ns::k_foo<true>(args...)' -> 'k_foo'."""
    n = name.split("(")[0]
    if n.startswith("void "):
        n = n[5:]
    return n.split("<")[0].rsplit("::", 1)[-1]


def box_counts(b):
    c = b.counts()
    return c["nlocal"], c["nghost"], c["ncells"], c["npairs"], c["rebuilds"]


def run_box(ctx, gen, steps, warmup, period, consumer=True):
    """Steps one box (plain form); returns its timing and counts."""
    import soil_amd
    d, pos, vel, mass, rad = gen
    n = len(pos)
    box = soil_amd.Box(ctx, d)
    box.upload(pos, mass, rad, vel=vel)
    if consumer:
        box.set_consumer(1000.0)
    box.setup()
    if warmup:
        box.step(warmup)
    box.sync()
    r0 = box.counts()["rebuilds"]
    box.set_timing(period)
    t0 = time.perf_counter()
    box.step(steps)
    box.sync()
    el = time.perf_counter() - t0
    tm = box.timing()
    tm["period"] = period
    N, G, C, P, r1 = box_counts(box)
    box.close()
    rebuilds = r1 - r0
    poly = bool(np.ptp(rad) > 1e-15)
    le = bool(d.has_shear)
    sub = rl.run_bytes(N, G, C, P, steps, rebuilds, poly, le, consumer=False)
    tot = rl.run_bytes(N, G, C, P, steps, rebuilds, poly, le, consumer=consumer)
    res = {
        "n": n, "seconds": el, "particle_steps_per_s": n * steps / el, "ms_per_step": 1e3 * el / steps,
        "rebuilds": rebuilds, "nghost": G, "npairs": P, "ncells": C, "csr": tm,
        "step_roofline": dict(rl.roofline(tot, el), bytes_per_step=tot / steps,
                              substrate_frac=rl.roofline(sub, el)["frac"],
                              model="soil_amd/roofline.py run_bytes: Verlet, trigger, zeroing, comm per step; "
                                    "pbc, sort, permute, borders, binning, CSR build, save per rebuild; "
                                    "consumer 84 B per atom and pair"),
    }
    return res


def csr_roofline(tm):
    """The whole CSR build (count + offset scan + fill), event-timed, against HBM peak."""
    if tm["builds"] <= 0 or tm["csr_ms"] <= 0:
        return None
    nb = tm["builds"]
    per = tm["csr_bytes"] / nb
    build_ms = tm["csr_ms"] / nb
    count_ms = tm["count_ms"] / nb
    fill_ms = tm["fill_ms"] / nb
    mode = csr_mode()
    traffic, src = measured_traffic([k for k in (COUNT_KERNEL[mode], FILL_KERNEL[mode]) if k])
    r = dict(rl.roofline(per, build_ms * 1e-3))
    r.update({
        "kernel": csr_kernel_text(mode),
        "bound": "hbm",
        "traffic": traffic,
        "traffic_source": src,
        "algorithmic_bytes_per_launch": per,
        "launches_timed": nb,
        "timing_period": tm["period"],
        "avg_launch_ms": build_ms,
        "count_ms": count_ms,
        "scan_ms": build_ms - count_ms - fill_ms,
        "fill_ms": fill_ms,
        "count_only_frac": per / (count_ms * 1e-3) / 1e9 / rl.HBM_PEAK_GBS,
        "csr_mode": mode,
    })
    return r


def run_sheared(ctx, steps, warmup):
    import soil_amd
    log("sheared: 4M box (generating)")
    gen = soil_amd.gen_box(3)
    res = run_box(ctx, gen, steps, warmup, 2)
    out = {
        "workload": "4,000,000 spheres r~U[0.45,0.55], box 3:1:1, Lees-Edwards xy strain rate 0.1, 64 Gaussian "
                    "blobs (sigma 4 diameters, half the spheres) over a uniform background, phi 0.40, "
                    "harmonic consumer; every step rebuilds",
        "particles": res["n"], "particle_steps_per_s": res["particle_steps_per_s"],
        "ms_per_step": res["ms_per_step"], "steps": steps, "rebuilds": res["rebuilds"],
        "nghost": res["nghost"], "npairs": res["npairs"], "step_roofline": res["step_roofline"],
        "csr_build": csr_roofline(res["csr"]),
    }
    return out, gen


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


def ensemble_bytes(boxes, gens, steps):
    tot = 0.0
    for b, g in zip(boxes, gens):
        N, G, C, P, _ = box_counts(b)
        tot += rl.run_bytes(N, G, C, P, steps, steps, True, bool(g[0].has_shear), consumer=True)
    return tot


def run_batched(ctx, steps_batched, steps_rr, steps_each, warmup, graph=True):
    import soil_amd
    gens = ensemble()
    ntot = sum(len(g[1]) for g in gens)
    # plain, one step per call: the boxes in turn, each call waiting for its box
    boxes = make_boxes(ctx, gens)
    for b in boxes:
        b.step(1)
        b.sync()
    t0 = time.perf_counter()
    for _ in range(steps_rr):
        for b in boxes:
            b.step(1)
    for b in boxes:
        b.sync()
    t_rr = time.perf_counter() - t0
    # plain, n steps per call: each box in turn runs its steps in one call
    t0 = time.perf_counter()
    for b in boxes:
        b.step(steps_each)
    for b in boxes:
        b.sync()
    t_each = time.perf_counter() - t0
    for b in boxes:
        b.close()
    del boxes
    # batched: one launch per phase for all boxes, graph-captured step
    boxes = make_boxes(ctx, gens)
    batch = soil_amd.Batch(ctx, boxes)
    batch.step(1, graph=graph)
    batch.sync()
    tuned = batch.autotune() if graph else None  # the autotuner times captured graphs
    batch.step(max(warmup, 1), graph=graph)
    batch.sync()
    t0 = time.perf_counter()
    batch.step(steps_batched, graph=graph)
    batch.sync()
    t_batch = time.perf_counter() - t0
    batch.close()
    step_b = ensemble_bytes(boxes, gens, steps_batched)
    for b in boxes:
        b.close()
    rr = ntot * steps_rr / t_rr
    each = ntot * steps_each / t_each
    batched = ntot * steps_batched / t_batch
    return {
        "workload": "256 periodic Lees-Edwards boxes, N log-uniform in [1k,64k], phi U[0.20,0.55], "
                    "r U[0.4,0.6], strain rate U[0,0.2], harmonic consumer, every step rebuilds",
        "particles": ntot,
        "batched_particle_steps_per_s": batched,
        "plain_particle_steps_per_s": each,
        "plain_one_step_per_call_particle_steps_per_s": rr,
        "speedup": batched / each,
        "speedup_vs_one_step_per_call": batched / rr,
        "batched_steps": steps_batched,
        "plain_steps_per_call": steps_each,
        "plain_one_step_rounds": steps_rr,
        "batched_ms_per_step": 1e3 * t_batch / steps_batched,
        "plain_ms_per_ensemble_step": 1e3 * t_each / steps_each,
        "step_roofline": dict(rl.roofline(step_b, t_batch), bytes_per_step=step_b / steps_batched),
        "autotuned": tuned,
        "note": "plain = each box stepped alone on the single-box path, n steps per call (one host wait per "
                "step's decision, ~30 launches per step); one_step_per_call adds a state read per call",
    }


def run_served(ctx):
    from soil_amd import serve
    sweep = []
    main = None
    for rate in SERVED_RATES:
        log(f"served: {rate} jobs/s")
        st = serve.bench_served(ctx, njobs=120, rate=rate)
        row = {k: st[k] for k in ("jobs_done", "jobs_failed", "throughput_particle_steps_per_s", "step_p50_ms",
                                  "step_p99_ms", "gpu_step_p50_ms", "gpu_step_p99_ms", "update_gap_p99_ms",
                                  "completion_p50_s", "completion_p99_s", "wait_p99_s", "mean_active_jobs",
                                  "reconfigs", "wall_s")}
        row["rate_jobs_per_s"] = rate
        row["p99_within_bound"] = st["step_p99_ms"] < LATENCY_BOUND_MS
        sweep.append(row)
        if main is None:
            main = st
    main["latency_bound_ms"] = LATENCY_BOUND_MS
    main["rate_sweep"] = sweep
    held = [r["rate_jobs_per_s"] for r in sweep if r["p99_within_bound"]]
    main["bound_holds_up_to_jobs_per_s"] = max(held) if held else None
    main["step_latency_note"] = ("step latency = per job, the wall time between its consecutive state updates "
                                 "(admission, then each chunk), re-batching and read-back gaps included, over "
                                 "the steps each update brought; gpu_step = a chunk's device time over its steps")
    return main


# ── CPU oracle lines (the restatement of the reference's algorithm) ──
def _ref_box(gen):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_ref import RefBox
    d, pos, vel, mass, rad = gen
    r = RefBox(d)
    r.add_atoms(pos, mass, rad, vel=vel)
    r.set_consumer(1000.0)
    r.setup()
    return r


def cpu_one_core(gen, target_s, max_steps):
    r = _ref_box(gen)
    n = len(gen[1])
    total, steps = 0.0, 0
    while total < target_s and steps < max_steps:
        total += r.time_steps(1)
        steps += 1
    return n * steps / total, steps, total


def cpu_k_cores(gens, steps, k):
    """Independent boxes on k threads (each thread steps its share of boxes;
    the oracle releases the interpreter while it runs)."""
    boxes = [_ref_box(g) for g in gens]
    ntot = sum(len(g[1]) for g in gens)
    shares = [boxes[t::k] for t in range(k)]
    barrier = threading.Barrier(k + 1)

    def work(share):
        barrier.wait()
        for b in share:
            b.time_steps(steps)

    th = [threading.Thread(target=work, args=(s,)) for s in shares]
    for t in th:
        t.start()
    barrier.wait()
    t0 = time.perf_counter()
    for t in th:
        t.join()
    el = time.perf_counter() - t0
    return ntot * steps / el, el


def run_cpu(c2_gen, c3_gen):
    import soil_amd
    out = {}
    model = cpu_model()
    k = host_workers()
    v, s, t = cpu_one_core(c2_gen, 12.0, 200)
    main = {"value": v, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"same 1,000,000-sphere box, first {s} steps from the initial state ({t:.1f} s, single "
                      f"thread, {model})"}
    g1 = soil_amd.gen_box(1)
    v1, s1, t1 = cpu_one_core(g1, 30.0, 100)
    out["c1_plain_4k"] = {"value": v1, "cores": 1, "sample": f"4,000 monodisperse spheres, {s1} steps ({t1:.2f} s)"}
    if c3_gen is not None:
        v3, s3, t3 = cpu_one_core(c3_gen, 8.0, 20)
        out["c3_sheared_4M"] = {"value": v3, "cores": 1,
                                "sample": f"4,000,000-sphere sheared box, first {s3} steps ({t3:.1f} s)"}
    gens4 = ensemble()
    v4, t4 = cpu_k_cores(gens4, 4, k)
    out["c4_ensemble"] = {"value": v4, "cores": k,
                          "sample": f"the 256-box ensemble, 4 steps per box, boxes dealt over {k} threads "
                                    f"({t4:.1f} s)"}
    gens5 = [soil_amd.gen_box(5, j) for j in range(2 * k)]
    v5, t5 = cpu_k_cores(gens5, 20, k)
    out["c5_jobs"] = {"value": v5, "cores": k,
                      "sample": f"{2 * k} served-stream jobs, 20 steps each, over {k} threads ({t5:.1f} s)"}
    main["other_lines"] = out
    main["host"] = f"{model}, {os.cpu_count()} cores visible"
    return main


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--parts", default="plain,sheared,batched,served,cpu")
    ap.add_argument("--sheared-steps", type=int, default=20)
    ap.add_argument("--csr-steps", type=int, default=1000,
                    help="steps of the plain box without the consumer, timing the CSR build alone (0: skip)")
    ap.add_argument("--batched-steps", type=int, default=200)
    ap.add_argument("--plain-ensemble-steps", type=int, default=10)
    ap.add_argument("--batched-graph", type=int, default=1,
                    help="0: launch the batch's phases directly, without the captured step graph or the "
                         "autotuner (for profilers that cannot follow graph launches)")
    ap.add_argument("--plain-steps-per-call", type=int, default=50)
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
    c2_gen = None
    if "plain" in parts or "cpu" in parts:
        c2_gen = soil_amd.gen_box(2)
    if "plain" in parts:
        log("plain: 1M box")
        res = run_box(ctx, c2_gen, args.steps, args.warmup, 0)
        # the build's roofline: the same box from the same start, every build timed
        rf = run_box(ctx, c2_gen, max(args.steps, ROOFLINE_STEPS), args.warmup, 1)
        out["value"] = res["particle_steps_per_s"]
        out["ms_per_step"] = res["ms_per_step"]
        out["config"]["rebuilds_in_timed_steps"] = res["rebuilds"]
        out["config"]["nghost"] = res["nghost"]
        out["config"]["npairs"] = res["npairs"]
        out["roofline"] = csr_roofline(rf["csr"])
        if out["roofline"] is not None:
            out["roofline"]["timed_run_steps"] = max(args.steps, ROOFLINE_STEPS)
        out["step_roofline"] = res["step_roofline"]
        # the same box without the consumer: the fill is then the CSR's own
        # placement (no transpose slots), so count + scan + fill is the build alone
        if args.csr_steps > 0 and out["roofline"] is not None:
            pure = run_box(ctx, c2_gen, args.csr_steps, args.warmup, 1, consumer=False)  # every build timed
            out["roofline"]["without_consumer"] = csr_roofline(pure["csr"])
        log(f"plain: {out['value']:.3e} particle-steps/s, {res['ms_per_step']:.3f} ms/step")
    c3_gen = None
    if "sheared" in parts:
        out["sheared"], c3_gen = run_sheared(ctx, args.sheared_steps, 3)
        log(f"sheared: {out['sheared']['particle_steps_per_s']:.3e} particle-steps/s")
    if "batched" in parts:
        log("batched: 256-box ensemble")
        out["batched"] = run_batched(ctx, args.batched_steps, args.plain_ensemble_steps, args.plain_steps_per_call,
                                     args.warmup, graph=bool(args.batched_graph))
        log(f"batched: {out['batched']['speedup']:.2f}x over plain")
    if "served" in parts:
        out["served"] = run_served(ctx)
    if "cpu" in parts:
        log("cpu baselines")
        out["cpu_baseline"] = run_cpu(c2_gen, c3_gen)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
