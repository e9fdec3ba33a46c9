"""Summarises rocprofv3 --pmc runs of SQ counters per kernel.

    python -m soil_amd.sqsum <results.db> [<results.db> ...] [--kernel SUBSTRING]

Each database is one counter pass (rocprofv3 collects a few SQ counters per
pass). Prints, per kernel, the average of every counter per dispatch and the
ratios derived from the ones present:
  valu_per_wave    SQ_INSTS_VALU / SQ_WAVES (VALU instructions a wave issues)
  lds_per_wave     SQ_INSTS_LDS / SQ_WAVES
  vmem_rd_per_wave SQ_INSTS_VMEM_RD / SQ_WAVES
  lane_util        SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU): the share of
                   lanes active while the wave executes VALU work
  valu_share       SQ_ACTIVE_INST_VALU / SQ_WAVE_CYCLES: cycles a resident
                   wave spends issuing VALU work
  waves_resident   SQ_WAVE_CYCLES / SQ_BUSY_CYCLES (waves in flight per busy
                   SQ cycle, summed over the SQs that count)
  wait_share       SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES, lds_wait_share likewise
"""
import sqlite3
import sys


def per_kernel(paths, like=None):
    out = {}
    for p in paths:
        c = sqlite3.connect(p)
        rows = c.execute("select name, counter_name, count(*), avg(counter_value) from pmc_events "
                         "group by name, counter_name")
        for name, counter, n, v in rows:
            if like and like not in name:
                continue
            out.setdefault(name, {})[counter] = v
            out[name].setdefault("_dispatches", n)
    return out


def derived(c):
    d = {}
    w = c.get("SQ_WAVES")
    if w:
        for src, key in (("SQ_INSTS_VALU", "valu_per_wave"), ("SQ_INSTS_LDS", "lds_per_wave"),
                         ("SQ_INSTS_VMEM_RD", "vmem_rd_per_wave"), ("SQ_INSTS_SALU", "salu_per_wave")):
            if src in c:
                d[key] = c[src] / w
    if c.get("SQ_ACTIVE_INST_VALU") and "SQ_THREAD_CYCLES_VALU" in c:
        d["lane_util"] = c["SQ_THREAD_CYCLES_VALU"] / (64.0 * c["SQ_ACTIVE_INST_VALU"])
    wc = c.get("SQ_WAVE_CYCLES")
    if wc:
        if "SQ_ACTIVE_INST_VALU" in c:
            d["valu_share"] = c["SQ_ACTIVE_INST_VALU"] / wc
        if "SQ_WAIT_INST_ANY" in c:
            d["wait_share"] = c["SQ_WAIT_INST_ANY"] / wc
        if "SQ_WAIT_INST_LDS" in c:
            d["lds_wait_share"] = c["SQ_WAIT_INST_LDS"] / wc
        if c.get("SQ_BUSY_CYCLES"):
            d["waves_resident"] = wc / c["SQ_BUSY_CYCLES"]
    return d


def summarize(paths, like=None):
    data = per_kernel(paths, like)
    lines = []
    for name, c in sorted(data.items(), key=lambda kv: -kv[1].get("SQ_WAVE_CYCLES", 0.0)):
        lines.append(f"{name[:90]}  ({c['_dispatches']} dispatches)")
        for k in sorted(x for x in c if not x.startswith("_")):
            lines.append(f"    {k:28s} {c[k]:16.1f}")
        for k, v in derived(c).items():
            lines.append(f"    {k:28s} {v:16.3f}")
    return "\n".join(lines)


if __name__ == "__main__":
    args = sys.argv[1:]
    like = None
    if "--kernel" in args:
        k = args.index("--kernel")
        like = args[k + 1]
        del args[k:k + 2]
    print(summarize(args, like))
