"""Summarises rocprofv3 --pmc runs (FETCH_SIZE / WRITE_SIZE, KiB per dispatch).

    python -m soil_amd.pmcsum <fetch_results.db> <write_results.db> <probe_bytes> [--json out.json]

The calibration copy kernel (soil_hip::k_copy_probe, `probe_bytes` read and
written per launch) gives the scale from counter KiB to bytes for each counter.
"""
import sqlite3
import sys


def per_kernel(path, counter):
    c = sqlite3.connect(path)
    rows = c.execute("select name, count(*), avg(counter_value), avg(duration) from pmc_events "
                     "where counter_name = ? group by name", (counter,))
    return {r[0]: (r[1], r[2], r[3]) for r in rows}


def summarize(fetch_db, write_db, probe_bytes):
    f = per_kernel(fetch_db, "FETCH_SIZE")
    w = per_kernel(write_db, "WRITE_SIZE")
    probe = [k for k in f if "k_copy_probe" in k][0]
    fs = probe_bytes / (f[probe][1] * 1024.0)
    ws = probe_bytes / (w[probe][1] * 1024.0)
    lines = [f"calibration: {probe_bytes} B copy -> FETCH_SIZE x{fs:.3f}, WRITE_SIZE x{ws:.3f} (KiB -> bytes)",
             f"{'kernel':56s} {'calls':>6s} {'read_MB':>9s} {'write_MB':>9s} {'total_MB':>9s}"]
    order = sorted(f, key=lambda k: -(f[k][1] * fs + w.get(k, (0, 0, 0))[1] * ws) * f[k][0])
    for k in order[:30]:
        rd = f[k][1] * 1024.0 * fs / 1e6
        wr = w.get(k, (0, 0.0, 0))[1] * 1024.0 * ws / 1e6
        lines.append(f"{k[:56]:56s} {f[k][0]:6d} {rd:9.2f} {wr:9.2f} {rd + wr:9.2f}")
    return "\n".join(lines), fs, ws


def per_kernel_bytes(fetch_db, write_db, probe_bytes):
    """{kernel: {"dispatches", "read_bytes", "write_bytes"}} per dispatch, calibrated."""
    f = per_kernel(fetch_db, "FETCH_SIZE")
    w = per_kernel(write_db, "WRITE_SIZE")
    probe = [k for k in f if "k_copy_probe" in k][0]
    fs = probe_bytes / (f[probe][1] * 1024.0)
    ws = probe_bytes / (w[probe][1] * 1024.0)
    out = {}
    for k in f:
        out[k] = {"dispatches": f[k][0], "read_bytes": f[k][1] * 1024.0 * fs,
                  "write_bytes": w.get(k, (0, 0.0, 0))[1] * 1024.0 * ws}
    return out, fs, ws


if __name__ == "__main__":
    text, _, _ = summarize(sys.argv[1], sys.argv[2], float(sys.argv[3]))
    print(text)
    if "--json" in sys.argv:
        import json
        path = sys.argv[sys.argv.index("--json") + 1]
        data, fs, ws = per_kernel_bytes(sys.argv[1], sys.argv[2], float(sys.argv[3]))
        with open(path, "w") as fh:
            json.dump({"fetch_scale": fs, "write_scale": ws, "kernels": data}, fh, indent=1)
