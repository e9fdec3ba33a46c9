"""Summarises a rocprofv3 kernel-trace database (rocpd SQLite) per kernel.

    python -m soil_amd.profsum <results.db> [top]
"""
import sqlite3
import sys


def summarize(path, top=40):
    c = sqlite3.connect(path)
    rows = list(c.execute(
        "select name, count(*), sum(duration), avg(duration), min(duration), max(duration) "
        "from kernels group by name order by sum(duration) desc"))
    tot = sum(r[2] for r in rows) or 1
    out = [f"{'kernel':64s} {'calls':>7s} {'total_ms':>9s} {'avg_us':>9s} {'min_us':>8s} {'max_us':>8s} {'%':>6s}"]
    for r in rows[:top]:
        out.append(f"{r[0][:64]:64s} {r[1]:7d} {r[2] / 1e6:9.2f} {r[3] / 1e3:9.2f} {r[4] / 1e3:8.2f} "
                   f"{r[5] / 1e3:8.2f} {100 * r[2] / tot:6.2f}")
    out.append(f"total kernel time {tot / 1e6:.2f} ms over {sum(r[1] for r in rows)} dispatches")
    return "\n".join(out)


if __name__ == "__main__":
    print(summarize(sys.argv[1], int(sys.argv[2]) if len(sys.argv) > 2 else 40))
