"""Summarises a rocprofv3 kernel-trace database (rocpd SQLite) per kernel.

    python -m soil_amd.profsum <results.db> [top] [--window SUBSTRING]

With --window, only dispatches that start and end between the first start
and the last end of a kernel whose name contains SUBSTRING are counted (for
example the batched steps of a run that also sets boxes up and steps them
one at a time).
"""
import sqlite3
import sys


def summarize(path, top=40, window=None):
    c = sqlite3.connect(path)
    where, head = "", ""
    if window:
        t0, t1, n = c.execute("select min(start), max(end), count(*) from kernels where instr(name, ?) > 0",
                              (window,)).fetchone()
        if not n:
            raise SystemExit(f"no kernel named like {window!r}")
        where = f" where start >= {int(t0)} and end <= {int(t1)}"
        head = (f"window: first to last dispatch of '{window}' ({n} dispatches of it), "
                f"span {(t1 - t0) / 1e6:.2f} ms\n")
    rows = list(c.execute(
        "select name, count(*), sum(duration), avg(duration), min(duration), max(duration) "
        f"from kernels{where} group by name order by sum(duration) desc"))
    tot = sum(r[2] for r in rows) or 1
    out = [f"{'kernel':64s} {'calls':>7s} {'total_ms':>9s} {'avg_us':>9s} {'min_us':>8s} {'max_us':>8s} {'%':>6s}"]
    for r in rows[:top]:
        out.append(f"{r[0][:64]:64s} {r[1]:7d} {r[2] / 1e6:9.2f} {r[3] / 1e3:9.2f} {r[4] / 1e3:8.2f} "
                   f"{r[5] / 1e3:8.2f} {100 * r[2] / tot:6.2f}")
    out.append(f"total kernel time {tot / 1e6:.2f} ms over {sum(r[1] for r in rows)} dispatches")
    return head + "\n".join(out)


if __name__ == "__main__":
    args = sys.argv[1:]
    window = None
    if "--window" in args:
        k = args.index("--window")
        window = args[k + 1]
        del args[k:k + 2]
    print(summarize(args[0], int(args[1]) if len(args) > 1 else 40, window))
