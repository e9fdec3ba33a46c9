"""Summarises a rocprofv3 kernel-trace database (rocpd SQLite) per kernel.

    python -m soil_amd.profsum <results.db> [top] [--window SUBSTRING]
    python -m soil_amd.profsum <results.db> --steps [--end SUBSTRING] [--mark SUBSTRING:US] [--skip K]

With --window, only dispatches that start and end between the first start
and the last end of a kernel whose name contains SUBSTRING are counted (for
example the batched steps of a run that also sets boxes up and steps them
one at a time).

With --steps, the dispatches are cut into steps, each ending with a kernel
whose name contains the --end substring (default k_step_end), and each step
is classed as a rebuild step when a kernel named like the --mark substring
ran longer than US microseconds in it (default k_csr_tcount:10: the CSR count
exits at once on a step that does not rebuild). Per class it prints the
step's span (from the end of the previous step to its own end; the first
step counted from its own first dispatch), the time
some kernel was running (busy), the rest (idle: launch gaps and host round
trips), and which kernels the idle time came before. The first K steps are
skipped (default 5: set-up and warm-up).
"""
import sqlite3
import sys
from collections import defaultdict


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


def short_name(name):
    """'void ns::k_foo<1>(args...)' -> 'k_foo<1>'."""
    n = name.split("(")[0]
    if n.startswith("void "):
        n = n[5:]
    head, lt, tail = n.partition("<")
    return head.rsplit("::", 1)[-1] + lt + tail


def step_rows(rows, end="k_step_end", mark="k_csr_tcount", mark_ns=10000, skip=5):
    """rows: (name, start_ns, end_ns) in start order. Returns {class: stats} over the steps after `skip`."""
    steps, cur = [], []
    for r in rows:
        cur.append(r)
        if end in r[0]:
            steps.append(cur)
            cur = []
    stats = {}
    # the first step counted starts at its own first dispatch: a host gap
    # before it (between two stepping calls) belongs to no step
    prev_end = None
    for s in steps[skip:]:
        kind = "rebuild" if any(mark in r[0] and r[2] - r[1] > mark_ns for r in s) else "no rebuild"
        t0 = prev_end if prev_end is not None else s[0][1]
        st = stats.setdefault(kind, {"steps": 0, "span": 0, "busy": 0, "dispatches": 0,
                                     "idle_before": defaultdict(int), "busy_by": defaultdict(int)})
        st["steps"] += 1
        st["dispatches"] += len(s)
        # busy: the union of the dispatch intervals inside the step
        covered = t0
        for r in s:
            a, b = max(r[1], covered), r[2]
            if r[1] > covered:
                st["idle_before"][short_name(r[0])] += r[1] - covered
            if b > a:
                st["busy"] += b - a
            covered = max(covered, r[2])
            st["busy_by"][short_name(r[0])] += r[2] - r[1]
        t1 = max(r[2] for r in s)
        st["span"] += max(t1 - t0, 0)
        prev_end = t1
    return stats


def steps_report(path, end="k_step_end", mark="k_csr_tcount", mark_us=10.0, skip=5, top=8):
    c = sqlite3.connect(path)
    rows = list(c.execute("select name, start, end from kernels order by start"))
    stats = step_rows(rows, end, mark, int(mark_us * 1000), skip)
    out = [f"steps cut at '{end}', rebuild steps marked by '{mark}' > {mark_us:g} us, first {skip} skipped"]
    for kind in ("rebuild", "no rebuild"):
        st = stats.get(kind)
        if not st:
            continue
        n = st["steps"]
        out.append(f"{kind}: {n} steps, span {st['span'] / n / 1e3:.1f} us, busy {st['busy'] / n / 1e3:.1f} us, "
                   f"idle {(st['span'] - st['busy']) / n / 1e3:.1f} us, {st['dispatches'] / n:.1f} dispatches per step")
        idle = sorted(st["idle_before"].items(), key=lambda x: -x[1])[:top]
        out.append("  idle before: " + ", ".join(f"{k} {v / n / 1e3:.1f}" for k, v in idle))
        busy = sorted(st["busy_by"].items(), key=lambda x: -x[1])[:top]
        out.append("  busy in: " + ", ".join(f"{k} {v / n / 1e3:.1f}" for k, v in busy))
    return "\n".join(out)


if __name__ == "__main__":
    args = sys.argv[1:]

    def opt(flag, default):
        if flag in args:
            k = args.index(flag)
            v = args[k + 1]
            del args[k:k + 2]
            return v
        return default

    if "--steps" in args:
        args.remove("--steps")
        end = opt("--end", "k_step_end")
        mark, _, us = opt("--mark", "k_csr_tcount:10").partition(":")
        skip = int(opt("--skip", "5"))
        print(steps_report(args[0], end, mark, float(us or 10), skip))
    else:
        window = opt("--window", None)
        print(summarize(args[0], int(args[1]) if len(args) > 1 else 40, window))
