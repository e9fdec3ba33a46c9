"""The kernel-trace summary (soil_amd.profsum) on a small synthetic rocpd-like database.

The per-kernel table and the per-step timeline are computed from a `kernels` table with
the columns rocprofv3's database view has (name, start, end, duration, in ns); the steps
here are laid out by hand so that every figure is known.
"""
import sqlite3

import pytest

from soil_amd import profsum

NS = 1000  # 1 us


def make_db(path, rows):
    c = sqlite3.connect(path)
    c.execute("create table kernels (name text, start integer, end integer, duration integer)")
    c.executemany("insert into kernels values (?, ?, ?, ?)", [(n, a, b, b - a) for n, a, b in rows])
    c.commit()
    c.close()


def timeline():
    """Three steps: a set-up step (skipped), a rebuild step, a step without one."""
    t, rows = 0, []

    def k(name, dur_us, gap_us=0):
        nonlocal t
        t += gap_us * NS
        rows.append((name, t, t + dur_us * NS))
        t += dur_us * NS

    k("void soil_hip::k_initial<true, 1>(soil_hip::BoxDev* const*)", 30)
    k("soil_hip::k_step_end(soil_hip::BoxDev* const*)", 2)
    # rebuild step: 40 + 60 + 2 us busy, 5 us idle (the 3 us before it precede the first step counted)
    k("void soil_hip::k_initial<true, 1>(soil_hip::BoxDev* const*)", 40, gap_us=3)
    k("soil_hip::k_csr_tcount(soil_hip::BoxDev* const*)", 60, gap_us=5)
    k("soil_hip::k_step_end(soil_hip::BoxDev* const*)", 2)
    # no rebuild: the count exits at once; 40 + 1 + 2 us busy, 20 + 1 us idle
    k("void soil_hip::k_initial<true, 1>(soil_hip::BoxDev* const*)", 40, gap_us=20)
    k("soil_hip::k_csr_tcount(soil_hip::BoxDev* const*)", 1)
    k("soil_hip::k_step_end(soil_hip::BoxDev* const*)", 2, gap_us=1)
    return rows


def test_short_names_keep_template_arguments():
    assert profsum.short_name("void soil_hip::k_initial<true, 1>(soil_hip::BoxDev* const*)") == "k_initial<true, 1>"
    assert profsum.short_name("soil_hip::k_csr_tcount(soil_hip::BoxDev* const*)") == "k_csr_tcount"


def test_step_timeline_splits_busy_and_idle_by_step_kind():
    st = profsum.step_rows(timeline(), skip=1)
    r, f = st["rebuild"], st["no rebuild"]
    assert r["steps"] == 1 and f["steps"] == 1
    # the first step counted starts at its first dispatch (the 3 us before it are not its own)
    assert (r["span"], r["busy"], r["dispatches"]) == (107 * NS, 102 * NS, 3)
    assert (f["span"], f["busy"], f["dispatches"]) == (64 * NS, 43 * NS, 3)
    assert r["idle_before"] == {"k_csr_tcount": 5 * NS}
    assert f["idle_before"] == {"k_initial<true, 1>": 20 * NS, "k_step_end": 1 * NS}


def test_overlapping_dispatches_count_once_as_busy():
    rows = [("k_step_end", 0, NS), ("k_step_end", NS, 2 * NS),
            ("k_a", 3 * NS, 10 * NS), ("k_b", 4 * NS, 12 * NS), ("k_step_end", 12 * NS, 13 * NS)]
    st = profsum.step_rows(rows, skip=1)["no rebuild"]
    assert st["steps"] == 2
    # second step: 1 us idle before k_a, k_a and k_b overlap (3..12 busy), then k_step_end
    assert st["span"] == NS + 11 * NS and st["busy"] == NS + 10 * NS


def test_reports_read_the_database(tmp_path):
    db = tmp_path / "t.db"
    make_db(str(db), timeline())
    rep = profsum.steps_report(str(db), skip=1)
    assert "rebuild: 1 steps, span 107.0 us, busy 102.0 us, idle 5.0 us" in rep
    assert "no rebuild: 1 steps, span 64.0 us, busy 43.0 us, idle 21.0 us" in rep
    table = profsum.summarize(str(db))
    assert "total kernel time 0.18 ms over 8 dispatches" in table


if __name__ == "__main__":
    pytest.main([__file__, "-q"])
