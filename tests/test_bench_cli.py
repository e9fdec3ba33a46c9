"""bench.py's command line on the CPU: the multi-GPU refusal and the defaults."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)
    assert r.stdout.strip() == "" or not r.stdout.strip().startswith("{")  # no bench line


def test_default_parts_cover_the_whole_metric():
    # the metric names plain, batched and the served stream's latencies; the
    # sheared 4M box and the CPU baselines ride along
    src = open(os.path.join(ROOT, "bench.py")).read()
    assert 'ap.add_argument("--parts", default="plain,sheared,batched,served,cpu")' in src
