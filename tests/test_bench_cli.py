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


def test_traffic_figures_are_found_for_templated_kernels():
    # the committed counter summary names the placement by its template
    # instantiation (k_csr_tplace<true>); the bench line must still find it
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    assert bench.kernel_base_name("void soil_hip::k_csr_tplace<true>(soil_hip::BoxDev* const*, int)") == "k_csr_tplace"
    assert bench.kernel_base_name("soil_hip::k_csr_tcount(soil_hip::BoxDev* const*)") == "k_csr_tcount"
    traffic, src = bench.measured_traffic(["k_csr_tcount", "k_csr_tplace"])
    assert src == "profiles/traffic_plain_1M.json"
    assert 100e6 < traffic < 300e6
    assert bench.ROOFLINE_STEPS >= 10  # the build's roofline rests on at least that many timed builds' run
