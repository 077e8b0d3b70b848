"""bench.py's algorithmic work figure equals what the oracle's op counter
measures on the benchmark workload (so the roofline cannot drift from the code)."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _load(path, name):
    spec = importlib.util.spec_from_file_location(name, path)
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_flops_per_sample_step_matches_counter(pkg):
    bench = _load(os.path.join(ROOT, "bench.py"), "bench_mod")
    count_ops = _load(os.path.join(ROOT, "oracle", "count_ops.py"), "count_ops_mod")
    measured = count_ops.count(pkg.scenarios.c1_four_drones())
    assert abs(bench.FLOPS_PER_SAMPLE_STEP - measured) < 0.01, measured


def test_stratrace_flops_match_counter(pkg):
    bench = _load(os.path.join(ROOT, "bench.py"), "bench_mod")
    count_ops = _load(os.path.join(ROOT, "oracle", "count_ops.py"), "count_ops_mod")
    measured = count_ops.count_sr(pkg)
    for key in ("sr_race", "pr_sr_race", "bm_race"):
        assert abs(bench.SR_FLOPS_PER_STEP[key] - measured[key]) < 0.01, (key, measured[key])


def test_hiddenobs_flops_match_counter(pkg):
    bench = _load(os.path.join(ROOT, "bench.py"), "bench_mod")
    count_ops = _load(os.path.join(ROOT, "oracle", "count_ops.py"), "count_ops_mod")
    measured = count_ops.count_ho(pkg)
    assert abs(bench.SR_FLOPS_PER_STEP["ho_race"] - measured["ho_race"]) < 0.01, measured


def test_dronerace_flops_match_counter(pkg):
    bench = _load(os.path.join(ROOT, "bench.py"), "bench_mod")
    count_ops = _load(os.path.join(ROOT, "oracle", "count_ops.py"), "count_ops_mod")
    measured = count_ops.count_drace(pkg)
    assert abs(bench.SR_FLOPS_PER_STEP["drace_race"] - measured["drace_race"]) < 0.01, measured


def test_eight_drone_flops_match_counter(pkg):
    bench = _load(os.path.join(ROOT, "bench.py"), "bench_mod")
    count_ops = _load(os.path.join(ROOT, "oracle", "count_ops.py"), "count_ops_mod")
    scn = pkg.scenarios.c4_eight_drones(4096)
    measured = count_ops.count(scn)
    assert abs(bench.C4_FLOPS_PER_SAMPLE_STEP - measured) < 0.01, measured
    assert bench.C4_REFINE_FLOPS == count_ops.count_refine(scn, n_grad=3, ck=10)


def test_noise_flops_match_counter():
    bench = _load(os.path.join(ROOT, "bench.py"), "bench_mod")
    count_ops = _load(os.path.join(ROOT, "oracle", "count_ops.py"), "count_ops_mod")
    assert bench.NOISE_FLOPS_PER_SAMPLE_STEP == count_ops.NOISE_FLOPS_PER_STEP
