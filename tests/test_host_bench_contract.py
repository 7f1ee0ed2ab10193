"""bench.py's command-line contract on a machine without a GPU: one GPU only (--gpus N>1
exits with a message naming the reason)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300, cwd=ROOT)
    assert r.returncode != 0
    msg = (r.stdout + r.stderr).lower()
    assert "one gpu" in msg or "single gpu" in msg or "1 gpu" in msg, msg[-500:]


def test_point_flop_model_matches_the_stated_counts():
    """bench.point_flops restates the reference algorithm's per-unit counts: with N = 12, four
    local iterations and five residual evaluations (the start at S = 0 and one trial per
    iteration), 5 F_res + 4 (F_jac + F_LU) + F_tan = 34,318 flops (FMA = 2) for a point with
    its tangent; with N = 24 each residual costs 41 x 12 more and each Jacobian 165 x 12 more;
    a residual-only pass adds no tangent."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    f12 = bench.point_flops(12, 1, 4, 5)
    assert f12 == 5 * (41 * 12 + 350) + 4 * (165 * 12 + 2574 + 648) + 9300 == 34318
    assert bench.point_flops(24, 1, 4, 5) - f12 == 5 * 41 * 12 + 4 * 165 * 12
    assert bench.point_flops(12, 10, 40, 50) == 10 * f12
    assert f12 - bench.point_flops(12, 0, 4, 5) == 9300
