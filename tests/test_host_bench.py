"""Host-side checks of bench.py: the graph chunking of the timed steps and the
single-GPU guard (no device needed)."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_graph_chunk_divides_the_timed_steps():
    b = load_bench()
    for steps in range(0, 1201):
        g = b.graph_chunk(steps)
        if steps < 2:
            assert g == 0
            continue
        even = steps - (steps % 2)
        assert g % 2 == 0 and 2 <= g <= 50 and even % g == 0, (steps, g)
        # no larger even divisor within the cap was skipped
        assert all(even % h for h in range(g + 2, min(50, even) + 1, 2)), (steps, g)
    assert b.graph_chunk(500) == 50 and b.graph_chunk(100) == 50 and b.graph_chunk(5) == 4


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
