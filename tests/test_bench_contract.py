"""bench.py's host-side contract, on the CPU: the roofline's `traffic` field
is read from the committed counter profile of the bench's grid (a number, of
the size the counters give for 136 algorithmic bytes per cell), and
`--gpus` other than 1 is refused with a message before any GPU work."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_traffic_line_parses():
    import bench
    t = bench.measured_traffic(545, 385)
    assert t is not None
    cells = 544 * 384
    assert 1.0 <= t / (136 * cells) <= 1.5, t
    assert bench.measured_traffic(17, 13) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
