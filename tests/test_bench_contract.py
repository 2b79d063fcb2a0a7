"""bench.py's command-line contract, CPU only: more than one GPU is refused
with a message (this build runs on one GPU), and the algorithmic byte count
of the roofline is the stated formula."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_algorithmic_bytes_formula():
    sys.path.insert(0, ROOT)
    import bench
    # fluid nodes x 2 x Q x bytes per population + 1 B node class per node
    assert bench.algorithmic_bytes(16777216, 11740933, 19, 4) == 11740933 * 2 * 19 * 4 + 16777216
    assert bench.algorithmic_bytes(16777216, 11740933, 19, 2) == 909088124
