"""bench.py's command line on the CPU: more than one GPU is refused with a
message (the benchmark measures one MI355X), and the algorithmic operation
count per voxel-iteration is the reference method's (QUATERNION.cpp:154-200,
375-393: 4 subtractions, a 23-operation polar power and a 28-operation
Hamilton product per root term; division 40, scaling 4, magnitude 8)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr or "single MI355X" in r.stderr
    assert r.stdout.strip() == ""


def test_basic_op_count():
    import bench
    # rational: per root term 4 + 23 + 28 = 55, minus the first product of each
    # polynomial (28 each) plus the division (40), the scaling (4) and the magnitude (8)
    assert bench.basic_ops_per_voxel_iter(32, 32) == 55 * 64 - 56 + 52
    assert bench.basic_ops_per_voxel_iter(32, 32) == 3516
    # polynomial only: no division
    assert bench.basic_ops_per_voxel_iter(8, 0) == 55 * 8 - 28 + 12
