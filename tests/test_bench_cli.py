"""bench.py's command-line contract on the CPU: more than one GPU is refused with a
message before any device is touched, and every workload is selectable."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=REPO)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2", "--no-traffic")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_workloads_are_listed():
    sys.path.insert(0, REPO)
    import bench

    r = run("--help")
    assert r.returncode == 0
    for name in bench.WORKLOADS:
        assert name in r.stdout
    assert {"lfts64", "replicas48", "realspace3d", "reduce128", "c1", "c2", "c2d", "c3", "c4"} <= set(bench.WORKLOADS)
