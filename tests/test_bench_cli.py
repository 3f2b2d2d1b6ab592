"""bench.py's command-line contract, checked without a GPU: more than one GPU
is refused with a message and a nonzero status, and the profiling driver
offers every operator it documents."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def run(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""   # no JSON line for a configuration that was not run


def test_profiling_driver_lists_its_operators():
    r = run("--help")
    assert r.returncode == 0
    for k in ("packed", "dense", "precond", "f32", "transpose", "spa", "rk4", "chebyshev", "bicg"):
        assert k in r.stdout
    bad = run("--apply-only", "--apply-kernel", "nonsense")
    assert bad.returncode != 0 and "invalid choice" in bad.stderr
