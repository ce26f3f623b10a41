"""bench.py's command line without a GPU: the one-GPU refusal and the help
text naming every leg."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, capture_output=True,
                          text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = run(["--gpus", "2"])
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout == ""


def test_help_lists_the_legs():
    r = run(["--help"])
    assert r.returncode == 0
    for flag in ("--workload", "--precision", "--adjoint-only", "--adjoint-falling", "--accel",
                 "--mixed-drift", "--accel-probe", "--traffic-probe", "--adjoint-probe", "--no-cpu-baseline"):
        assert flag in r.stdout, flag
