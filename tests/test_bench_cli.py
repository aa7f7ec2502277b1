"""bench.py's command-line contract that needs no GPU: more than one GPU is
refused with a message and exit status 2 before anything touches a device;
the help lists the profiling modes."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    p = _run("--gpus", "2")
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""


def test_help_lists_profiling_modes():
    p = _run("--help")
    assert p.returncode == 0
    for flag in ("--kernel-only", "--solve-only", "--dense-only", "--clouds-only",
                 "--batch-kernel-only", "--steps", "--warmup", "--gpus"):
        assert flag in p.stdout, flag


def test_more_than_one_gpu_is_refused_in_profiling_modes():
    """The refusal comes before the profiling modes are dispatched, so none
    of them silently runs on one GPU."""
    for mode in ("--kernel-only", "--solve-only", "--dense-only", "--clouds-only",
                 "--batch-kernel-only"):
        p = _run("--gpus", "4", mode)
        assert p.returncode == 2, mode
        assert "one GPU" in p.stderr, mode
        assert p.stdout == "", mode
