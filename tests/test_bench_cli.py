"""bench.py's command-line contract on the CPU: a run on more than one GPU (by flag or by a
launcher's WORLD_SIZE) is refused with exit status 2 and a message, before any device work;
--help lists the profiling modes."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, capture_output=True, text=True,
                          env=env, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _run(["--gpus", "2"])
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""  # no JSON line


def test_launcher_world_size_is_refused():
    r = _run([], {"WORLD_SIZE": "4", "RANK": "0", "LOCAL_RANK": "0"})
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_help_lists_the_modes():
    r = _run(["--help"])
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--gpus", "--profile-solve", "--profile-batch", "--layout"):
        assert opt in r.stdout
