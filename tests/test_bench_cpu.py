"""bench.py's command-line contract, without a GPU: one GPU only (any other
--gpus value exits with status 2 and a JSON error line), and every workload
named in the help."""

import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    p = run("--gpus", "2")
    assert p.returncode == 2
    line = json.loads(p.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_help_names_every_workload():
    p = run("--help")
    assert p.returncode == 0
    text = " ".join(p.stdout.split())
    for w in ("multislice", "adjoint", "large", "batch", "plasma", "cbbpm", "history", "grating"):
        assert f"{w}:" in text, w
