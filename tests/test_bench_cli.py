"""The benchmark's command-line contract, checked without a GPU: the documented flags parse,
and a run asked for more than one GPU exits with a one-line JSON message (status 2) before
importing torch or touching a device."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_help_lists_the_contract_flags():
    p = run("--help")
    assert p.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup", "--no-cpu", "--no-build-scaling", "--build-sizes"):
        assert flag in p.stdout


def test_more_than_one_gpu_is_refused_with_json():
    p = run("--gpus", "2", "--steps", "1", "--warmup", "0")
    assert p.returncode == 2
    line = p.stdout.strip().splitlines()[-1]
    d = json.loads(line)
    assert d["n_gpus"] == 2 and "one GPU" in d["error"]
