"""bench.py's command line without a GPU: more than one GPU is refused with
a message (the benchmark measures one), so is an unknown form name, and its
documented options are listed."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def run(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2", "--steps", "1", "--warmup", "0")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_help_lists_the_contract_options():
    r = run("--help")
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--forms", "--no-cpu", "--batch-nets", "--no-batch-plain"):
        assert opt in r.stdout


def test_unknown_form_is_refused():
    r = run("--forms", "adjoint,treee")
    assert r.returncode == 2
    assert "treee" in r.stderr and r.stdout == ""
