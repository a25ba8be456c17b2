"""bench.py without a GPU: more than one GPU is refused with
a message (the benchmark measures one), so is an unknown form name, its
documented options are listed, and its reference CPU leg agrees with the
oracle."""
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


def test_reference_cpu_leg_matches_the_oracle():
    """The bench's reference leg (the reference's own loop, compiled from its
    sources) gives the oracle's beliefs bit for bit on a small network."""
    import pytest

    if not (ROOT / "oracle" / "_ref" / "ref_harness_omp").exists():
        pytest.skip("the reference harness is not built")
    sys.path.insert(0, str(ROOT))
    import bench
    import bp_ctypes as B
    import oracle_ctypes as O

    bn = B.generate("c1")
    h = B.bn_to_graph(bn)
    O.init_previous_edge(h)
    for _ in range(bench.ITERATIONS):
        O.loopy_iteration(h, threads=2)
    r = bench.reference_cpu(bn, 2, B.states(h))
    B.load().graph_destroy(h)
    B.load().bp_bn_free(bn)
    assert r["kind"] == "reference" and "error" not in r, r
    assert r["parity_vs_gpu"] and r["value"] > 0
