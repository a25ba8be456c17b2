"""Shared test helpers.

- `gpu` marks tests that need an MI355X; everything else runs on the CPU.
- Golden fixtures (tests/golden/*.npz) hold what the compiled reference
  computed (tests/golden/make_golden.py).  A case's input graph is rebuilt
  either from its stored structure (replayed through the graph API, edge by
  edge, which reproduces the reference's arrays exactly) or, for generated
  networks, from the same seeded preset.
"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
INPUTS = GOLDEN / "inputs"
REFERENCE = Path("/root/reference/src/benchmark_files")
for p in (ROOT / "belief-propagation_amd", ROOT / "oracle", GOLDEN, ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

import bp_ctypes as B  # noqa: E402

MAX_STATES = 10


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def lib():
    lib = B.load()
    lib.bp_set_abort_on_error(0)
    return lib


@pytest.fixture(scope="session")
def orc():
    import oracle_ctypes

    oracle_ctypes.load()
    return oracle_ctypes


def golden(name: str) -> dict:
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def graph_from_structure(d: dict):
    """Replays a stored structure through create_graph / graph_add_node /
    graph_set_node_state / graph_add_edge, then builds both CSR views."""
    lib = B.load()
    meta = d["meta"]
    V, E, total_e = int(meta[0]), int(meta[1]), int(meta[3])
    g = lib.create_graph(V, total_e)
    nv = d["node_num_vars"]
    states = d["node_states"]
    names = d["node_names"].reshape(V, 50)
    obs = d["observed"]
    buf = (C.c_float * MAX_STATES)()
    for v in range(V):
        n = int(nv[v])
        lib.graph_add_node(g, n, bytes(names[v]).split(b"\0", 1)[0][:50])
        if obs[v]:
            for k in range(n):
                buf[k] = float(states[v, k])
            lib.graph_set_node_state(g, v, n, buf)
    J = d["edges_J"]
    S = J.shape[1]
    tab = (C.c_float * (MAX_STATES * MAX_STATES))()
    src, dst, xs, ys = d["edges_src"], d["edges_dst"], d["edges_x"], d["edges_y"]
    for e in range(E):
        x, y = int(xs[e]), int(ys[e])
        for i in range(x):
            for j in range(y):
                tab[i * MAX_STATES + j] = float(J[e, i, j])
        lib.graph_add_edge(g, int(src[e]), int(dst[e]), x, y, tab)
    lib.set_up_src_nodes_to_edges(g)
    lib.set_up_dest_nodes_to_edges(g)
    assert S >= 1
    return g


def overrides(d: dict) -> dict:
    import ast

    if "overrides" not in d:
        return {}
    return dict(ast.literal_eval(str(d["overrides"])))


def graph_for_case(d: dict, levels: bool = False):
    """The input graph of a fixture case."""
    lib = B.load()
    if "preset" in d:
        bn = B.generate(str(d["preset"]), **overrides(d))
        g = B.bn_to_graph(bn, levels=levels)
        lib.bp_bn_free(bn)
        return g
    g = graph_from_structure(d)
    if levels:
        lib.init_levels_to_nodes(g)
    return g


def states_of(g, S: int) -> np.ndarray:
    return B.states(g)[:, :S]


def messages_of(g, which: str, S: int) -> np.ndarray:
    return B.messages(g, which)[:, :S]


def same_bits(a: np.ndarray, b: np.ndarray) -> bool:
    a = np.ascontiguousarray(a, dtype=b.dtype)
    return a.shape == b.shape and a.tobytes() == np.ascontiguousarray(b).tobytes()


def nan_equal_bits(a: np.ndarray, b: np.ndarray) -> bool:
    """same_bits, except that any NaN equals any NaN: x86 produces the
    negative default NaN for inf / inf or inf - inf, the GPU the positive
    one.  NaN positions must agree."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    if a.shape != b.shape:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    return bool(np.array_equal(na, nb)) and a[~na].tobytes() == b[~nb].tobytes()


def bits_report(a: np.ndarray, b: np.ndarray) -> str:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return f"shape {a.shape} vs {b.shape}"
    diff = np.abs(a - b)
    idx = np.unravel_index(np.argmax(diff), diff.shape)
    return f"{int((a != b).sum())} entries differ, max |diff| {diff.max():.3e} at {idx}: {a[idx]!r} vs {b[idx]!r}"


SMALL_LOOPY_CASES = ["example", "dog_xml", "bf_1000_2000_1", "c1", "lattice16", "c5_small", "c4net_small"]
TREE_CASES = ["example", "dog_xml", "bf_1000_2000_1", "c1", "lattice16", "tree10k"]
