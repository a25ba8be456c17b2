"""Non-finite table entries: an infinite weight in a network file reads as
+inf, and the messages it feeds become inf or NaN.  The reference skips NaN
factors in the node products and counts a NaN message difference as 0
(src/graph/graph.c: the combine and delta loops); the GPU kernels must give
the oracle's values with NaNs at the same places.  This exercises the
paths whose products run over zero padding instead of masks (mixed-state
tables, the batch kernel's pooled rows) and the node phases that replace a
message's NaN entries by 1 once (one lane per node, and two lanes per node
in the batch kernel).

NaN bit patterns are not compared: x86 produces the negative default NaN
for inf - inf, the GPU the positive one.  Every other value is compared bit
for bit."""
from __future__ import annotations

import numpy as np
import pytest

import bp_ctypes as B
from conftest import nan_equal_bits

pytestmark = pytest.mark.gpu


def mismatch_report(a: np.ndarray, b: np.ndarray, limit: int = 6) -> str:
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    na, nb = np.isnan(a), np.isnan(b)
    bad = (na != nb) | (~na & ~nb & (a.view(np.uint32) != b.view(np.uint32)))
    idx = np.argwhere(bad)
    rows = ", ".join(f"{tuple(int(v) for v in i)}: {a[tuple(i)]!r} vs {b[tuple(i)]!r}" for i in idx[:limit])
    return f"{len(idx)} entries differ; {rows}"


def poisoned_graphs(lib, copies: int, seed: int, states_hi: int, fraction: float):
    """copies of one seeded mixed-state network, each with the same table
    entries (within the entry's x rows and y columns) set to +inf"""
    bn = B.generate("c4net", num_nodes=1200, states_hi=states_hi, seed=seed)
    gs = [B.bn_to_graph(bn) for _ in range(copies)]
    lib.bp_bn_free(bn)
    G = gs[0].contents
    E = G.current_num_edges
    x = B.view_u32(G.edges_x_dim, E).copy()
    y = B.view_u32(G.edges_y_dim, E).copy()
    rng = np.random.default_rng(seed)
    picks = rng.choice(E, size=max(1, int(E * fraction)), replace=False)
    cells = [(int(e), int(rng.integers(x[e])), int(rng.integers(y[e]))) for e in picks]
    for g in gs:
        J = B.view_f32(g.contents.edges_joint_probabilities, E * B.MAX_STATES * B.MAX_STATES)
        for e, i, j in cells:
            J[e * B.MAX_STATES * B.MAX_STATES + i * B.MAX_STATES + j] = np.inf
    return gs


@pytest.mark.parametrize("states_hi", [4, 8])
def test_plain_path_matches_oracle_with_infinite_entries(lib, orc, states_hi):
    print(f"plain, states 2-{states_hi}", flush=True)
    g, h = poisoned_graphs(lib, 2, 4100 + states_hi, states_hi, 0.02)
    d = B.Device(g, 0)
    d.init()
    d.iterate(6)
    d.download(g)
    d.free()
    orc.init_previous_edge(h)
    for _ in range(6):
        orc.loopy_iteration(h)
    got, want = B.states(g), B.states(h)
    assert np.isnan(B.messages(h, "cur")).any(), "the poisoned entries should produce NaN messages"
    assert nan_equal_bits(got, want), mismatch_report(got, want)
    mg, mh = B.messages(g, "cur"), B.messages(h, "cur")
    assert nan_equal_bits(mg, mh), mismatch_report(mg, mh)
    for x in (g, h):
        lib.graph_destroy(x)


def test_batch_matches_plain_with_infinite_entries(lib):
    """the two-lane shape (its node phase takes two lanes per node in chunks
    of at most 32 nodes) against the plain path"""
    print("batch, states 2-8", flush=True)
    gs = poisoned_graphs(lib, 2, 4200, 8, 0.02)
    b = B.Batch()
    b.add(gs[0])
    b.upload(0)
    b.set_config(1, 2, 0)
    b.init()
    b.iterate(6)
    b.download(0, gs[0])
    b.free()
    d = B.Device(gs[1], 0)
    d.init()
    d.iterate(6)
    d.download(gs[1])
    d.free()
    assert np.isnan(B.messages(gs[1], "cur")).any()
    assert nan_equal_bits(B.states(gs[0]), B.states(gs[1])), mismatch_report(B.states(gs[0]), B.states(gs[1]))
    m0, m1 = B.messages(gs[0], "cur"), B.messages(gs[1], "cur")
    assert nan_equal_bits(m0, m1), mismatch_report(m0, m1)
    for g in gs:
        lib.graph_destroy(g)
