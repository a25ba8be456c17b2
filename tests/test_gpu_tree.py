"""GPU tree BP against the reference (src/graph/graph.c:376-605, driver
src/c_benchmark/main.c:205-215), bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

import bp_ctypes as B
from conftest import TREE_CASES, bits_report, golden, graph_for_case, messages_of, nan_equal_bits, same_bits, states_of

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", TREE_CASES)
def test_tree_handle_bitwise(lib, case):
    print(f"case {case}", flush=True)
    d = golden(case)
    g = graph_for_case(d, levels=True)
    S = d["tree_states"].shape[-1]
    dev = B.Device(g)
    dev.tree()
    dev.download(g)
    assert same_bits(messages_of(g, "tree", S), d["tree_messages"])
    got = states_of(g, S)
    assert same_bits(got, d["tree_states"]), bits_report(got, d["tree_states"])
    dev.free()
    lib.graph_destroy(g)


@pytest.mark.parametrize("case", TREE_CASES)
def test_tree_dropin_sequence_bitwise(lib, case):
    """The caller's own level loop through graph.h, as the reference drivers do."""
    print(f"case {case}", flush=True)
    d = golden(case)
    g = graph_for_case(d, levels=True)
    S = d["tree_states"].shape[-1]
    L = g.contents.num_levels
    lib.propagate_using_levels_start(g)
    for i in range(1, L - 1):
        lib.propagate_using_levels(g, i)
    lib.reset_visited(g)
    for i in range(L - 1, 0, -1):
        lib.propagate_using_levels(g, i)
    assert same_bits(messages_of(g, "tree", S), d["tree_messages"])
    lib.marginalize(g)
    assert same_bits(states_of(g, S), d["tree_states"])
    lib.graph_destroy(g)


def test_tree_known_answer(lib):
    """src/example-belief-propagation/example-belief-propagation.c:72-107."""
    d = golden("example")
    g = graph_for_case(d, levels=True)
    dev = B.Device(g)
    dev.tree()
    dev.download(g)
    st = states_of(g, 2)
    want = np.array([[0.521531100478469, 0.47846889952153115], [0.9090909090909091, 0.09090909090909091],
                     [0.1652892561983471, 0.8347107438016529], [1.0, 0.0]])
    assert np.abs(st - want).max() < 1e-4
    dev.free()
    lib.graph_destroy(g)


def test_tree_mixed_states_match_oracle(lib, orc):
    bn = B.generate("c1", states_hi=4, seed=1002)
    g = B.bn_to_graph(bn, levels=True)
    h = B.bn_to_graph(bn, levels=True)
    dev = B.Device(g)
    dev.tree()
    dev.download(g)
    orc.tree_run(h)
    assert same_bits(B.messages(g, "tree"), B.messages(h, "tree"))
    assert same_bits(B.states(g), B.states(h))
    dev.free()
    lib.graph_destroy(g)
    lib.graph_destroy(h)
    lib.bp_bn_free(bn)


@pytest.mark.parametrize("case", ["finite", "inf"])
@pytest.mark.parametrize("path", ["handle", "dropin"])
def test_tree_mixed_states_hand_worked(lib, orc, path, case):
    """Zero-extension on the GPU against exact hand-worked values
    (tests/mixed_hand.py: a 3-state node combining 2-entry messages and
    sending on a 3x3 edge, every padded slot filled with 7.0), and bit for
    bit against the oracle.  The inf case puts an infinite entry in a's
    table to b (a NaN message entry; NaN positions compared, not NaN bit
    patterns)."""
    import mixed_hand as MH

    print(f"mixed star, {path}, {case}", flush=True)
    exp = MH.star_expected() if case == "finite" else MH.star_inf_expected()
    build = MH.star_graph if case == "finite" else MH.star_graph_inf
    eq = same_bits if case == "finite" else nan_equal_bits
    g = build()
    if path == "handle":
        dev = B.Device(g)
        dev.tree()
        dev.download(g)
        dev.free()
    else:
        L = g.contents.num_levels
        lib.propagate_using_levels_start(g)
        for i in range(1, L - 1):
            lib.propagate_using_levels(g, i)
        lib.reset_visited(g)
        for i in range(L - 1, 0, -1):
            lib.propagate_using_levels(g, i)
        lib.marginalize(g)
    assert MH.check_star(g, exp) == []
    h = build()
    orc.tree_run(h)
    assert eq(B.messages(g, "tree"), B.messages(h, "tree"))
    assert eq(B.states(g), B.states(h))
    lib.graph_destroy(g)
    lib.graph_destroy(h)


def test_duplicate_level_entries_are_rejected(lib):
    """levels_to_nodes naming a node twice is refused (the per-level plan
    derives each node's sub-wave from its position in the level)."""
    d = golden("tree10k")
    g = graph_for_case(d, levels=True)
    G = g.contents
    V = G.current_num_vertices
    lv = B.view_u32(G.levels_to_nodes, 2 * V)
    lv[V + 1] = lv[V]
    lib.propagate_using_levels_start(g)
    assert "twice" in B.last_error()
    lib.graph_destroy(g)
