"""The CPU oracle (oracle/bp_oracle.c) pinned against the reference.

Each check compares the oracle with what the compiled reference wrote
(tests/golden), bit for bit, or with a closed form.  A planted mistake --
the classic transposed table of textbook BP -- is shown to fail the
fixtures, so they do pin the J.b orientation the reference uses
(src/graph/graph.c:899).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import bp_ctypes as B
import mixed_hand as MH
from conftest import (SMALL_LOOPY_CASES, TREE_CASES, bits_report, golden, graph_for_case, messages_of,
                      same_bits, states_of)


def _S(d):
    return int(d["node_num_vars"].max()) if "node_num_vars" in d else int(d["loopy_states"].shape[-1])


@pytest.mark.parametrize("case", SMALL_LOOPY_CASES)
def test_loopy_iterations_bitwise(lib, orc, case):
    d = golden(case)
    g = graph_for_case(d)
    S = d["loopy_states"].shape[-1]
    orc.init_previous_edge(g)
    assert same_bits(messages_of(g, "prev", S), d["loopy_init_prev"])
    state_iters = list(d["loopy_state_iters"])
    msg_iters = list(d["loopy_msg_iters"])
    n = int(max(state_iters))
    for it in range(1, n + 1):
        orc.loopy_iteration(g)
        # the reference's delta after this iteration, f32 sequential
        assert np.float32(orc.delta(g)).tobytes() == d["loopy_deltas"][it - 1].tobytes(), f"delta {it}"
        if it in state_iters:
            got = states_of(g, S)
            want = d["loopy_states"][state_iters.index(it)]
            assert same_bits(got, want), f"{case} iteration {it}: {bits_report(got, want)}"
        if it in msg_iters:
            k = msg_iters.index(it)
            assert same_bits(messages_of(g, "cur", S), d["loopy_cur"][k])
            assert same_bits(messages_of(g, "prev", S), d["loopy_prev"][k])
    lib.graph_destroy(g)


@pytest.mark.parametrize("case", SMALL_LOOPY_CASES)
def test_openmp_matches_serial(lib, orc, case):
    d = golden(case)
    g = graph_for_case(d)
    S = d["loopy_states"].shape[-1]
    orc.init_previous_edge(g)
    for _ in range(int(d["loopy_deltas"].shape[0])):
        orc.loopy_iteration(g, threads=4)
    assert same_bits(states_of(g, S), d["omp_states"])
    lib.graph_destroy(g)


@pytest.mark.parametrize("case", SMALL_LOOPY_CASES)
def test_until_matches_reference(lib, orc, case):
    d = golden(case)
    g = graph_for_case(d)
    S = d["until_states"].shape[-1]
    orc.init_previous_edge(g)
    it, _ = orc.loopy_until(g, float(d["until_conv"]), int(d["until_max"]))
    assert it == int(d["until_iterations"][0])
    assert same_bits(states_of(g, S), d["until_states"])
    assert same_bits(messages_of(g, "cur", S), d["until_cur"])
    assert same_bits(messages_of(g, "prev", S), d["until_prev"])
    lib.graph_destroy(g)


@pytest.mark.parametrize("case", TREE_CASES)
def test_tree_bitwise(lib, orc, case):
    d = golden(case)
    g = graph_for_case(d, levels=True)
    S = d["tree_states"].shape[-1]
    orc.tree_run(g)
    assert same_bits(messages_of(g, "tree", S), d["tree_messages"])
    got = states_of(g, S)
    assert same_bits(got, d["tree_states"]), bits_report(got, d["tree_states"])
    lib.graph_destroy(g)


def test_known_answer_tree(lib, orc):
    """src/example-belief-propagation/example-belief-propagation.c:72-107 (+-1e-4)."""
    d = golden("example")
    g = graph_for_case(d, levels=True)
    orc.tree_run(g)
    st = states_of(g, 2)
    want = np.array([[0.521531100478469, 0.47846889952153115], [0.9090909090909091, 0.09090909090909091],
                     [0.1652892561983471, 0.8347107438016529], [1.0, 0.0]])
    assert np.abs(st - want).max() < 1e-4
    lib.graph_destroy(g)


# The example is left out of the FP64 comparisons: under -Dfloat=double the
# reference builds its tables from double literals (0.9), while the FP64
# oracle widens the f32 graph (0.9f).  The other cases replay f32 inputs.
@pytest.mark.parametrize("case", ["c1", "lattice16", "c5_small", "c4net_small"])
def test_fp64_oracle_bitwise(lib, orc, case):
    """The double instantiation against the reference's graph core built with -Dfloat=double."""
    d = golden(case)
    g = graph_for_case(d)
    S = d["fp64_states"].shape[-1]
    o = orc.FP64(g)
    o.init()
    iters = list(d["fp64_state_iters"])
    for it in range(1, int(max(iters)) + 1):
        o.iterate()
        if it in iters:
            got = o.states()[:, :S]
            want = d["fp64_states"][iters.index(it)]
            assert same_bits(got, want), f"{case} iteration {it}: {bits_report(got, want)}"
    assert same_bits(o.messages(previous=True)[:, :S], d["fp64_prev"])
    assert same_bits(o.messages(previous=False)[:, :S], d["fp64_cur"])
    del o
    lib.graph_destroy(g)


@pytest.mark.parametrize("case", ["c1", "lattice16", "tree10k"])
def test_fp64_tree_bitwise(lib, orc, case):
    d = golden(case)
    g = graph_for_case(d, levels=True)
    S = d["fp64_tree_states"].shape[-1]
    o = orc.FP64(g)
    o.tree()
    assert same_bits(o.states()[:, :S], d["fp64_tree_states"])
    del o
    lib.graph_destroy(g)


# ---------------------------------------------------------------------------
# closed forms


def _two_node_graph(lib, J, b_src, observed_dst=None):
    """src -> dst with table J (x*y); src state b_src."""
    x, y = J.shape
    g = lib.create_graph(2, 1)
    st = (C.c_float * 10)(*([float(v) for v in b_src] + [0.0] * (10 - x)))
    lib.graph_add_and_set_node_state(g, x, b"src", st)
    lib.graph_add_node(g, y, b"dst")
    if observed_dst is not None:
        o = (C.c_float * 10)(*([float(v) for v in observed_dst] + [0.0] * (10 - y)))
        lib.graph_set_node_state(g, 1, y, o)
    tab = (C.c_float * 100)()
    for i in range(x):
        for j in range(y):
            tab[i * 10 + j] = float(J[i, j])
    lib.graph_add_edge(g, 0, 1, x, y, tab)
    lib.set_up_src_nodes_to_edges(g)
    lib.set_up_dest_nodes_to_edges(g)
    return g


def test_single_edge_closed_form(lib, orc):
    """init message m = N(J.b) with J[src_state][dst_state] (graph.c:896-909), f32 left-to-right."""
    J = np.array([[0.7, 0.2, 0.1], [0.05, 0.9, 0.3], [0.4, 0.4, 0.2]], dtype=np.float32)
    b = np.array([0.6, 0.3, 0.1], dtype=np.float32)
    g = _two_node_graph(lib, J, b)
    orc.init_previous_edge(g)
    m = messages_of(g, "prev", 3)[0]
    acc = np.zeros(3, dtype=np.float32)
    for i in range(3):
        for j in range(3):
            acc[i] = np.float32(acc[i] + np.float32(J[i, j] * b[j]))
    s = np.float32(0)
    for i in range(3):
        s = np.float32(s + acc[i])
    assert same_bits(m, (acc / s).astype(np.float32))
    # the transposed (textbook) orientation gives a different message
    mt = (J.T @ b) / (J.T @ b).sum()
    assert np.abs(m - mt).max() > 1e-3
    lib.graph_destroy(g)


def test_node_without_in_edges_keeps_normalised_state(lib, orc):
    J = np.array([[0.5, 0.5], [0.1, 0.9]], dtype=np.float32)
    g = _two_node_graph(lib, J, [3.0, 1.0])
    orc.init_previous_edge(g)
    for _ in range(3):
        orc.loopy_iteration(g)
    st = states_of(g, 2)
    assert same_bits(st[0], np.array([0.75, 0.25], dtype=np.float32))
    lib.graph_destroy(g)


def test_zero_product_gives_zero_message(lib, orc):
    """sum <= 0 is replaced by 1 (graph.c:442-444): an all-zero product stays all-zero."""
    J = np.zeros((2, 2), dtype=np.float32)
    g = _two_node_graph(lib, J, [1.0, 1.0])
    orc.init_previous_edge(g)
    assert same_bits(messages_of(g, "prev", 2)[0], np.zeros(2, dtype=np.float32))
    lib.graph_destroy(g)


def test_observed_one_hot_leaf_stays_one_hot(lib, orc):
    J = np.array([[0.8, 0.2], [0.3, 0.7]], dtype=np.float32)
    g = _two_node_graph(lib, J, [0.5, 0.5], observed_dst=[0.0, 1.0])
    orc.init_previous_edge(g)
    for _ in range(5):
        orc.loopy_iteration(g)
    st = states_of(g, 2)
    assert st[1, 0] == 0.0 and st[1, 1] == 1.0
    lib.graph_destroy(g)


def test_planted_transpose_is_caught(lib, orc):
    """Plant textbook BP's orientation (J^T.b) in the oracle's input: the
    reference fixtures reject it, so they pin the orientation."""
    d = golden("c1")
    g = graph_for_case(d)
    G = g.contents
    E = G.current_num_edges
    J = B.view_f32(G.edges_joint_probabilities, E * 100).reshape(E, 10, 10)
    J[:, :2, :2] = np.transpose(J[:, :2, :2], (0, 2, 1)).copy()
    orc.init_previous_edge(g)
    for _ in range(3):
        orc.loopy_iteration(g)
    assert not same_bits(states_of(g, 2), d["loopy_states"][2])
    lib.graph_destroy(g)


def test_planted_belief_reset_is_caught(lib, orc):
    """Plant the textbook non-accumulating belief (states reset to the prior
    each iteration): the fixtures reject it (graph.c:1010 accumulates)."""
    d = golden("c1")
    g = graph_for_case(d)
    S = 2
    prior = states_of(g, S).copy()
    orc.init_previous_edge(g)
    st_ptr = B.view_f32(g.contents.node_states, g.contents.current_num_vertices * 10).reshape(-1, 10)
    for _ in range(3):
        st_ptr[:, :S] = prior
        orc.loopy_iteration(g)
    assert not same_bits(states_of(g, S), d["loopy_states"][2])
    lib.graph_destroy(g)


def test_mixed_state_counts_zero_extension(lib, orc):
    """Networks with different neighbouring state counts are undefined in the
    reference (uninitialised reads, graph.c:856/899, :103-104/455); the
    oracle defines them as zero-extension: OpenMP and serial agree and every
    value is finite."""
    d = golden("c1_mixed")
    assert not bool(d["square"])
    bn = B.generate("c1", states_hi=4, seed=1002)
    g = B.bn_to_graph(bn)
    h = B.bn_to_graph(bn)
    orc.init_previous_edge(g)
    orc.init_previous_edge(h)
    for _ in range(10):
        orc.loopy_iteration(g)
        orc.loopy_iteration(h, threads=4)
    a, b = B.states(g), B.states(h)
    assert same_bits(a, b)
    assert np.isfinite(a).all()
    lib.graph_destroy(g)
    lib.graph_destroy(h)
    lib.bp_bn_free(bn)


def test_mixed_states_hand_worked_loopy(lib, orc):
    """Zero-extension pinned by hand: a 2-state node sending on a 2x3 edge
    (graph.c:896-903, the third buffer entry reads 0), a 3-state node
    combining a 2-entry message (graph.c:450-459, the third entry reads 0)
    and the loopy belief update over it (graph.c:985-1024), against exact
    values worked out in tests/mixed_hand.py, with every padded slot
    filled with 7.0; serial and OpenMP node loops."""
    exp = MH.chain_expected()
    for threads in (1, 2):
        g = MH.chain_graph()
        orc.init_previous_edge(g)
        assert MH.check_chain_init(g, exp) == []
        orc.loopy_iteration(g, threads=threads)
        assert MH.check_chain_iteration(g, exp) == []
        lib.graph_destroy(g)


def test_mixed_states_hand_worked_nan_message(lib, orc):
    """Zero extension past a NaN message: with an infinite entry in a's
    table to b, a's message is (NaN, 0); b's products skip the NaN entry,
    multiply by the 0 and by the 0 read past the message's two entries
    (graph.c:450-459, 985-1024), against values worked out by hand in
    tests/mixed_hand.py; serial and OpenMP node loops."""
    exp = MH.chain_inf_expected()
    for threads in (1, 2):
        g = MH.chain_graph_inf()
        orc.init_previous_edge(g)
        assert MH.check_chain_init(g, exp) == []
        orc.loopy_iteration(g, threads=threads)
        assert MH.check_chain_iteration(g, exp) == []
        lib.graph_destroy(g)


def test_mixed_states_hand_worked_tree(lib, orc):
    """Zero-extension in tree BP (graph.c:461-538, 540-596): a 3-state node
    combining two 2-entry messages and sending on a 3x3 edge, against exact
    hand-worked values (tests/mixed_hand.py)."""
    exp = MH.star_expected()
    g = MH.star_graph()
    G = g.contents
    assert G.num_levels == exp["num_levels"]
    assert list(B.view_u32(G.levels_to_nodes, 2 * G.current_num_vertices)[G.current_num_vertices:]) == exp["levels"]
    orc.tree_run(g)
    assert MH.check_star(g, exp) == []
    lib.graph_destroy(g)


def test_mixed_states_hand_worked_tree_nan_message(lib, orc):
    """The star with an infinite entry in a's table to b: tree BP skips the
    NaN entry of a's message at b and still reads 0 past its two entries
    (graph.c:450-459, 472-521), against tests/mixed_hand.py."""
    exp = MH.star_inf_expected()
    g = MH.star_graph_inf()
    orc.tree_run(g)
    assert MH.check_star(g, exp) == []
    lib.graph_destroy(g)


@pytest.mark.parametrize("plant", ["pad_one", "stale", "norm_stride"])
def test_mixed_hand_worked_rejects_planted_mistakes(lib, orc, plant):
    """Each plausible mistake in the mixed-state branch, planted in a copy of
    the oracle (ORC_PLANT_* in oracle/bp_oracle_core.h) -- reads beyond a
    state count padded with 1.0, reads of whatever the padded row holds,
    normalisation over the padded row width -- fails every hand-worked
    check: the init messages, the first iteration and the tree."""
    pl = orc.load_plant(plant)
    exp = MH.chain_expected()
    g = MH.chain_graph()
    orc.init_previous_edge(g, lib=pl)
    assert MH.check_chain_init(g, exp) != []
    lib.graph_destroy(g)
    g = MH.chain_graph()
    orc.init_previous_edge(g)           # correct init, then the planted iteration
    orc.loopy_iteration(g, lib=pl)
    assert MH.check_chain_iteration(g, exp) != []
    lib.graph_destroy(g)
    g = MH.star_graph()
    orc.tree_run(g, lib=pl)
    assert MH.check_star(g, MH.star_expected()) != []
    lib.graph_destroy(g)


@pytest.mark.parametrize("case", ["example", "dog_xml", "c1", "lattice16", "c4net_small"])
def test_edge_schedule_against_reference(lib, orc, case):
    """The edge-centric schedule with its defects fixed vs the reference's
    (src/graph/graph.c:1186-1227), which normalises node 0 only: its beliefs
    are the oracle's up to a per-node scale, so rows are compared after
    normalising them (rounding differs by that rescaling: rtol 1e-5); the
    messages, which are scale-free, match to the same tolerance."""
    d = golden(case)
    g = graph_for_case(d)
    S = d["edge3_states"].shape[-1]
    orc.init_previous_edge(g)
    for _ in range(3):
        orc.edge_iteration(g)
    mine = states_of(g, S).astype(np.float64)
    ref = d["edge3_states"].astype(np.float64)
    sums = ref.sum(axis=1, keepdims=True)
    ref_n = np.where(sums > 0, ref / np.where(sums > 0, sums, 1), ref)
    np.testing.assert_allclose(mine, ref_n, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(messages_of(g, "cur", S), d["edge3_cur"], rtol=1e-5, atol=1e-7)
    lib.graph_destroy(g)
