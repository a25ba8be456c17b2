"""GPU loopy BP (libbp_amd.so) against the reference's results.

The bar is bit-exactness: the kernels perform the reference's float
operations in its order (src/graph/graph.c:426-459,851-1032), so beliefs and
messages must equal the reference's to the last bit.  The convergence delta
is accumulated in FP64 on the GPU; it is compared with the same sum computed
in double by the oracle (relative 1e-12: only the order of the FP64 additions
differs).
"""
from __future__ import annotations

import numpy as np
import pytest

import bp_ctypes as B
from conftest import (SMALL_LOOPY_CASES, bits_report, golden, graph_for_case, messages_of, nan_equal_bits, same_bits,
                      states_of)

pytestmark = pytest.mark.gpu


def f32_sum_bound(g, total: float) -> float:
    """Error bound of a sequential f32 sum of Sum_e x_e non-negative terms."""
    G = g.contents
    terms = int(B.view_u32(G.edges_x_dim, G.current_num_edges).sum())
    return max(terms, 1) * 2.0 ** -24 * abs(total) + 1e-30


@pytest.mark.parametrize("case", SMALL_LOOPY_CASES)
def test_handle_iterations_bitwise(lib, orc, case):
    print(f"case {case}", flush=True)
    d = golden(case)
    g = graph_for_case(d)
    S = d["loopy_states"].shape[-1]
    dev = B.Device(g)
    dev.init()
    state_iters = list(d["loopy_state_iters"])
    msg_iters = list(d["loopy_msg_iters"])
    n = int(max(state_iters))
    # the oracle, in lockstep, provides the FP64 deltas to compare with
    h = graph_for_case(d)
    orc.init_previous_edge(h)
    dev.download(g)
    assert same_bits(messages_of(g, "prev", S), d["loopy_init_prev"])
    for it in range(1, n + 1):
        deltas = dev.iterate(1, want_deltas=True)
        orc.loopy_iteration(h)
        want = orc.delta64(h)
        assert abs(deltas[0] - want) <= 1e-12 * max(1.0, abs(want)), f"delta {it}: {deltas[0]} vs {want}"
        # the reference's own f32 running sum of the same non-negative terms:
        # within (terms - 1) * 2^-24 of the total
        assert abs(deltas[0] - float(d["loopy_deltas"][it - 1])) <= f32_sum_bound(g, want)
        if it in state_iters or it in msg_iters:
            dev.download(g)
        if it in state_iters:
            got = states_of(g, S)
            ref = d["loopy_states"][state_iters.index(it)]
            assert same_bits(got, ref), f"{case} iteration {it}: {bits_report(got, ref)}"
        if it in msg_iters:
            k = msg_iters.index(it)
            assert same_bits(messages_of(g, "cur", S), d["loopy_cur"][k])
            assert same_bits(messages_of(g, "prev", S), d["loopy_prev"][k])
    dev.free()
    lib.graph_destroy(g)
    lib.graph_destroy(h)


@pytest.mark.parametrize("case", SMALL_LOOPY_CASES)
def test_batched_iterations_equal_single_steps(lib, case):
    """n iterations from one cached HIP graph == the reference after n."""
    print(f"case {case}", flush=True)
    d = golden(case)
    g = graph_for_case(d)
    S = d["loopy_states"].shape[-1]
    dev = B.Device(g)
    dev.init()
    n = int(d["loopy_deltas"].shape[0])
    deltas = dev.iterate(n, want_deltas=True)
    dev.download(g)
    it = list(d["loopy_state_iters"])
    assert same_bits(states_of(g, S), d["loopy_states"][it.index(n)])
    for k in range(n):
        assert abs(deltas[k] - float(d["loopy_deltas"][k])) <= f32_sum_bound(g, deltas[k]), f"delta {k + 1}"
    dev.free()
    lib.graph_destroy(g)


@pytest.mark.parametrize("case", SMALL_LOOPY_CASES)
def test_dropin_api_bitwise(lib, case):
    """init_previous_edge + loopy_propagate_one_iteration through graph.h."""
    print(f"case {case}", flush=True)
    d = golden(case)
    g = graph_for_case(d)
    S = d["loopy_states"].shape[-1]
    lib.init_previous_edge(g)
    assert same_bits(messages_of(g, "prev", S), d["loopy_init_prev"])
    msg_iters = list(d["loopy_msg_iters"])
    for it in range(1, max(msg_iters) + 1):
        lib.loopy_propagate_one_iteration(g)
        if it in msg_iters:
            k = msg_iters.index(it)
            assert same_bits(messages_of(g, "cur", S), d["loopy_cur"][k]), f"iteration {it}"
            assert same_bits(messages_of(g, "prev", S), d["loopy_prev"][k]), f"iteration {it}"
    st_iters = list(d["loopy_state_iters"])
    if max(msg_iters) in st_iters:
        assert same_bits(states_of(g, S), d["loopy_states"][st_iters.index(max(msg_iters))])
    lib.graph_destroy(g)


@pytest.mark.parametrize("case", SMALL_LOOPY_CASES)
def test_until_bitwise(lib, case):
    """loopy_propagate_until and loopy_propagate_until_hip: iteration count,
    beliefs and both message arrays (with the reference's pointer swaps)."""
    print(f"case {case}", flush=True)
    d = golden(case)
    S = d["until_states"].shape[-1]
    conv, mx = float(d["until_conv"]), int(d["until_max"])
    for fn in ("dropin", "hip"):
        g = graph_for_case(d)
        lib.init_previous_edge(g)
        if fn == "dropin":
            it = lib.loopy_propagate_until(g, conv, mx)
        else:
            it = lib.loopy_propagate_until_hip(g, conv, mx)
        assert it == int(d["until_iterations"][0]), fn
        assert same_bits(states_of(g, S), d["until_states"]), fn
        assert same_bits(messages_of(g, "cur", S), d["until_cur"]), fn
        assert same_bits(messages_of(g, "prev", S), d["until_prev"]), fn
        lib.graph_destroy(g)


@pytest.mark.parametrize("case", ["finite", "inf"])
@pytest.mark.parametrize("path", ["handle", "dropin"])
def test_mixed_states_hand_worked(lib, orc, path, case):
    """Zero-extension on the GPU against exact hand-worked values
    (tests/mixed_hand.py: a 2-state node sending on a 2x3 edge, a 3-state
    node combining a 2-entry message, the belief update over it; every
    padded slot filled with 7.0), and bit for bit against the oracle.  The
    inf case puts an infinite entry in that edge's table, so the 2-entry
    message is NaN in one entry: the zero extension past it must still read
    0 (NaN positions compared, not NaN bit patterns)."""
    import mixed_hand as MH

    print(f"mixed chain, {path}, {case}", flush=True)
    exp = MH.chain_expected() if case == "finite" else MH.chain_inf_expected()
    build = MH.chain_graph if case == "finite" else MH.chain_graph_inf
    eq = same_bits if case == "finite" else nan_equal_bits
    g = build()
    h = build()
    if path == "handle":
        dev = B.Device(g)
        dev.init()
        dev.download(g)
    else:
        lib.init_previous_edge(g)
    assert MH.check_chain_init(g, exp) == []
    orc.init_previous_edge(h)
    assert eq(B.messages(g, "prev"), B.messages(h, "prev"))
    if path == "handle":
        dev.iterate(1)
        dev.download(g)
        dev.free()
    else:
        lib.loopy_propagate_one_iteration(g)
    assert MH.check_chain_iteration(g, exp) == []
    orc.loopy_iteration(h)
    assert eq(B.states(g), B.states(h))
    assert eq(B.messages(g, "prev"), B.messages(h, "prev"))
    assert eq(B.messages(g, "cur"), B.messages(h, "cur"))
    lib.graph_destroy(g)
    lib.graph_destroy(h)


def test_mixed_state_counts_match_oracle(lib, orc):
    """Different state counts per node: zero-extension, defined by the oracle."""
    bn = B.generate("c1", states_hi=4, seed=1002)
    g = B.bn_to_graph(bn)
    h = B.bn_to_graph(bn)
    dev = B.Device(g)
    assert dev.info()["mixed"] == 1
    dev.init()
    orc.init_previous_edge(h)
    dev.iterate(15)
    for _ in range(15):
        orc.loopy_iteration(h)
    dev.download(g)
    assert same_bits(B.states(g), B.states(h))
    assert same_bits(B.messages(g, "cur"), B.messages(h, "cur"))
    assert same_bits(B.messages(g, "prev"), B.messages(h, "prev"))
    dev.free()
    for x in (g, h):
        lib.graph_destroy(x)
    lib.bp_bn_free(bn)


def _star(lib, n_leaves: int, states: int):
    """A hub with n_leaves in-edges (more than one wave of slots)."""
    import ctypes as C

    rng = np.random.default_rng(n_leaves)
    g = lib.create_graph(n_leaves + 1, 2 * n_leaves)
    lib.graph_add_node(g, states, b"hub")
    st = (C.c_float * 10)()
    for i in range(n_leaves):
        lib.graph_add_node(g, states, f"leaf{i}".encode())
        if i % 3 == 0:
            p = rng.random(states).astype(np.float32)
            for k in range(states):
                st[k] = float(p[k])
            lib.graph_set_node_state(g, i + 1, states, st)
    tab = (C.c_float * 100)()
    for i in range(n_leaves):
        J = rng.random((states, states)).astype(np.float32)
        for a in range(states):
            for b in range(states):
                tab[a * 10 + b] = float(J[a, b])
        lib.graph_add_edge(g, i + 1, 0, states, states, tab)
        lib.graph_add_edge(g, 0, i + 1, states, states, tab)
    lib.set_up_src_nodes_to_edges(g)
    lib.set_up_dest_nodes_to_edges(g)
    return g


@pytest.mark.parametrize("n_leaves,states", [(63, 2), (64, 3), (65, 2), (200, 4), (1000, 2)])
def test_high_in_degree_matches_oracle(lib, orc, n_leaves, states):
    """Nodes with more in-edges than a wave has lanes (the reference caps
    lists at MAX_DEGREE; the drop-in does not)."""
    print(f"star {n_leaves} x {states}", flush=True)
    g = _star(lib, n_leaves, states)
    h = _star(lib, n_leaves, states)
    dev = B.Device(g)
    dev.init()
    orc.init_previous_edge(h)
    dev.iterate(6)
    for _ in range(6):
        orc.loopy_iteration(h)
    dev.download(g)
    assert same_bits(B.states(g), B.states(h)), bits_report(B.states(g), B.states(h))
    assert same_bits(B.messages(g, "prev"), B.messages(h, "prev"))
    dev.free()
    lib.graph_destroy(g)
    lib.graph_destroy(h)


def _pair_graph(lib, n_nodes: int, pairs, states, seed: int):
    """Nodes with states[v % len(states)] states and, for every (a, b) in
    pairs, an edge a -> b and its reverse b -> a carrying the transposed
    table, as the parsers emit them (a handle may then share one stored
    table per pair)."""
    import ctypes as C

    rng = np.random.default_rng(seed)
    g = lib.create_graph(n_nodes, 2 * len(pairs))
    ns = [states[v % len(states)] for v in range(n_nodes)]
    st = (C.c_float * 10)()
    for v in range(n_nodes):
        lib.graph_add_node(g, ns[v], f"n{v}".encode())
        if v % 3 == 0:
            p = rng.random(ns[v]).astype(np.float32)
            for k in range(ns[v]):
                st[k] = float(p[k])
            lib.graph_set_node_state(g, v, ns[v], st)
    tab = (C.c_float * 100)()
    for a_, b_ in pairs:
        J = rng.random((ns[a_], ns[b_])).astype(np.float32)
        for i in range(ns[a_]):
            for j in range(ns[b_]):
                tab[i * 10 + j] = float(J[i, j])
        lib.graph_add_edge(g, a_, b_, ns[a_], ns[b_], tab)
        for i in range(ns[b_]):
            for j in range(ns[a_]):
                tab[i * 10 + j] = float(J[j, i])
        lib.graph_add_edge(g, b_, a_, ns[b_], ns[a_], tab)
    lib.set_up_src_nodes_to_edges(g)
    lib.set_up_dest_nodes_to_edges(g)
    return g


def _lattice_with_hub(side: int, hub_leaves: int):
    """A side x side lattice (node 0 is an extra hub) whose hub is joined to
    the first hub_leaves lattice nodes: nearly every pair is local, and the
    hub has more in-slots than a chunk holds."""
    pairs = []
    for r in range(side):
        for c in range(side):
            v = 1 + r * side + c
            if c + 1 < side:
                pairs.append((v, v + 1))
            if r + 1 < side:
                pairs.append((v, v + side))
    # the hub's neighbours are consecutive lattice nodes: a hub joined to
    # nodes far apart would be a shortcut that scatters the device order
    pairs += [(1 + k, 0) for k in range(hub_leaves)]
    return 1 + side * side, pairs


@pytest.mark.parametrize("shape,states,shared,mixed", [
    (("lattice", 128, 150), (2,), 1, 0), (("lattice", 128, 150), (4,), 1, 0),
    (("star", 300), (4, 2, 3), 0, 1), (("star", 90), (3, 2, 3), 0, 1)])
def test_high_in_degree_pairs_match_oracle(lib, orc, shape, states, shared, mixed):
    """A node above the chunk cap on a handle that shares one stored table
    per transposed pair (the shared-pool kernels: the later slot of each
    pair reads its partner's table transposed, the hub through the one-node
    path), and with mixed state counts (tables of padded rows, read whole by
    the narrow path and row by row by the one-node path)."""
    print(f"pairs {shape} states {states}", flush=True)
    if shape[0] == "lattice":
        n, pairs = _lattice_with_hub(shape[1], shape[2])
    else:
        n, pairs = shape[1] + 1, [(k + 1, 0) for k in range(shape[1])]
    g = _pair_graph(lib, n, pairs, states, seed=n)
    h = _pair_graph(lib, n, pairs, states, seed=n)
    dev = B.Device(g)
    info = dev.info()
    assert info["shared_tables"] == shared and info["mixed"] == mixed, info
    dev.init()
    orc.init_previous_edge(h)
    dev.iterate(6)
    for _ in range(6):
        orc.loopy_iteration(h)
    dev.download(g)
    assert same_bits(B.states(g), B.states(h)), bits_report(B.states(g), B.states(h))
    assert same_bits(B.messages(g, "prev"), B.messages(h, "prev"))
    assert same_bits(B.messages(g, "cur"), B.messages(h, "cur"))
    # the host's arrays back on the device: the combine buffer is rebuilt
    # from them (PREP) before two more iterations
    B.check(lib.bp_upload_state(dev.h, g), "bp_upload_state")
    dev.iterate(2)
    for _ in range(2):
        orc.loopy_iteration(h)
    dev.download(g)
    assert same_bits(B.states(g), B.states(h)), bits_report(B.states(g), B.states(h))
    assert same_bits(B.messages(g, "prev"), B.messages(h, "prev"))
    dev.free()
    # the edge-centric schedule on the same handle kind
    g2 = _pair_graph(lib, n, pairs, states, seed=n)
    h2 = _pair_graph(lib, n, pairs, states, seed=n)
    dev = B.Device(g2)
    dev.init()
    B.check(lib.bp_edge_iterate(dev.h, 4), "bp_edge_iterate")
    orc.init_previous_edge(h2)
    for _ in range(4):
        orc.edge_iteration(h2)
    dev.download(g2)
    assert same_bits(B.states(g2), B.states(h2)), bits_report(B.states(g2), B.states(h2))
    for which in ("cur", "prev"):
        assert same_bits(B.messages(g2, which), B.messages(h2, which)), which
    dev.free()
    for x in (g, h, g2, h2):
        lib.graph_destroy(x)


def test_send_message_dropin(lib, orc):
    """send_message (graph.c:426-448), one call per edge over every edge of
    the lattice, run on the GPU, each equal to the oracle's send."""
    import ctypes as C

    d = golden("lattice16")
    g = graph_for_case(d)
    G = g.contents
    E = G.current_num_edges
    h = graph_for_case(d)
    for e in range(E):
        src = G.edges_src_index[e]
        lib.send_message(G.node_states, 10 * src, e, G.edges_joint_probabilities, G.edges_messages,
                         G.edges_x_dim, G.edges_y_dim)
        b = (C.c_float * 10)(*B.view_f32(G.node_states, G.current_num_vertices * 10)[10 * src:10 * src + 10])
        out = (C.c_float * 10)()
        orc.load().orc_send_message(b, G.node_num_vars[src], C.cast(h, C.c_void_p), e, out)
        got = B.view_f32(G.edges_messages, E * 10)[10 * e:10 * e + 4]
        assert same_bits(got, np.array(out[:4], dtype=np.float32))
    lib.graph_destroy(g)
    lib.graph_destroy(h)


def test_upload_errors_are_reported(lib):
    """Status codes, not aborts, at the C-ABI."""
    import ctypes as C

    g = lib.create_graph(3, 2)
    lib.graph_add_node(g, 2, b"a")
    lib.graph_add_node(g, 2, b"b")
    h = C.c_void_p()
    assert lib.bp_upload(g, -1, C.byref(h)) == -1  # 2 of 3 nodes added
    assert b"nodes" in lib.bp_last_error()
    lib.graph_destroy(g)


@pytest.mark.parametrize("case", ["example", "dog_xml", "c1", "lattice16", "c4net_small"])
def test_edge_schedule_bitwise(lib, orc, case):
    """Edge-centric iterations on the GPU == the oracle's, bit for bit,
    through the handle and through loopy_propagate_edge_one_iteration."""
    print(f"case {case}", flush=True)
    d = golden(case)
    g = graph_for_case(d)
    h = graph_for_case(d)
    k = graph_for_case(d)
    orc.init_previous_edge(h)
    for _ in range(5):
        orc.edge_iteration(h)
    dev = B.Device(g)
    dev.init()
    B.check(lib.bp_edge_iterate(dev.h, 5), "bp_edge_iterate")
    dev.download(g)
    for which in ("cur", "prev"):
        assert same_bits(B.messages(g, which), B.messages(h, which)), which
    assert same_bits(B.states(g), B.states(h)), bits_report(B.states(g), B.states(h))
    dev.free()
    lib.init_previous_edge(k)
    for _ in range(5):
        lib.loopy_propagate_edge_one_iteration(k)
    assert same_bits(B.states(k), B.states(h))
    assert same_bits(B.messages(k, "cur"), B.messages(h, "cur"))
    for x in (g, h, k):
        lib.graph_destroy(x)


def _expected_until(orc, g, step, conv, max_iter, batched):
    """Return value of the until drivers, from FP64 deltas of the oracle."""
    prev = -1.0
    i = 0
    num_iter = 0
    while i < max_iter:
        n = 10 if batched else 1
        for _ in range(n):
            step(g)
        delta = orc.delta64(g)
        if delta < conv or abs(delta - prev) < conv:
            return num_iter if batched else i
        prev = delta
        num_iter += n
        i += n
    return num_iter if batched else max_iter


@pytest.mark.parametrize("variant", ["until_edge", "acc", "edge_acc", "hip", "hip_edge"])
def test_convergence_drivers(lib, orc, variant):
    """loopy_propagate_until_edge, loopy_progagate_until_acc / _edge_acc and
    the HIP entry points: return value and results against the oracle run
    for the same number of iterations."""
    d = golden("c1")
    conv, mx = 1e-3, 200
    edge = variant in ("until_edge", "edge_acc", "hip_edge")
    batched = variant in ("acc", "edge_acc")
    h = graph_for_case(d)
    orc.init_previous_edge(h)
    step = orc.edge_iteration if edge else orc.loopy_iteration
    want = _expected_until(orc, h, step, conv, mx, batched)
    g = graph_for_case(d)
    lib.init_previous_edge(g)
    fn = {"until_edge": lib.loopy_propagate_until_edge, "acc": lib.loopy_progagate_until_acc,
          "edge_acc": lib.loopy_progagate_until_edge_acc, "hip": lib.loopy_propagate_until_hip,
          "hip_edge": lib.loopy_propagate_until_hip_edge}[variant]
    got = fn(g, conv, mx)
    assert got == want
    # the oracle ran exactly as many iterations as the driver did
    assert same_bits(B.states(g), B.states(h))
    assert same_bits(B.messages(g, "cur"), B.messages(h, "cur"))
    assert same_bits(B.messages(g, "prev"), B.messages(h, "prev"))
    lib.graph_destroy(g)
    lib.graph_destroy(h)


@pytest.mark.parametrize("case", ["c1", "lattice16", "bf_1000_2000_1", "c2b_60k", "c2_60k"])
def test_device_node_order_does_not_change_results(lib, case):
    """Handles renumber nodes on the device (Cuthill-McKee, then grouped by
    the XCDs that gather them, by default); Cuthill-McKee alone and the
    reference's order give the same bits, and deltas that differ only in
    the FP64 summation order."""
    out = []
    for mode in (0, 1, 2):
        prev = lib.bp_set_node_order(mode)
        if case.endswith("_60k"):
            bn = B.generate(case.split("_")[0], num_nodes=60000)
            g = B.bn_to_graph(bn)
            lib.bp_bn_free(bn)
        else:
            g = graph_for_case(golden(case))
        dev = B.Device(g)
        lib.bp_set_node_order(prev)
        dev.init()
        deltas = dev.iterate(15, want_deltas=True)
        dev.download(g)
        out.append((B.states(g).copy(), B.messages(g, "prev").copy(), deltas))
        dev.free()
        lib.graph_destroy(g)
    for k in (1, 2):
        assert same_bits(out[0][0], out[k][0]) and same_bits(out[0][1], out[k][1]), f"mode {k}"
        assert np.allclose(out[0][2], out[k][2], rtol=1e-12, atol=0), f"mode {k}"


@pytest.mark.parametrize("edges", [0, 6])
def test_isolated_nodes_and_chunks_without_slots(lib, orc, edges):
    """Chunks whose nodes have no in-edges (every chunk, when the graph has
    no edges) take no slot loads; results equal the oracle's."""
    V = 300
    import ctypes as C

    def build():
        g = lib.create_graph(V, max(edges, 1))
        st = (C.c_float * 10)()
        for v in range(V):
            lib.graph_add_node(g, 2, f"n{v}".encode())
            if v % 7 == 0:
                st[0], st[1] = 0.25, 0.75
                lib.graph_set_node_state(g, v, 2, st)
        tab = (C.c_float * 100)()
        for i in range(edges):
            tab[0], tab[1], tab[10], tab[11] = 0.3 + 0.1 * i, 0.7, 0.6, 0.4 - 0.05 * i
            lib.graph_add_edge(g, i, i + 1, 2, 2, tab)
        lib.set_up_src_nodes_to_edges(g)
        lib.set_up_dest_nodes_to_edges(g)
        return g

    g, h = build(), build()
    dev = B.Device(g)
    dev.init()
    dev.iterate(5)
    dev.download(g)
    orc.init_previous_edge(h)
    for _ in range(5):
        orc.loopy_iteration(h)
    assert same_bits(B.states(g), B.states(h)), bits_report(B.states(g), B.states(h))
    dev.free()
    lib.graph_destroy(g)
    lib.graph_destroy(h)


def test_hip_entry_point_rereads_the_arrays_on_a_second_call(lib, orc):
    """loopy_propagate_until_hip keeps the graph's device mirror between
    calls; a table the caller changes between calls is used by the next
    call, as the oracle doing the same sequence shows."""
    d = golden("c1")
    conv, mx = 1e-3, 200
    g = graph_for_case(d)
    h = graph_for_case(d)
    e = 7
    for rnd in range(2):
        if rnd == 1:
            for x in (g, h):
                J = B.view_f32(x.contents.edges_joint_probabilities, x.contents.current_num_edges * 100)
                J[e * 100 + 0] *= 0.25
                J[e * 100 + 11] += 0.5
        orc.init_previous_edge(h)
        want = _expected_until(orc, h, orc.loopy_iteration, conv, mx, False)
        lib.init_previous_edge(g)
        got = lib.loopy_propagate_until_hip(g, conv, mx)
        assert got == want, f"round {rnd}"
        assert same_bits(B.states(g), B.states(h)), f"round {rnd}"
        assert same_bits(B.messages(g, "cur"), B.messages(h, "cur")), f"round {rnd}"
        assert same_bits(B.messages(g, "prev"), B.messages(h, "prev")), f"round {rnd}"
    lib.graph_destroy(g)
    lib.graph_destroy(h)
