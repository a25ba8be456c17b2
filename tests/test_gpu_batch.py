"""The batched form on the GPU (include/bp_batch.h): every instance's
beliefs, messages and deltas equal running it alone through the plain path
(bp_upload + bp_loopy_init + bp_loopy_iterate), for every kernel shape the
tuner can pick; and an instance equals the oracle."""
from __future__ import annotations

import ctypes as C
import json

import numpy as np
import pytest

import bp_ctypes as B
from conftest import same_bits, bits_report
from workloads import batched_networks

pytestmark = pytest.mark.gpu


def build(count, evidence_sets, lo, hi, preset_graphs=None):
    b = B.Batch()
    graphs = []
    for i, e, g in batched_networks(count=count, evidence_sets=evidence_sets, lo=lo, hi=hi):
        b.add(g)
        graphs.append(g)
    return b, graphs


def plain_run(g, n):
    d = B.Device(g, 0)
    d.init()
    deltas = d.iterate(n, want_deltas=True)
    d.download(g)
    d.free()
    return deltas


def copy_graph_results(lib, g):
    return B.states(g).copy(), B.messages(g, "prev").copy(), B.messages(g, "cur").copy()


def test_batch_matches_plain_per_instance(lib):
    n = 30
    b, graphs = build(10, 3, 200, 3000)
    print(f"batch: {b.info()}", flush=True)
    b.upload(0)
    b.init()
    bd = b.iterate(n, want_deltas=True)
    for i, g in enumerate(graphs):
        b.download(i, g)
    batch_res = [copy_graph_results(lib, g) for g in graphs]
    for g in graphs:
        lib.graph_destroy(g)
    # the same instances, one at a time
    for i, e, g in batched_networks(count=10, evidence_sets=3, lo=200, hi=3000):
        k = i * 3 + e
        pd = plain_run(g, n)
        st, mp, mc = copy_graph_results(lib, g)
        assert same_bits(batch_res[k][0], st), f"instance {k}: {bits_report(batch_res[k][0], st)}"
        assert same_bits(batch_res[k][1], mp) and same_bits(batch_res[k][2], mc), f"instance {k} messages"
        assert np.allclose(bd[:, k], pd, rtol=1e-12, atol=1e-300), f"instance {k} deltas"
        lib.graph_destroy(g)
    b.free()


@pytest.mark.parametrize("preset_hi", [4, 5, 7, 8])
def test_every_shape_gives_the_same_bits(lib, preset_hi):
    """Every launch shape gives the bits of the first, which are those of
    the plain run; with 5 to 7 states the shape with two lanes per slot
    runs the 8-state kernel."""
    print(f"states 2-{preset_hi}", flush=True)
    b = B.Batch()
    gs = []
    for s in range(6):
        bn = B.generate("c4net", num_nodes=400 + 300 * s, states_hi=preset_hi, seed=900 + s)
        g = B.bn_to_graph(bn)
        lib.bp_bn_free(bn)
        b.add(g)
        gs.append(g)
    b.upload(0)
    ref = None
    shapes = [(1, 1, 0), (1, 1, 1), (1, 1, 4), (1, 1, 16), (1, 2, 0), (1, 4, 4), (1, 8, 0)]
    shapes += [(2, 1, 0), (2, 1, 4)] if preset_hi <= 6 else []
    for R, L, q in shapes:
        b.set_config(R, L, q)
        b.init()
        d = b.iterate(12, want_deltas=True)
        b.download(len(gs) - 1, gs[-1])
        got = (B.states(gs[-1]).copy(), d)
        if ref is None:
            ref = got
        else:
            assert same_bits(got[0], ref[0]), f"shape {R},{L},{q}"
            assert np.array_equal(got[1], ref[1]), f"shape {R},{L},{q} deltas"
    b.free()
    # the plain handle on the same network from its original beliefs
    bn = B.generate("c4net", num_nodes=400 + 300 * 5, states_hi=preset_hi, seed=900 + 5)
    g = B.bn_to_graph(bn)
    lib.bp_bn_free(bn)
    pd = plain_run(g, 12)
    assert same_bits(B.states(g), ref[0]), "plain run"
    assert np.allclose(ref[1][:, -1], pd, rtol=1e-12, atol=1e-300), "plain deltas"
    lib.graph_destroy(g)
    for g in gs:
        lib.graph_destroy(g)


def test_tuner_picks_and_caches_a_shape(lib, tmp_path):
    b, graphs = build(6, 2, 300, 2000)
    b.upload(0)
    cache = tmp_path / "tune.json"
    b.tune(3, str(cache))
    first = b.config()
    data = json.loads(cache.read_text())
    assert len(data) == 1
    (entry,) = data.values()
    assert (entry["slots_per_lane"], entry["lanes_per_slot"], entry["queue_block"]) == first
    b.set_config(1, 1, 16 if first != (1, 1, 16) else 0)
    b.tune(3, str(cache))  # a hit: no timing, same choice
    assert b.config() == first
    b.init()
    b.iterate(5)
    b.sync()
    b.free()
    for g in graphs:
        lib.graph_destroy(g)


def test_instance_matches_oracle(lib, orc):
    b, graphs = build(3, 2, 500, 1500)
    b.upload(0)
    b.init()
    b.iterate(20)
    k = 4
    b.download(k, graphs[k])
    got = B.states(graphs[k]).copy()
    for i, e, g in batched_networks(count=3, evidence_sets=2, lo=500, hi=1500):
        if i * 2 + e == k:
            orc.init_previous_edge(g)
            for _ in range(20):
                orc.loopy_iteration(g)
            assert same_bits(got, B.states(g)), bits_report(got, B.states(g))
        lib.graph_destroy(g)
    b.free()
    for g in graphs:
        lib.graph_destroy(g)


@pytest.mark.parametrize("shape", [(1, 1, 0), (1, 4, 4)])
def test_interleaved_evidence_groups_match_plain(lib, shape):
    """20 evidence sets per network: the batch runs each network's copies
    merged by node range, in groups of at most 16 (here 16 + 4); every
    instance still equals its plain run, deltas included."""
    n, sets = 12, 20
    b, graphs = build(3, sets, 200, 900)
    b.upload(0)
    b.set_config(*shape)
    b.init()
    bd = b.iterate(n, want_deltas=True)
    for i, g in enumerate(graphs):
        b.download(i, g)
    got = [copy_graph_results(lib, g) for g in graphs]
    for g in graphs:
        lib.graph_destroy(g)
    for i, e, g in batched_networks(count=3, evidence_sets=sets, lo=200, hi=900):
        k = i * sets + e
        pd = plain_run(g, n)
        st, mp, mc = copy_graph_results(lib, g)
        lib.graph_destroy(g)
        assert same_bits(got[k][0], st), f"instance {k}: {bits_report(got[k][0], st)}"
        assert same_bits(got[k][1], mp) and same_bits(got[k][2], mc), f"instance {k} messages"
        assert np.allclose(bd[:, k], pd, rtol=1e-12, atol=1e-300), f"instance {k} deltas"
    b.free()


def test_device_node_order_does_not_change_batch_results(lib):
    """Instances are renumbered on the device (Cuthill-McKee, once per node
    set); with the graphs' own order the downloads are the same bits."""
    n = 15
    out = {}
    for mode in (0, 1):
        prev = lib.bp_set_node_order(mode)
        try:
            b, graphs = build(4, 3, 200, 1500)
        finally:
            lib.bp_set_node_order(prev)
        b.upload(0)
        b.init()
        d = b.iterate(n, want_deltas=True)
        res = []
        for i, g in enumerate(graphs):
            b.download(i, g)
            res.append(copy_graph_results(lib, g))
            lib.graph_destroy(g)
        b.free()
        out[mode] = (res, d)
    for k, (a, c) in enumerate(zip(out[0][0], out[1][0])):
        assert all(same_bits(x, y) for x, y in zip(a, c)), f"instance {k}"
    assert np.allclose(out[0][1], out[1][1], rtol=1e-12, atol=1e-300)


class _ResultShell:
    """A Graph_t holding only what bp_batch_download writes: node and edge
    counts, node_states and the two message arrays (zeroed, so entries
    beyond each node's or edge's count stay 0, as B.states/B.messages mask
    them), with the message pointers as create_graph sets them."""

    def __init__(self, V: int, E: int):
        self.states = np.zeros(max(V, 1) * 10, dtype=np.float32)
        self.m_a = np.zeros(max(E, 1) * 10, dtype=np.float32)
        self.m_b = np.zeros(max(E, 1) * 10, dtype=np.float32)
        g = B.Graph()
        g.total_num_vertices = g.current_num_vertices = V
        g.total_num_edges = g.current_num_edges = E
        f32p = C.POINTER(C.c_float)
        g.node_states = self.states.ctypes.data_as(f32p)
        g.edges_messages = self.m_a.ctypes.data_as(f32p)
        g.last_edges_messages = self.m_b.ctypes.data_as(f32p)
        self.g = g
        self.ptr = C.pointer(g)
        base = C.addressof(g)
        g.current_edge_messages = C.cast(base + B.Graph.edges_messages.offset, C.POINTER(f32p))
        g.previous_edge_messages = C.cast(base + B.Graph.last_edges_messages.offset, C.POINTER(f32p))
        self.V, self.E = V, E

    def arrays(self):
        g = self.g
        prev = np.ctypeslib.as_array(g.previous_edge_messages.contents, shape=(self.E * 10,))
        cur = np.ctypeslib.as_array(g.current_edge_messages.contents, shape=(self.E * 10,))
        return (self.states[: self.V * 10].reshape(self.V, 10), prev.reshape(self.E, 10), cur.reshape(self.E, 10))


def _sha(a) -> str:
    import hashlib

    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.timeout(1200)
def test_full_size_batch_matches_plain_runs_and_oracle(lib, orc, request):
    """The batched workload at its stated size: 512 seeded DAGs (node counts
    log-uniform in 200-50k, in-degree 0-4, states 2-8) x 8 evidence sets
    (3% one-hot), init + 30 iterations in one batch, shape tuned.  Every
    one of the 4096 instances -- beliefs, both message arrays and its FP64
    deltas -- equals its plain run (bp_upload + bp_loopy_init +
    bp_loopy_iterate, the network alone), and the first instance of the
    largest network equals the oracle.  The plain runs are taken while the
    batch is built (one pass over the networks) and kept as digests."""
    import time

    from workloads import _network_graphs, network_sizes

    tr = request.config.pluginmanager.get_plugin("terminalreporter")

    def progress(msg):
        if tr is not None:
            tr.write_line(f"full-size batch: {msg}")
            tr._tw.flush()

    count, sets, n = 512, 8, 30
    sizes = network_sizes(count, 200, 50000, 5005)
    big = int(np.argmax(sizes))
    b = B.Batch()
    plain = []
    t0 = time.time()
    for i, e, g in batched_networks(count=count, evidence_sets=sets):
        b.add(g)
        pd = plain_run(g, n)
        st, mp, mc = copy_graph_results(lib, g)
        plain.append((g.contents.current_num_vertices, g.contents.current_num_edges, _sha(st), _sha(mp), _sha(mc), pd))
        lib.graph_destroy(g)
        if e == sets - 1 and (i + 1) % 32 == 0:
            progress(f"{i + 1} of {count} networks built and run alone ({time.time() - t0:.0f} s)")
    b.upload(0)
    b.tune(3)
    progress(f"uploaded, shape {b.config()}")
    b.init()
    bd = b.iterate(n, want_deltas=True)
    b.sync()
    info = b.info()
    assert info["instances"] == count * sets
    for k, (V, E, s_st, s_mp, s_mc, pd) in enumerate(plain):
        sh = _ResultShell(V, E)
        b.download(k, sh.ptr)
        st, mp, mc = sh.arrays()
        assert _sha(st) == s_st, f"instance {k}: beliefs differ from its plain run"
        assert _sha(mp) == s_mp and _sha(mc) == s_mc, f"instance {k}: messages differ from its plain run"
        assert np.allclose(bd[:, k], pd, rtol=1e-12, atol=1e-300), f"instance {k}: deltas"
        if k == big * sets:
            big_states = st.copy()
        if (k + 1) % 1024 == 0:
            progress(f"{k + 1} of {count * sets} instances equal their plain runs")
    (h,) = _network_graphs(big, sizes[big], 1, 0.03, 5005)
    orc.init_previous_edge(h)
    for _ in range(n):
        orc.loopy_iteration(h, threads=8)
    want = B.states(h)
    assert same_bits(big_states, want), bits_report(big_states, want)
    progress(f"largest network ({h.contents.current_num_vertices} nodes) equals the oracle")
    lib.graph_destroy(h)
    b.free()
