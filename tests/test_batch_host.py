"""Host side of the batched form (include/bp_batch.h), no GPU needed:
instance bookkeeping, the shared table pool, and error paths."""
from __future__ import annotations

import numpy as np

import bp_ctypes as B
from workloads import batched_networks, network_sizes


def test_sizes_are_log_uniform_and_seeded():
    a = network_sizes(512, 200, 50000, 5005)
    assert np.array_equal(a, network_sizes(512, 200, 50000, 5005))
    assert a.min() >= 200 and a.max() < 50000
    # log-uniform: the median sits near the geometric mean of the bounds
    assert 0.5 < np.median(a) / np.sqrt(200 * 50000) < 2.0


def test_tables_are_shared_across_edges_and_evidence_sets(lib):
    b = B.Batch()
    arcs_xy = 0
    edges = 0
    for i, e, g in batched_networks(count=3, evidence_sets=4, lo=300, hi=900):
        G = g.contents
        if e == 0:
            # every arc of the network appears as a forward edge whatever the evidence
            src = B.view_u32(G.edges_src_index, G.current_num_edges)
            dst = B.view_u32(G.edges_dest_index, G.current_num_edges)
            x = B.view_u32(G.edges_x_dim, G.current_num_edges)
            y = B.view_u32(G.edges_y_dim, G.current_num_edges)
            fwd = src < dst  # parents have lower indices in these DAGs
            arcs_xy += int((x[fwd].astype(np.int64) * y[fwd]).sum())
        edges += G.current_num_edges
        b.add(g)
        lib.graph_destroy(g)
    info = b.info()
    assert info["instances"] == 12 and info["E"] == edges
    # one table per arc (reverse edges and other evidence sets reuse it)
    assert info["table_floats"] == arcs_xy
    b.free()


def test_upload_needs_a_gpu_here(lib):
    if lib.bp_device_count() > 0:
        return
    b = B.Batch()
    for i, e, g in batched_networks(count=1, evidence_sets=1, lo=200, hi=300):
        b.add(g)
        lib.graph_destroy(g)
    assert lib.bp_batch_upload(b.h, 0) == -1
    assert "no HIP device" in B.last_error()
    b.free()


def test_add_rejects_graph_without_csr(lib):
    g = lib.create_graph(3, 4)
    for k in range(3):
        lib.graph_add_node(g, 2, f"n{k}".encode())
    tab = (B.C.c_float * 100)(*([0.5] * 100))
    lib.graph_add_edge(g, 0, 1, 2, 2, tab)
    lib.graph_add_edge(g, 1, 2, 2, 2, tab)
    b = B.Batch()
    # dest CSR not set up: the offsets are zero while there are edges
    assert lib.bp_batch_add(b.h, g, None) == -1
    assert "CSR" in B.last_error()
    lib.set_up_src_nodes_to_edges(g)
    lib.set_up_dest_nodes_to_edges(g)
    assert lib.bp_batch_add(b.h, g, None) == 0
    b.free()
    lib.graph_destroy(g)
