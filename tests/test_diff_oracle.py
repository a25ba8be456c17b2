"""Pins the differentiable form's oracle (oracle/bp_diff_oracle.py) on the CPU.

- Its parameter map, applied to a network's CPT tables in float64, gives the
  pairwise tables the library's XMLBIF reader built (bitwise equal to the
  reference's) to within the f32 rounding of their sums; a transposed map
  fails that.
- Its forward recursion equals the C FP64 oracle (bitwise equal to the
  reference built with -Dfloat=double) on the same tables.
- Its gradient equals central finite differences of its float64 loss.
Also: the library's own structure checks of the adjoint, which need no GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

import bp_ctypes as B
import bp_diff_oracle as D


def network(preset: str, **kw):
    bn = B.generate(preset, **kw)
    A = B.bn_arrays(bn)
    g = B.bn_to_graph(bn)
    return bn, A, g, B.graph_arrays(g)


CASES = [("c5", dict(num_nodes=1500)), ("c1", dict())]


@pytest.mark.parametrize("preset,kw", CASES)
def test_map_matches_reader_tables(lib, preset, kw):
    bn, A, g, G = network(preset, **kw)
    st = D.DiffStructure(A, G, len(A["num_states"]))
    P = torch.tensor(A["table"][: st.num_params].astype(np.float64))
    J = st.tables(P).numpy()
    E, S = st.E, st.S
    J32 = G["edges_J"].reshape(E, 10, 10)[:, :S, :S].astype(np.float64)
    # every table entry is a sum of at most 3^6 positive terms
    bound = 730 * 2.0 ** -24 * np.abs(J) + 1e-30
    assert np.all(np.abs(J - J32) <= bound)
    # a transposed map is caught
    bad = torch.zeros(E * S * S, dtype=torch.float64)
    e, rem = st.map_tgt // (S * S), st.map_tgt % (S * S)
    bad = bad.index_add(0, e * S * S + (rem % S) * S + rem // S, P[st.map_src]).view(E, S, S).numpy()
    assert not np.all(np.abs(bad - J32) <= 730 * 2.0 ** -24 * np.abs(bad) + 1e-30)
    # initial beliefs: priors and evidence as the reader set them
    s0 = st.initial(P).numpy()
    assert np.array_equal(s0, G["node_states"].reshape(-1, 10)[:, :S].astype(np.float64))
    lib.graph_destroy(g)
    lib.bp_bn_free(bn)


@pytest.mark.parametrize("preset,kw", CASES)
def test_forward_matches_c_fp64_oracle(lib, orc, preset, kw):
    bn, A, g, G = network(preset, **kw)
    st = D.DiffStructure(A, G, len(A["num_states"]))
    E, S = st.E, st.S
    J = torch.tensor(G["edges_J"].reshape(E, 10, 10)[:, :S, :S].astype(np.float64))
    K = 30
    with torch.no_grad():
        b = D.unrolled(st, J, st.base_st0, K).numpy()
    o = orc.FP64(g)
    o.init()
    o.iterate(K)
    assert np.abs(b - o.states()[:, :S]).max() < 1e-12
    del o
    lib.graph_destroy(g)
    lib.bp_bn_free(bn)


def test_gradient_matches_finite_differences(lib):
    bn, A, g, G = network("c5", num_nodes=800)
    st = D.DiffStructure(A, G, len(A["num_states"]))
    P = A["table"][: st.num_params].astype(np.float64)
    rng = np.random.default_rng(5)
    cand = np.nonzero(G["observed"] == 0)[0]
    q = rng.choice(cand, 200, replace=False)
    t = rng.integers(0, st.S, 200)
    K = 12
    _, grad = D.loss_and_grad(st, P, K, q, t)
    for _ in range(3):
        v = rng.standard_normal(len(P))
        eps = 1e-6
        fd = (D.loss_only(st, P + eps * v, K, q, t) - D.loss_only(st, P - eps * v, K, q, t)) / (2 * eps)
        assert abs(fd - grad @ v) <= 1e-5 * max(1.0, abs(fd))
    lib.graph_destroy(g)
    lib.bp_bn_free(bn)


def diff_create_error(lib, bn, npd):
    h = C.c_void_p()
    rc = lib.bp_diff_create(bn, npd, 0, C.byref(h))
    assert rc == -1
    return B.last_error()


def test_adjoint_rejects_mixed_state_counts(lib):
    bn = B.generate("c1", num_nodes=300, states_hi=4, seed=77)
    assert "one state count" in diff_create_error(lib, bn, 300)
    lib.bp_bn_free(bn)


def test_adjoint_rejects_givens_after_the_parameters(lib):
    bn = B.generate("c1", num_nodes=300)
    # declaring only the first 10 definitions as parameters leaves CPTs with
    # GIVENs among the constants
    assert "after the parameters has GIVENs" in diff_create_error(lib, bn, 10)
    lib.bp_bn_free(bn)


def test_adjoint_rejects_wide_nodes(lib):
    # node 1 is the single parent of 199 children and is not observed, so
    # each child sends it a reverse edge: 200 in-edges at node 1
    bn = lib.bp_bn_create(201)
    for v in range(201):
        assert lib.bp_bn_set_var(bn, v, f"n{v}".encode(), 2) == 0
    assert lib.bp_bn_add_def(bn, 0, None, 0, (C.c_float * 2)(0.3, 0.7), 2) == 0
    tab = (C.c_float * 4)(0.2, 0.6, 0.8, 0.4)
    assert lib.bp_bn_add_def(bn, 1, (C.c_uint * 1)(0), 1, tab, 4) == 0
    for v in range(2, 201):
        assert lib.bp_bn_add_def(bn, v, (C.c_uint * 1)(1), 1, tab, 4) == 0
    assert "in-edges" in diff_create_error(lib, bn, 201)
    lib.bp_bn_free(bn)


def test_adjoint_needs_a_gpu_here(lib):
    if lib.bp_device_count() > 0:
        pytest.skip("a GPU is visible")
    bn = B.generate("c5", num_nodes=300)
    assert "no HIP device" in diff_create_error(lib, bn, 300)
    lib.bp_bn_free(bn)
