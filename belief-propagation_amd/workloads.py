"""Seeded workloads for the batched form.

``batched_networks`` yields the networks and evidence sets of a batch: DAG
networks of the "c4net" family (in-degree uniform 0-4, states uniform 2-8,
CPT rows U(0,1)), node counts log-uniform in [lo, hi], each with several
one-hot evidence sets of the given fraction.  Network i uses seed
base + i; its evidence set e uses seed (base + i) * 16 + e + 1.  Each
(network, evidence set) pair is one instance: the XMLBIF reader's graph of
that network with those observations (roots keep their priors), so two
evidence sets of one network differ in beliefs and in which reverse edges
exist (xml-expression.c:409: none into an observed parent).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import bp_ctypes as B


def network_sizes(count: int, lo: int, hi: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=count)).astype(np.int64)


def batched_networks(count: int = 512, evidence_sets: int = 8, lo: int = 200, hi: int = 50000,
                     fraction: float = 0.03, base: int = 5005):
    """Yields (network index, evidence index, Graph_t); the caller destroys each graph."""
    lib = B.load()
    sizes = network_sizes(count, lo, hi, base)
    for i, n in enumerate(sizes):
        bn = B.generate("c4net", num_nodes=int(n), evidence_fraction=0.0, seed=base + i)
        V = C.cast(bn, C.POINTER(B.BN)).contents.num_defs
        try:
            for e in range(evidence_sets):
                B.check(lib.bp_bn_add_evidence(bn, fraction, (base + i) * 16 + e + 1), "bp_bn_add_evidence")
                g = B.bn_to_graph(bn)
                lib.bp_bn_truncate_defs(bn, V)
                yield i, e, g
        finally:
            lib.bp_bn_free(bn)
