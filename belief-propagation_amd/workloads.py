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

The graphs are built by a small pool of threads (the library's graph API
is reentrant and ctypes releases the GIL during each call), a bounded
number of networks ahead of the consumer, and yielded in order.
"""
from __future__ import annotations

import ctypes as C
import os
from collections import deque
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import bp_ctypes as B


def network_sizes(count: int, lo: int, hi: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=count)).astype(np.int64)


def _network_graphs(i: int, n: int, evidence_sets: int, fraction: float, base: int) -> list:
    lib = B.load()
    bn = B.generate("c4net", num_nodes=int(n), evidence_fraction=0.0, seed=base + i)
    ndefs = C.cast(bn, C.POINTER(B.BN)).contents.num_defs
    graphs = []
    try:
        for e in range(evidence_sets):
            B.check(lib.bp_bn_add_evidence(bn, fraction, (base + i) * 16 + e + 1), "bp_bn_add_evidence")
            graphs.append(B.bn_to_graph(bn))
            lib.bp_bn_truncate_defs(bn, ndefs)
    finally:
        lib.bp_bn_free(bn)
    return graphs


def batched_networks(count: int = 512, evidence_sets: int = 8, lo: int = 200, hi: int = 50000,
                     fraction: float = 0.03, base: int = 5005, threads: int | None = None, ahead: int = 8):
    """Yields (network index, evidence index, Graph_t) in order; the caller
    destroys each graph."""
    sizes = network_sizes(count, lo, hi, base)
    if threads is None:
        try:
            threads = max(1, min(16, len(os.sched_getaffinity(0))))
        except AttributeError:
            threads = 4
    with ThreadPoolExecutor(max_workers=threads) as pool:
        pending = deque()
        nxt = 0
        while nxt < count and len(pending) < ahead:
            pending.append(pool.submit(_network_graphs, nxt, sizes[nxt], evidence_sets, fraction, base))
            nxt += 1
        for i in range(count):
            graphs = pending.popleft().result()
            if nxt < count:
                pending.append(pool.submit(_network_graphs, nxt, sizes[nxt], evidence_sets, fraction, base))
                nxt += 1
            for e, g in enumerate(graphs):
                yield i, e, g
