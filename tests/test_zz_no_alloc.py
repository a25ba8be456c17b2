"""bp_loopy_iterate_noalloc (include/bp_hip.h): the allocation-free variant
of bp_loopy_iterate gives the same beliefs, messages and FP64 deltas, bit
for bit, on two calls in a row -- the second on a smaller graph, so the
shared fixed scratch array is reused with fewer partials -- and refuses a
request beyond its fixed scratch with an error."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import bp_ctypes as B
from conftest import same_bits

pytestmark = pytest.mark.gpu


def _run(lib, g, n, noalloc):
    dev = B.Device(g)
    dev.init()
    d = (C.c_double * n)()
    fn = lib.bp_loopy_iterate_noalloc if noalloc else lib.bp_loopy_iterate
    rc = fn(dev.h, n, d)
    assert rc == 0, B.last_error()
    dev.download(g)
    dev.free()
    return np.array(d[:n]), B.states(g).copy(), B.messages(g, "prev").copy(), B.messages(g, "cur").copy()


@pytest.mark.parametrize("order", ["larger_first"])
def test_noalloc_matches_iterate_on_two_calls(lib, order):
    cases = [("c1", dict(num_nodes=60000, states_hi=4, seed=77), 20),  # mixed 2-4 states, ~60k nodes
             ("c1", dict(), 12)]                                       # 1000 binary nodes
    for preset, kw, n in cases:
        print(f"noalloc vs iterate: {preset} {kw} n={n}", flush=True)
        bn = B.generate(preset, **kw)
        g, h = B.bn_to_graph(bn), B.bn_to_graph(bn)
        lib.bp_bn_free(bn)
        a = _run(lib, g, n, noalloc=True)
        b = _run(lib, h, n, noalloc=False)
        assert np.array_equal(a[0], b[0]), f"deltas {a[0]} vs {b[0]}"
        for x, y in zip(a[1:], b[1:]):
            assert same_bits(x, y)
        lib.graph_destroy(g)
        lib.graph_destroy(h)


def test_noalloc_refuses_beyond_its_scratch(lib):
    bn = B.generate("c1")
    g = B.bn_to_graph(bn)
    lib.bp_bn_free(bn)
    dev = B.Device(g)
    dev.init()
    rc = lib.bp_loopy_iterate_noalloc(dev.h, 1 << 20, None)
    assert rc == -1
    assert "scratch" in B.last_error()
    dev.free()
    lib.graph_destroy(g)
