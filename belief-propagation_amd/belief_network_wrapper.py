"""Python front end of the MI355X belief-propagation engine.

The reference's belief_network_wrapper.py only drives its network generator;
here the same name holds the Python interface to libbp_amd.so:

- ``BeliefNetwork``: load (XMLBIF / BIF) or generate a network and run loopy
  or tree BP on the GPU, mirroring the reference's C entry points
  (src/graph/graph.h:110-126).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

import bp_ctypes as B


class BeliefNetwork:
    """A Graph_t owned from Python, with its CSR views set up."""

    def __init__(self, graph: B.GraphP):
        if not graph:
            raise RuntimeError(f"graph construction failed: {B.last_error()}")
        self.lib = B.load()
        self.g = graph
        self._dev = None

    # -- construction ---------------------------------------------------
    @classmethod
    def from_xmlbif(cls, path: str | Path) -> "BeliefNetwork":
        lib = B.load()
        g = lib.parse_xml_file(str(path).encode())
        net = cls(g)
        net._setup()
        return net

    @classmethod
    def from_bif(cls, path: str | Path) -> "BeliefNetwork":
        lib = B.load()
        g = lib.parse_bif_file(str(path).encode())
        net = cls(g)
        net._setup()
        return net

    @classmethod
    def generate(cls, preset: str, **overrides) -> "BeliefNetwork":
        bn = B.generate(preset, **overrides)
        try:
            g = B.bn_to_graph(bn)
        finally:
            B.load().bp_bn_free(bn)
        return cls(g)

    def _setup(self):
        self.lib.set_up_src_nodes_to_edges(self.g)
        self.lib.set_up_dest_nodes_to_edges(self.g)

    @property
    def num_nodes(self) -> int:
        return self.g.contents.current_num_vertices

    @property
    def num_edges(self) -> int:
        return self.g.contents.current_num_edges

    # -- propagation (GPU) ------------------------------------------------
    def device(self) -> B.Device:
        if self._dev is None:
            self._dev = B.Device(self.g)
        return self._dev

    def loopy(self, iterations: int, with_deltas: bool = False):
        """init_previous_edge + `iterations` loopy iterations on the GPU."""
        d = self.device()
        d.init()
        deltas = d.iterate(iterations, want_deltas=with_deltas)
        d.download(self.g)
        return deltas

    def loopy_until(self, convergence: float = 1e-3, max_iterations: int = 200) -> int:
        d = self.device()
        d.init()
        it = d.until(convergence, max_iterations)
        d.download(self.g)
        return it

    def tree(self):
        if self.g.contents.num_levels == 0:
            self.lib.init_levels_to_nodes(self.g)
            if self._dev is not None:
                self._dev.free()
                self._dev = None
        d = self.device()
        d.tree()
        d.download(self.g)

    def beliefs(self) -> np.ndarray:
        """[V][MAX_STATES] node beliefs (entries beyond a node's state count are 0)."""
        return B.states(self.g)

    def close(self):
        if self._dev is not None:
            self._dev.free()
            self._dev = None
        if self.g:
            self.lib.graph_destroy(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["BeliefNetwork", "C"]
