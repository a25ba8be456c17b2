"""Python front end of the MI355X belief-propagation engine.

The reference's belief_network_wrapper.py only drives its network generator;
here the same name holds the Python interface to libbp_amd.so:

- ``BeliefNetwork``: load (XMLBIF / BIF) or generate a network and run loopy
  or tree BP on the GPU, mirroring the reference's C entry points
  (src/graph/graph.h:110-126).
- ``DifferentiableNetwork`` and ``UnrolledLoopyBP``: the beliefs after K
  unrolled loopy iterations (init_previous_edge, src/graph/graph.c:725, then
  loopy_propagate_one_iteration, graph.c:1121) as a differentiable function
  of the network's conditional probability tables, with the hand-written
  adjoint of libbp_amd.so (include/bp_diff.h) as its backward.
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


def _stream_ptr():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class DifferentiableNetwork:
    """A network whose CPTs are parameters, on one GPU (bp_diff handle).

    The parameter vector is the concatenation of the first ``num_param_defs``
    definitions' TABLEs (child-state-major, as in the XMLBIF file); later
    definitions are observations (constants).  ``softmax_tables`` maps
    log-parameters (one row per definition and parent configuration) to it.
    """

    def __init__(self, bn, num_param_defs: int | None = None, device: int = 0, own_bn: bool = False):
        import torch

        self.lib = B.load()
        self.bn = bn
        self._own_bn = own_bn
        if num_param_defs is None:
            num_param_defs = C.cast(bn, C.POINTER(B.BN)).contents.num_vars
        self.num_param_defs = num_param_defs
        h = C.c_void_p()
        with torch.cuda.device(device):
            B.check(self.lib.bp_diff_create(bn, num_param_defs, device, C.byref(h)), "bp_diff_create")
        self.h = h
        self.device = torch.device("cuda", device)
        info = (C.c_uint64 * 6)()
        B.check(self.lib.bp_diff_info(self.h, info), "bp_diff_info")
        self.V, self.E, self.S, self.stride, self.num_params, self.num_rows = (int(x) for x in info)
        first = np.zeros(self.num_rows, dtype=np.uint64)
        stride = np.zeros(self.num_rows, dtype=np.uint32)
        length = np.zeros(self.num_rows, dtype=np.uint32)
        B.check(self.lib.bp_diff_rows(self.h, first.ctypes.data_as(C.POINTER(C.c_uint64)),
                                      stride.ctypes.data_as(C.POINTER(C.c_uint)),
                                      length.ctypes.data_as(C.POINTER(C.c_uint))), "bp_diff_rows")
        if self.num_rows and not np.all(length == self.S):
            raise ValueError("rows of unequal length")
        idx = first.astype(np.int64)[:, None] + stride.astype(np.int64)[:, None] * np.arange(self.S)[None, :]
        inv = np.empty(self.num_params, dtype=np.int64)
        inv[idx.ravel()] = np.arange(idx.size)
        self._rows = torch.from_numpy(idx).to(self.device)
        self._inv = torch.from_numpy(inv).to(self.device)
        self._generation = 0

    @classmethod
    def generate(cls, preset: str = "c5", device: int = 0, **overrides) -> "DifferentiableNetwork":
        return cls(B.generate(preset, **overrides), device=device, own_bn=True)

    def softmax_tables(self, theta):
        """CPT entries = softmax of theta over the child states of each row."""
        import torch

        p = torch.softmax(theta[self._rows], dim=1)
        return p.reshape(-1)[self._inv]

    def beliefs(self, tables, K: int, checkpoint_every: int = 5):
        """[V, S] beliefs after K iterations; differentiable in `tables`."""
        return UnrolledLoopyBP.apply(tables, self, K, checkpoint_every)[:, : self.S]

    def beliefs_from_log(self, theta, K: int, checkpoint_every: int = 5):
        """[V, S] beliefs after K iterations with tables = softmax_tables(theta),
        the softmax fused into the library; differentiable in `theta`."""
        return UnrolledLoopyBPLog.apply(theta, self, K, checkpoint_every)[:, : self.S]

    def _set_log_params(self, theta):
        B.check(self.lib.bp_diff_set_log_params(self.h, C.c_void_p(theta.data_ptr()), _stream_ptr()),
                "bp_diff_set_log_params")

    def _backward_log(self, grad_beliefs):
        import torch

        g = torch.empty(self.num_params, dtype=torch.float32, device=self.device)
        B.check(self.lib.bp_diff_backward_log(self.h, C.c_void_p(grad_beliefs.data_ptr()), C.c_void_p(g.data_ptr()),
                                              _stream_ptr()), "bp_diff_backward_log")
        return g

    # raw handle calls (device pointers on torch's current stream)
    def _set_params(self, tables):
        B.check(self.lib.bp_diff_set_params(self.h, C.c_void_p(tables.data_ptr()), _stream_ptr()),
                "bp_diff_set_params")

    def _forward(self, K: int, c: int):
        import torch

        out = torch.empty(self.V, self.stride, dtype=torch.float32, device=self.device)
        B.check(self.lib.bp_diff_forward(self.h, K, c, C.c_void_p(out.data_ptr()), _stream_ptr()),
                "bp_diff_forward")
        self._generation += 1
        return out

    def _backward(self, grad_beliefs, tables_grad=False, initial_grad=False):
        import torch

        g = torch.empty(self.num_params, dtype=torch.float32, device=self.device)
        gt = torch.empty(self.E, self.S * self.S, dtype=torch.float32, device=self.device) if tables_grad else None
        gi = torch.empty(self.V, self.stride, dtype=torch.float32, device=self.device) if initial_grad else None
        B.check(self.lib.bp_diff_backward_ex(self.h, C.c_void_p(grad_beliefs.data_ptr()), C.c_void_p(g.data_ptr()),
                                             C.c_void_p(gt.data_ptr() if gt is not None else 0),
                                             C.c_void_p(gi.data_ptr() if gi is not None else 0), _stream_ptr()),
                "bp_diff_backward")
        return g, gt, gi

    def tables_host(self) -> np.ndarray:
        """[E, S, S] pairwise tables of the current parameters (reference edge order)."""
        import torch

        torch.cuda.synchronize(self.device)
        out = np.zeros(self.E * self.S * self.S, dtype=np.float32)
        B.check(self.lib.bp_diff_tables(self.h, out.ctypes.data_as(B.f32p)), "bp_diff_tables")
        return out.reshape(self.E, self.S, self.S)

    def close(self):
        if getattr(self, "h", None):
            self.lib.bp_diff_free(self.h)
            self.h = None
        if getattr(self, "_own_bn", False) and self.bn:
            self.lib.bp_bn_free(self.bn)
            self.bn = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _functions():
    import torch

    class UnrolledLoopyBP(torch.autograd.Function):
        """beliefs[V, stride] = K unrolled loopy iterations of tables[num_params].

        Forward: bp_diff_set_params + bp_diff_forward (checkpoints every c
        iterations).  Backward: bp_diff_backward, the adjoint kernel through
        the K iterations, recomputing between checkpoints.  A network holds
        one forward at a time: the backward must come before the network's
        next forward.
        """

        @staticmethod
        def forward(ctx, tables, net, K, c):
            if tables.dtype != torch.float32 or not tables.is_contiguous() or tables.numel() != net.num_params:
                raise ValueError("tables must be a contiguous float32 vector of num_params entries")
            net._set_params(tables)
            out = net._forward(int(K), int(c))
            ctx.net = net
            ctx.generation = net._generation
            return out

        @staticmethod
        def backward(ctx, grad):
            net = ctx.net
            if net._generation != ctx.generation:
                raise RuntimeError("another forward ran on this network before the backward")
            g, _, _ = net._backward(grad.contiguous())
            return g, None, None, None

    class UnrolledLoopyBPLog(torch.autograd.Function):
        """beliefs[V, stride] = K unrolled loopy iterations of
        tables = softmax(theta) per row, the softmax fused into the library;
        backward: bp_diff_backward_log."""

        @staticmethod
        def forward(ctx, theta, net, K, c):
            if theta.dtype != torch.float32 or not theta.is_contiguous() or theta.numel() != net.num_params:
                raise ValueError("theta must be a contiguous float32 vector of num_params entries")
            net._set_log_params(theta)
            out = net._forward(int(K), int(c))
            ctx.net = net
            ctx.generation = net._generation
            return out

        @staticmethod
        def backward(ctx, grad):
            net = ctx.net
            if net._generation != ctx.generation:
                raise RuntimeError("another forward ran on this network before the backward")
            return net._backward_log(grad.contiguous()), None, None, None

    return {"UnrolledLoopyBP": UnrolledLoopyBP, "UnrolledLoopyBPLog": UnrolledLoopyBPLog}


class _Lazy:
    """torch is imported only when the differentiable form is used."""

    _classes = None

    def __init__(self, name):
        self._name = name

    def __getattr__(self, attr):
        if _Lazy._classes is None:
            _Lazy._classes = _functions()
        return getattr(_Lazy._classes[self._name], attr)


UnrolledLoopyBP = _Lazy("UnrolledLoopyBP")
UnrolledLoopyBPLog = _Lazy("UnrolledLoopyBPLog")

__all__ = ["BeliefNetwork", "DifferentiableNetwork", "UnrolledLoopyBP", "UnrolledLoopyBPLog", "C"]
