"""TEST INFRASTRUCTURE: ctypes bindings of liboracle.so (the CPU oracle).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker or the timed CPU baseline.  Functions take the
Graph_t of libbp_amd.so (the structure is the reference's public struct
layout) and compute on its host arrays; see bp_oracle.c.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_lib = None
_plants: dict = {}

# deliberately wrong builds of the oracle (ORC_PLANT_* in bp_oracle_core.h),
# loaded only by the tests that show the checks reject them
PLANTS = ("pad_one", "stale", "norm_stride")


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    _lib = _bind(C.CDLL(str(LIB_PATH)))
    return _lib


def load_plant(name: str):
    """A planted-mistake build of the oracle (built on first use)."""
    if name not in PLANTS:
        raise ValueError(f"unknown planted oracle {name!r}")
    if name not in _plants:
        path = HERE / f"liboracle_plant_{name}.so"
        if not path.exists():
            subprocess.run(["make", "-s", "-C", str(HERE), "plants"], check=True)
        _plants[name] = _bind(C.CDLL(str(path)))
    return _plants[name]


def _bind(lib):
    vp = C.c_void_p
    sig = {
        "orc_init_previous_edge": (None, [vp]),
        "orc_loopy_iteration": (None, [vp]),
        "orc_loopy_iteration_omp": (None, [vp, C.c_int]),
        "orc_loopy_until": (C.c_uint, [vp, C.c_float, C.c_uint, C.c_int, C.POINTER(C.c_float)]),
        "orc_edge_iteration": (None, [vp]),
        "orc_delta": (C.c_float, [vp]),
        "orc_delta64": (C.c_double, [vp]),
        "orc_send_message": (None, [C.POINTER(C.c_float), C.c_uint, vp, C.c_uint, C.POINTER(C.c_float)]),
        "orc_tree_start": (None, [vp]),
        "orc_tree_level": (None, [vp, C.c_uint]),
        "orc_tree_marginalize": (None, [vp]),
        "orc_tree_run": (None, [vp]),
        "orc64_create": (vp, [vp]),
        "orc64_free": (None, [vp]),
        "orc64_init_previous_edge": (None, [vp]),
        "orc64_loopy_iteration": (None, [vp, C.c_int]),
        "orc64_tree_run": (None, [vp]),
        "orc64_get_states": (None, [vp, C.POINTER(C.c_double)]),
        "orc64_get_messages": (None, [vp, C.c_int, C.POINTER(C.c_double)]),
        "orc64_get_tree_messages": (None, [vp, C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def _vp(g):
    return C.cast(g, C.c_void_p)


def init_previous_edge(g, lib=None):
    (lib or load()).orc_init_previous_edge(_vp(g))


def loopy_iteration(g, threads: int = 1, lib=None):
    lib = lib or load()
    if threads > 1:
        lib.orc_loopy_iteration_omp(_vp(g), threads)
    else:
        lib.orc_loopy_iteration(_vp(g))


def loopy_until(g, conv: float, max_iter: int, threads: int = 1):
    d = (C.c_float * max(max_iter, 1))()
    it = load().orc_loopy_until(_vp(g), conv, max_iter, threads, d)
    return it, np.array(d[: min(it + 1, max_iter)], dtype=np.float32)


def edge_iteration(g):
    load().orc_edge_iteration(_vp(g))


def delta(g) -> float:
    return float(load().orc_delta(_vp(g)))


def delta64(g) -> float:
    return float(load().orc_delta64(_vp(g)))


def tree_run(g, lib=None):
    (lib or load()).orc_tree_run(_vp(g))


class FP64:
    """The FP64 oracle on a widened copy of a graph (the graph must outlive it)."""

    def __init__(self, g):
        self.lib = load()
        self.g = g
        self.h = self.lib.orc64_create(_vp(g))
        if not self.h:
            raise MemoryError("orc64_create")
        G = g.contents
        self.V, self.E = G.current_num_vertices, G.current_num_edges

    def init(self):
        self.lib.orc64_init_previous_edge(self.h)

    def iterate(self, n: int = 1, threads: int = 1):
        for _ in range(n):
            self.lib.orc64_loopy_iteration(self.h, threads)

    def tree(self):
        self.lib.orc64_tree_run(self.h)

    def states(self) -> np.ndarray:
        out = np.zeros(self.V * 10, dtype=np.float64)
        self.lib.orc64_get_states(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out.reshape(self.V, 10)

    def messages(self, previous: bool) -> np.ndarray:
        out = np.zeros(self.E * 10, dtype=np.float64)
        self.lib.orc64_get_messages(self.h, int(previous), out.ctypes.data_as(C.POINTER(C.c_double)))
        return out.reshape(self.E, 10)

    def tree_messages(self) -> np.ndarray:
        out = np.zeros(self.E * 10, dtype=np.float64)
        self.lib.orc64_get_tree_messages(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out.reshape(self.E, 10)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc64_free(self.h)
            self.h = None
