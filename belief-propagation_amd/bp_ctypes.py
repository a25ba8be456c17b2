"""ctypes bindings of libbp_amd.so.

Mirrors the reference's C interface -- ``struct graph`` and the graph API of
src/graph/graph.h:22-133, ``parse_xml_file`` (src/bnf-xml-parser/
xml-expression.h:9), ``build_graph`` (src/bnf-parser/expression.h:58) -- plus
the device handle API of include/bp_hip.h and the generators of
include/bp_model.h.  The library is loaded from this directory; there is no
fallback: if it is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libbp_amd.so"

MAX_STATES = 10
CHAR_BUFFER_SIZE = 50

u32p = C.POINTER(C.c_uint)
f32p = C.POINTER(C.c_float)


class Graph(C.Structure):
    """struct graph of src/graph/graph.h:22-74 (field order is ABI)."""

    _fields_ = [
        ("total_num_vertices", C.c_uint),
        ("total_num_edges", C.c_uint),
        ("current_num_vertices", C.c_uint),
        ("current_num_edges", C.c_uint),
        ("max_degree", C.c_uint),
        ("edges_src_index", u32p),
        ("edges_dest_index", u32p),
        ("edges_x_dim", u32p),
        ("edges_y_dim", u32p),
        ("edges_joint_probabilities", f32p),
        ("edges_messages", f32p),
        ("last_edges_messages", f32p),
        ("current_edge_messages", C.POINTER(f32p)),
        ("previous_edge_messages", C.POINTER(f32p)),
        ("node_states", f32p),
        ("node_num_vars", u32p),
        ("src_nodes_to_edges_node_list", u32p),
        ("src_nodes_to_edges_edge_list", u32p),
        ("dest_nodes_to_edges_node_list", u32p),
        ("dest_nodes_to_edges_edge_list", u32p),
        ("levels_to_nodes", u32p),
        ("num_levels", C.c_uint),
        ("diameter", C.c_int),
        ("visited", C.POINTER(C.c_char)),
        ("node_names", C.POINTER(C.c_char)),
        ("variable_names", C.POINTER(C.c_char)),
        ("observed_nodes", C.POINTER(C.c_char)),
        ("graph_name", C.c_char * CHAR_BUFFER_SIZE),
        ("node_hash_table_created", C.c_char),
        ("node_hash_table", C.c_void_p),
        ("src_node_to_edge_table", C.c_void_p),
        ("dest_node_to_edge_table", C.c_void_p),
        ("edge_tables_created", C.c_char),
    ]


GraphP = C.POINTER(Graph)


class GenParams(C.Structure):
    """bp_gen_params of include/bp_model.h."""

    _fields_ = [
        ("kind", C.c_int),
        ("num_nodes", C.c_uint),
        ("rows", C.c_uint),
        ("cols", C.c_uint),
        ("indegree_dist", C.c_int),
        ("indeg_lo", C.c_uint),
        ("indeg_hi", C.c_uint),
        ("indeg_mean", C.c_double),
        ("states_lo", C.c_uint),
        ("states_hi", C.c_uint),
        ("cpt_dist", C.c_int),
        ("alpha", C.c_double),
        ("evidence_fraction", C.c_double),
        ("max_children", C.c_uint),
        ("seed", C.c_uint64),
    ]


class BN(C.Structure):
    """bp_bn of include/bp_model.h (a network as an XMLBIF file holds it)."""

    _fields_ = [
        ("num_vars", C.c_uint),
        ("names", C.c_char_p),
        ("num_states", u32p),
        ("num_defs", C.c_uint),
        ("def_for", u32p),
        ("def_given_begin", C.POINTER(C.c_uint64)),
        ("def_given", u32p),
        ("def_table_begin", C.POINTER(C.c_uint64)),
        ("def_table", f32p),
        ("given_cap", C.c_uint64),
        ("table_cap", C.c_uint64),
        ("def_cap", C.c_uint64),
    ]


_SIGNATURES = {
    # graph.h
    "create_graph": (GraphP, [C.c_uint, C.c_uint]),
    "graph_add_node": (None, [GraphP, C.c_uint, C.c_char_p]),
    "graph_add_and_set_node_state": (None, [GraphP, C.c_uint, C.c_char_p, f32p]),
    "graph_set_node_state": (None, [GraphP, C.c_uint, C.c_uint, f32p]),
    "graph_add_edge": (None, [GraphP, C.c_uint, C.c_uint, C.c_uint, C.c_uint, f32p]),
    "set_up_src_nodes_to_edges": (None, [GraphP]),
    "set_up_dest_nodes_to_edges": (None, [GraphP]),
    "init_levels_to_nodes": (None, [GraphP]),
    "calculate_diameter": (None, [GraphP]),
    "initialize_node": (None, [GraphP, C.c_uint, C.c_uint]),
    "node_set_state": (None, [GraphP, C.c_uint, C.c_uint, f32p]),
    "init_edge": (None, [GraphP, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_uint, f32p]),
    "send_message": (None, [f32p, C.c_uint, C.c_uint, f32p, f32p, u32p, u32p]),
    "fill_in_node_table": (None, [GraphP]),
    "fill_in_node_hash_table": (None, [GraphP]),
    "find_node_by_name": (C.c_uint, [C.c_char_p, GraphP]),
    "graph_vertex_count": (C.c_int, [GraphP]),
    "graph_edge_count": (C.c_int, [GraphP]),
    "graph_destroy": (None, [GraphP]),
    "fill_in_leaf_nodes_in_index": (None, [GraphP, u32p, u32p, C.c_uint]),
    "visit_node": (None, [GraphP, C.c_uint, u32p]),
    "propagate_using_levels_start": (None, [GraphP]),
    "propagate_using_levels": (None, [GraphP, C.c_uint]),
    "reset_visited": (None, [GraphP]),
    "marginalize": (None, [GraphP]),
    "init_previous_edge": (None, [GraphP]),
    "loopy_propagate_one_iteration": (None, [GraphP]),
    "loopy_propagate_until": (C.c_uint, [GraphP, C.c_float, C.c_uint]),
    "loopy_propagate_edge_one_iteration": (None, [GraphP]),
    "loopy_propagate_until_edge": (C.c_uint, [GraphP, C.c_float, C.c_uint]),
    "loopy_progagate_until_acc": (C.c_uint, [GraphP, C.c_float, C.c_uint]),
    "loopy_progagate_until_edge_acc": (C.c_uint, [GraphP, C.c_float, C.c_uint]),
    "print_node": (None, [GraphP, C.c_uint]),
    "print_edge": (None, [GraphP, C.c_uint]),
    "print_nodes": (None, [GraphP]),
    "print_edges": (None, [GraphP]),
    "print_src_nodes_to_edges": (None, [GraphP]),
    "print_dest_nodes_to_edges": (None, [GraphP]),
    "print_levels_to_nodes": (None, [GraphP]),
    # xml-expression.h / expression.h
    "parse_xml_file": (GraphP, [C.c_char_p]),
    "parse_bif_file": (GraphP, [C.c_char_p]),
    "parse_bif_expression": (C.c_void_p, [C.c_char_p]),
    "parse_bif_string": (C.c_void_p, [C.c_char_p]),
    "build_graph": (GraphP, [C.c_void_p]),
    "create_expression": (C.c_void_p, [C.c_int, C.c_void_p, C.c_void_p]),
    "delete_expression": (None, [C.c_void_p]),
    "print_expression": (None, [C.c_void_p]),
    # bp_hip.h
    "bp_last_error": (C.c_char_p, []),
    "bp_set_abort_on_error": (C.c_int, [C.c_int]),
    "bp_device_count": (C.c_int, []),
    "bp_set_node_order": (C.c_int, [C.c_int]),
    "loopy_propagate_until_hip": (C.c_uint, [GraphP, C.c_float, C.c_uint]),
    "loopy_propagate_until_hip_edge": (C.c_uint, [GraphP, C.c_float, C.c_uint]),
    "bp_upload": (C.c_int, [GraphP, C.c_int, C.POINTER(C.c_void_p)]),
    "bp_free": (None, [C.c_void_p]),
    "bp_upload_state": (C.c_int, [C.c_void_p, GraphP]),
    "bp_download": (C.c_int, [C.c_void_p, GraphP]),
    "bp_download_states": (C.c_int, [C.c_void_p, f32p]),
    "bp_loopy_init": (C.c_int, [C.c_void_p]),
    "bp_restore_states": (C.c_int, [C.c_void_p]),
    "bp_loopy_iterate": (C.c_int, [C.c_void_p, C.c_uint, C.POINTER(C.c_double)]),
    "bp_loopy_iterate_noalloc": (C.c_int, [C.c_void_p, C.c_uint, C.POINTER(C.c_double)]),
    "bp_loopy_until": (C.c_int, [C.c_void_p, C.c_float, C.c_uint, C.POINTER(C.c_uint)]),
    "bp_edge_iterate": (C.c_int, [C.c_void_p, C.c_uint]),
    "bp_edge_until": (C.c_int, [C.c_void_p, C.c_float, C.c_uint, C.POINTER(C.c_uint)]),
    "bp_tree": (C.c_int, [C.c_void_p]),
    "bp_sync": (C.c_int, [C.c_void_p]),
    "bp_dev_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "bp_loopy_algorithmic_bytes": (C.c_double, [C.c_void_p]),
    "bp_profile_loopy": (C.c_int, [C.c_void_p, C.c_uint, f32p]),
    # bp_model.h
    "bp_gen_preset": (C.c_int, [C.c_char_p, C.POINTER(GenParams)]),
    "bp_bn_generate": (C.c_void_p, [C.POINTER(GenParams)]),
    "bp_bn_add_evidence": (C.c_int, [C.c_void_p, C.c_double, C.c_uint64]),
    "bp_bn_truncate_defs": (None, [C.c_void_p, C.c_uint]),
    "bp_bn_to_graph": (GraphP, [C.c_void_p]),
    "bp_bn_write_xmlbif": (C.c_int, [C.c_void_p, C.c_char_p]),
    "bp_bn_free": (None, [C.c_void_p]),
    "bp_bn_create": (C.c_void_p, [C.c_uint]),
    "bp_bn_set_var": (C.c_int, [C.c_void_p, C.c_uint, C.c_char_p, C.c_uint]),
    "bp_bn_add_def": (C.c_int, [C.c_void_p, C.c_uint, u32p, C.c_uint, f32p, C.c_uint64]),
    # bp_diff.h
    "bp_diff_create": (C.c_int, [C.c_void_p, C.c_uint, C.c_int, C.POINTER(C.c_void_p)]),
    "bp_diff_free": (None, [C.c_void_p]),
    "bp_diff_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "bp_diff_rows": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), u32p, u32p]),
    "bp_diff_set_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "bp_diff_forward": (C.c_int, [C.c_void_p, C.c_uint, C.c_uint, C.c_void_p, C.c_void_p]),
    "bp_diff_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "bp_diff_backward_ex": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "bp_diff_tables": (C.c_int, [C.c_void_p, f32p]),
    "bp_diff_set_log_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "bp_diff_backward_log": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    # bp_batch.h
    "bp_batch_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "bp_batch_free": (None, [C.c_void_p]),
    "bp_batch_add": (C.c_int, [C.c_void_p, GraphP, C.POINTER(C.c_uint)]),
    "bp_batch_upload": (C.c_int, [C.c_void_p, C.c_int]),
    "bp_batch_set_config": (C.c_int, [C.c_void_p, C.c_uint, C.c_uint, C.c_uint]),
    "bp_batch_get_config": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint), C.POINTER(C.c_uint), C.POINTER(C.c_uint)]),
    "bp_batch_tune": (C.c_int, [C.c_void_p, C.c_uint, C.c_char_p]),
    "bp_batch_init": (C.c_int, [C.c_void_p]),
    "bp_batch_iterate": (C.c_int, [C.c_void_p, C.c_uint, C.POINTER(C.c_double)]),
    "bp_batch_download": (C.c_int, [C.c_void_p, C.c_uint, GraphP]),
    "bp_batch_sync": (C.c_int, [C.c_void_p]),
    "bp_batch_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "bp_batch_profile": (C.c_int, [C.c_void_p, C.c_uint, f32p]),
}

_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Loads libbp_amd.so once (path, else $BP_AMD_LIB, else the in-tree
    build); raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else Path(os.environ.get("BP_AMD_LIB") or LIB_PATH)
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(str(p))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def last_error() -> str:
    return (load().bp_last_error() or b"").decode(errors="replace")


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {last_error()}")


# ---------------------------------------------------------------------------
# numpy views of the public arrays

def view_u32(ptr, n: int) -> np.ndarray:
    if n == 0:
        return np.zeros(0, dtype=np.uint32)
    return np.ctypeslib.as_array(ptr, shape=(n,))


def view_f32(ptr, n: int) -> np.ndarray:
    if n == 0:
        return np.zeros(0, dtype=np.float32)
    return np.ctypeslib.as_array(ptr, shape=(n,))


def view_bytes(ptr, n: int) -> np.ndarray:
    if n == 0:
        return np.zeros(0, dtype=np.uint8)
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint8)), shape=(n,))


def graph_arrays(g: GraphP) -> dict:
    """Copies of every array of the graph, in the reference layout."""
    G = g.contents
    V, E = G.current_num_vertices, G.current_num_edges
    out = {
        "meta": np.array([V, E, G.total_num_vertices, G.total_num_edges, G.max_degree, G.num_levels], dtype=np.uint32),
        "node_num_vars": view_u32(G.node_num_vars, V).copy(),
        "node_states": view_f32(G.node_states, V * MAX_STATES).copy(),
        "observed": view_bytes(G.observed_nodes, V).copy(),
        "node_names": view_bytes(G.node_names, V * CHAR_BUFFER_SIZE).copy(),
        "edges_src": view_u32(G.edges_src_index, E).copy(),
        "edges_dst": view_u32(G.edges_dest_index, E).copy(),
        "edges_x": view_u32(G.edges_x_dim, E).copy(),
        "edges_y": view_u32(G.edges_y_dim, E).copy(),
        "edges_J": view_f32(G.edges_joint_probabilities, E * MAX_STATES * MAX_STATES).copy(),
        "src_node_list": view_u32(G.src_nodes_to_edges_node_list, V).copy(),
        "src_edge_list": view_u32(G.src_nodes_to_edges_edge_list, E).copy(),
        "dst_node_list": view_u32(G.dest_nodes_to_edges_node_list, V).copy(),
        "dst_edge_list": view_u32(G.dest_nodes_to_edges_edge_list, E).copy(),
        "levels_to_nodes": view_u32(G.levels_to_nodes, 2 * V).copy(),
        "num_levels": np.array([G.num_levels], dtype=np.uint32),
    }
    return out


def masked_rows(flat: np.ndarray, count: int, valid: np.ndarray) -> np.ndarray:
    """[count][MAX_STATES] with entries beyond each row's valid prefix zeroed."""
    a = flat.reshape(count, MAX_STATES).copy()
    mask = np.arange(MAX_STATES)[None, :] < valid[:, None]
    a[~mask] = 0
    return a


def states(g: GraphP) -> np.ndarray:
    G = g.contents
    V = G.current_num_vertices
    return masked_rows(view_f32(G.node_states, V * MAX_STATES), V, view_u32(G.node_num_vars, V))


def messages(g: GraphP, which: str) -> np.ndarray:
    """'cur' / 'prev' (through the pointer the reference swaps) or 'tree' (edges_messages)."""
    G = g.contents
    E = G.current_num_edges
    if which == "cur":
        ptr = G.current_edge_messages.contents
    elif which == "prev":
        ptr = G.previous_edge_messages.contents
    else:
        ptr = G.edges_messages
    return masked_rows(view_f32(ptr, E * MAX_STATES), E, view_u32(G.edges_x_dim, E))


def generate(preset: str, **overrides):
    """Seeded network of a preset ("c1", "c2", "c2b", "tree", "c3", "c4net", "c5"); returns a bp_bn handle."""
    lib = load()
    p = GenParams()
    if lib.bp_gen_preset(preset.encode(), C.byref(p)) != 0:
        raise ValueError(f"unknown preset {preset}")
    for k, v in overrides.items():
        setattr(p, k, v)
    bn = lib.bp_bn_generate(C.byref(p))
    if not bn:
        raise RuntimeError(f"generator failed: {last_error()}")
    return bn


def bn_arrays(bn) -> dict:
    """Copies of a bp_bn's arrays (state counts, definitions, tables)."""
    b = C.cast(bn, C.POINTER(BN)).contents
    nd = b.num_defs
    gb = np.ctypeslib.as_array(b.def_given_begin, shape=(nd + 1,)).copy()
    tb = np.ctypeslib.as_array(b.def_table_begin, shape=(nd + 1,)).copy()
    return {
        "num_states": view_u32(b.num_states, b.num_vars).copy(),
        "def_for": view_u32(b.def_for, nd).copy(),
        "given_begin": gb,
        "given": view_u32(b.def_given, int(gb[-1])).copy() if gb[-1] else np.zeros(0, np.uint32),
        "table_begin": tb,
        "table": view_f32(b.def_table, int(tb[-1])).copy() if tb[-1] else np.zeros(0, np.float32),
    }


def bn_to_graph(bn, csr: bool = True, levels: bool = False) -> GraphP:
    lib = load()
    g = lib.bp_bn_to_graph(bn)
    if not g:
        raise RuntimeError(f"bp_bn_to_graph failed: {last_error()}")
    if csr:
        lib.set_up_src_nodes_to_edges(g)
        lib.set_up_dest_nodes_to_edges(g)
    if levels:
        lib.init_levels_to_nodes(g)
    return g


class Batch:
    """Many networks run together (bp_batch); frees itself."""

    def __init__(self):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.bp_batch_create(C.byref(h)), "bp_batch_create")
        self.h = h
        self.n = 0

    def add(self, g: GraphP) -> int:
        i = C.c_uint()
        check(self.lib.bp_batch_add(self.h, g, C.byref(i)), "bp_batch_add")
        self.n += 1
        return i.value

    def upload(self, device: int = -1):
        check(self.lib.bp_batch_upload(self.h, device), "bp_batch_upload")

    def config(self) -> tuple[int, int, int]:
        """(slots per lane, lanes per slot, queue block)"""
        r, l, q = C.c_uint(), C.c_uint(), C.c_uint()
        check(self.lib.bp_batch_get_config(self.h, C.byref(r), C.byref(l), C.byref(q)), "bp_batch_get_config")
        return r.value, l.value, q.value

    def set_config(self, slots_per_lane: int, lanes_per_slot: int, queue_block: int):
        check(self.lib.bp_batch_set_config(self.h, slots_per_lane, lanes_per_slot, queue_block),
              "bp_batch_set_config")

    def tune(self, iterations: int = 4, cache: str | None = None):
        check(self.lib.bp_batch_tune(self.h, iterations, cache.encode() if cache else None), "bp_batch_tune")

    def init(self):
        check(self.lib.bp_batch_init(self.h), "bp_batch_init")

    def iterate(self, n: int, want_deltas: bool = False):
        if want_deltas:
            d = np.zeros(max(n, 1) * self.n, dtype=np.float64)
            check(self.lib.bp_batch_iterate(self.h, n, d.ctypes.data_as(C.POINTER(C.c_double))), "bp_batch_iterate")
            return d[: n * self.n].reshape(n, self.n)
        check(self.lib.bp_batch_iterate(self.h, n, None), "bp_batch_iterate")
        return None

    def download(self, index: int, g: GraphP):
        check(self.lib.bp_batch_download(self.h, index, g), "bp_batch_download")

    def sync(self):
        check(self.lib.bp_batch_sync(self.h), "bp_batch_sync")

    def info(self) -> dict:
        a = (C.c_uint64 * 8)()
        check(self.lib.bp_batch_info(self.h, a), "bp_batch_info")
        return {"instances": a[0], "V": a[1], "E": a[2], "S": a[3], "table_floats": a[4], "chunks": a[5],
                "alg_bytes": a[6]}

    def profile(self, n: int) -> float:
        ms = C.c_float()
        check(self.lib.bp_batch_profile(self.h, n, C.byref(ms)), "bp_batch_profile")
        return ms.value

    def free(self):
        if self.h:
            self.lib.bp_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Device:
    """A device-resident graph (bp_dev); frees itself."""

    def __init__(self, g: GraphP, device: int = -1):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.bp_upload(g, device, C.byref(h)), "bp_upload")
        self.h = h

    def info(self) -> dict:
        a = (C.c_uint64 * 8)()
        self.lib.bp_dev_info(self.h, a)
        return {"V": a[0], "E": a[1], "stride": a[2], "S": a[3], "chunks": a[4], "mixed": a[5], "tree_waves": a[6],
                "shared_tables": a[7]}

    def init(self):
        check(self.lib.bp_loopy_init(self.h), "bp_loopy_init")

    def restore(self):
        check(self.lib.bp_restore_states(self.h), "bp_restore_states")

    def iterate(self, n: int, want_deltas: bool = False):
        if want_deltas:
            d = (C.c_double * max(n, 1))()
            check(self.lib.bp_loopy_iterate(self.h, n, d), "bp_loopy_iterate")
            return np.array(d[:n])
        check(self.lib.bp_loopy_iterate(self.h, n, None), "bp_loopy_iterate")
        return None

    def until(self, conv: float, max_iter: int) -> int:
        r = C.c_uint()
        check(self.lib.bp_loopy_until(self.h, conv, max_iter, C.byref(r)), "bp_loopy_until")
        return r.value

    def tree(self):
        check(self.lib.bp_tree(self.h), "bp_tree")

    def sync(self):
        check(self.lib.bp_sync(self.h), "bp_sync")

    def download(self, g: GraphP):
        check(self.lib.bp_download(self.h, g), "bp_download")

    def algorithmic_bytes(self) -> float:
        return float(self.lib.bp_loopy_algorithmic_bytes(self.h))

    def profile(self, n: int) -> np.ndarray:
        ms = (C.c_float * n)()
        check(self.lib.bp_profile_loopy(self.h, n, ms), "bp_profile_loopy")
        return np.array(ms[:n], dtype=np.float64)

    def free(self):
        if self.h:
            self.lib.bp_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
