"""Host side of the drop-in: construction, CSR, levels and the XMLBIF / BIF
readers against what the reference built (bit for bit), plus the generator
and the error paths.  CPU only."""
from __future__ import annotations

import ctypes as C
import glob
import hashlib
import subprocess
from pathlib import Path

import numpy as np
import pytest

import bp_ctypes as B
from conftest import GOLDEN, INPUTS, REFERENCE, ROOT, golden, graph_for_case, overrides, same_bits

STRUCT_KEYS = ["node_num_vars", "observed", "node_names", "edges_src", "edges_dst", "edges_x", "edges_y",
               "src_node_list", "src_edge_list", "dst_node_list", "dst_edge_list"]


def mine_struct(g, S: int, levels: bool) -> dict:
    a = B.graph_arrays(g)
    V, E = int(a["meta"][0]), int(a["meta"][1])
    out = {k: a[k] for k in STRUCT_KEYS}
    out["meta"] = a["meta"]
    out["node_states"] = B.masked_rows(a["node_states"], V, a["node_num_vars"])[:, :S]
    J = a["edges_J"].reshape(E, 10, 10).copy()
    for e in range(E):
        J[e, a["edges_x"][e]:, :] = 0
        J[e, :, a["edges_y"][e]:] = 0
    out["edges_J"] = J[:, :S, :S]
    if levels:
        lv = a["levels_to_nodes"].copy()
        lv[int(a["num_levels"][0]):V] = 0
        out["levels_to_nodes"] = lv
        out["num_levels"] = a["num_levels"]
    return out


def compare_struct(mine: dict, d: dict, levels: bool = True):
    keys = STRUCT_KEYS + ["node_states", "edges_J"] + (["levels_to_nodes", "num_levels"] if levels else [])
    for k in keys:
        assert same_bits(mine[k], d[k]), f"{k} differs"
    assert (mine["meta"][:5] == d["meta"][:5]).all(), f"meta {mine['meta']} vs {d['meta']}"


def digest(mine: dict, d: dict, levels: bool = True):
    for k in STRUCT_KEYS + ["node_states", "edges_J"] + (["levels_to_nodes"] if levels else []):
        h = hashlib.sha256(np.ascontiguousarray(mine[k]).tobytes()).hexdigest()
        assert h == str(d[f"sha_{k}"]), f"{k} digest differs"


# ---------------------------------------------------------------------------
# layout and exports


def test_struct_graph_layout_matches_header(tmp_path):
    """ctypes mirror == include/graph.h, field by field (the reference's layout)."""
    src = tmp_path / "off.c"
    fields = [f for f, _ in B.Graph._fields_]
    body = "\n".join(f'printf("%zu\\n", offsetof(struct graph, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "graph.h"\nint main(void){\n'
                   f'printf("%zu\\n", sizeof(struct graph));\n{body}\nreturn 0;}}\n')
    exe = tmp_path / "off"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(B.Graph)] + [getattr(B.Graph, f).offset for f in fields]
    assert got == want


@pytest.mark.skipif(not Path("/root/reference/src/graph/graph.h").exists(), reason="reference sources absent")
def test_struct_graph_layout_matches_reference(tmp_path):
    src = tmp_path / "off.c"
    fields = [f for f, _ in B.Graph._fields_]
    body = "\n".join(f'printf("%zu\\n", offsetof(struct graph, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "graph/graph.h"\nint main(void){\n'
                   f'printf("%zu\\n", sizeof(struct graph));\n{body}\nreturn 0;}}\n')
    exe = tmp_path / "off"
    subprocess.run(["gcc", "-D_GNU_SOURCE", "-I", "/root/reference/src", str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(B.Graph)] + [getattr(B.Graph, f).offset for f in fields]
    assert got == want


def declared_functions() -> list[str]:
    import re

    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\([^;{]*\)\s*;", text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_library_exports_every_declared_function(lib):
    names = declared_functions()
    assert "create_graph" in names and "bp_upload" in names and "build_graph" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared in include/ but not exported: {missing}"


# ---------------------------------------------------------------------------
# construction against the reference


@pytest.mark.parametrize("case", ["example", "dog_xml", "bf_1000_2000_1", "c1", "c1_mixed", "lattice16"])
def test_structure_bitwise(lib, case):
    d = golden(case)
    S = d["node_states"].shape[1]
    g = graph_for_case(d, levels=True)
    compare_struct(mine_struct(g, S, True), d)
    lib.graph_destroy(g)


@pytest.mark.parametrize("case", ["tree10k", "c5_small", "c4net_small"])
def test_generated_structure_digest(lib, case):
    d = golden(case)
    g = graph_for_case(d, levels=True)
    S = int(B.view_u32(g.contents.node_num_vars, g.contents.current_num_vertices).max())
    digest(mine_struct(g, S, True), d)
    lib.graph_destroy(g)


def test_dog_xml_reader(lib):
    d = golden("dog_xml")
    g = lib.parse_xml_file(str(INPUTS / "dog.xml").encode())
    lib.set_up_src_nodes_to_edges(g)
    lib.set_up_dest_nodes_to_edges(g)
    lib.init_levels_to_nodes(g)
    compare_struct(mine_struct(g, 2, True), d)
    lib.graph_destroy(g)


@pytest.mark.skipif(not REFERENCE.exists(), reason="reference data files absent")
@pytest.mark.parametrize("path", sorted(glob.glob(str(REFERENCE / "xml" / "bf_*.xml"))))
def test_bundled_xml_reader(lib, path):
    name = Path(path).stem
    d = golden(name if name == "bf_1000_2000_1" else "struct_" + name)
    g = lib.parse_xml_file(path.encode())
    lib.set_up_src_nodes_to_edges(g)
    lib.set_up_dest_nodes_to_edges(g)
    lib.init_levels_to_nodes(g)
    compare_struct(mine_struct(g, int(d["node_states"].shape[1]), True), d)
    lib.graph_destroy(g)


BIF_INPUTS = {"bif_dog": "dog.bif", "bif_small_asia": "asia.bif", "bif_small_cancer": "cancer.bif",
              "bif_small_earthquake": "earthquake.bif", "bif_small_sachs": "sachs.bif",
              "bif_small_survey": "survey.bif", "bif_medium_alarm": "alarm.bif"}


@pytest.mark.parametrize("case", sorted(BIF_INPUTS))
def test_bif_reader(lib, case):
    d = golden(case)
    g = lib.parse_bif_file(str(INPUTS / BIF_INPUTS[case]).encode())
    assert g, B.last_error()
    lib.set_up_src_nodes_to_edges(g)
    lib.set_up_dest_nodes_to_edges(g)
    lib.init_levels_to_nodes(g)
    compare_struct(mine_struct(g, int(d["node_states"].shape[1]), True), d)
    lib.graph_destroy(g)


@pytest.mark.skipif(not REFERENCE.exists(), reason="reference data files absent")
@pytest.mark.parametrize("case", ["bif_medium_insurance", "bif_medium_water", "bif_large_hepar2",
                                  "bif_large_win95pts", "bif_carpo"])
def test_bundled_bif_reader(lib, case):
    d = golden(case)
    g = lib.parse_bif_file(str(REFERENCE / str(d["source"])).encode())
    assert g, B.last_error()
    lib.set_up_src_nodes_to_edges(g)
    lib.set_up_dest_nodes_to_edges(g)
    lib.init_levels_to_nodes(g)
    S = int(B.view_u32(g.contents.node_num_vars, g.contents.current_num_vertices).max())
    m = mine_struct(g, S, True)
    if "sha_node_num_vars" in d:
        digest(m, d)
    else:
        compare_struct(m, d)
    lib.graph_destroy(g)


def test_bif_rejects_what_the_reference_rejects(lib):
    """child.bif has a syntax error at line 31 for the reference's parser too."""
    assert not lib.parse_bif_file(str(INPUTS / "child.bif").encode())
    assert "line 31" in B.last_error()


def test_bif_state_names_partial_swap(lib):
    """Outcome names are stored last-first, then swapped j <-> n/2-j
    (src/bnf-parser/expression.c:902-917): a, b, c -> b, c, a."""
    text = b"""network n { }
variable x { type discrete [ 3 ] { a, b, c }; }
variable y { type discrete [ 2 ] { p, q }; }
probability ( x ) { table 0.2 0.3 0.5; }
probability ( y | x ) { (a) 0.1 0.9; (b) 0.4 0.6; (c) 0.7 0.3; }
"""
    e = lib.parse_bif_string(text)
    assert e
    g = lib.build_graph(e)
    lib.delete_expression(e)
    G = g.contents
    names = B.view_bytes(G.variable_names, 2 * 500).reshape(2, 10, 50)
    # node 0 is the last declared variable (y)
    got_y = [bytes(names[0, k]).split(b"\0")[0] for k in range(2)]
    got_x = [bytes(names[1, k]).split(b"\0")[0] for k in range(3)]
    assert got_y == [b"p", b"q"] and got_x == [b"b", b"c", b"a"]
    lib.graph_destroy(g)


# ---------------------------------------------------------------------------
# generator and XMLBIF writer


@pytest.mark.parametrize("case", ["c1", "c1_mixed", "lattice16", "tree10k", "c5_small", "c4net_small", "c2b", "c2"])
def test_generator_reproduces_fixture_input(lib, tmp_path, case):
    """The seeded generator writes byte-identical XMLBIF to what the
    reference read when the fixture was made."""
    d = golden(case)
    bn = B.generate(str(d["preset"]), **(overrides(d) if "overrides" in d else {}))
    out = tmp_path / "net.xml"
    assert lib.bp_bn_write_xmlbif(bn, str(out).encode()) == 0
    assert hashlib.sha256(out.read_bytes()).hexdigest() == str(d["xml_sha256"])
    lib.bp_bn_free(bn)


def test_generated_graph_equals_parsed_xml(lib, tmp_path):
    bn = B.generate("c1", num_nodes=3000, states_hi=4)
    out = tmp_path / "net.xml"
    lib.bp_bn_write_xmlbif(bn, str(out).encode())
    g = B.bn_to_graph(bn, levels=True)
    h = lib.parse_xml_file(str(out).encode())
    lib.set_up_src_nodes_to_edges(h)
    lib.set_up_dest_nodes_to_edges(h)
    lib.init_levels_to_nodes(h)
    a, b = B.graph_arrays(g), B.graph_arrays(h)
    for k in a:
        assert same_bits(a[k], b[k]), k
    for x in (g, h):
        lib.graph_destroy(x)
    lib.bp_bn_free(bn)


def test_c2b_full_scale_structure(lib):
    """500k-node generated network: structure digest equals the reference's parse."""
    d = golden("c2b")
    g = graph_for_case(d)
    m = mine_struct(g, 2, False)
    digest(m, d, levels=False)
    lib.graph_destroy(g)


# ---------------------------------------------------------------------------
# levels


def test_levels_match_reference_scan(lib):
    """init_levels_to_nodes (O(V+E)) == the reference's round structure run
    with the exported scan-based helpers (fill_in_leaf_nodes_in_index,
    visit_node), on a network with loops."""
    d = golden("bf_1000_2000_1")
    g = graph_for_case(d)
    G = g.contents
    V = G.current_num_vertices
    lib.reset_visited(g)
    start, end = C.c_uint(V), C.c_uint(V)
    lib.fill_in_leaf_nodes_in_index(g, C.byref(start), C.byref(end), 2)
    levels = 0
    L = B.view_u32(G.levels_to_nodes, 2 * V)
    while end.value < 2 * V:
        copy_end = end.value
        for i in range(start.value, copy_end):
            lib.visit_node(g, i, C.byref(end))
        levels += 1
        L[levels] = copy_end
    ref = L.copy()
    ref[levels + 1:V] = 0
    lib.init_levels_to_nodes(g)
    got = B.view_u32(G.levels_to_nodes, 2 * V).copy()
    got[G.num_levels:V] = 0
    assert G.num_levels == levels + 1
    assert same_bits(got, ref)
    assert same_bits(got, d["levels_to_nodes"])
    lib.graph_destroy(g)


def test_levels_every_node_once(lib):
    bn = B.generate("tree", num_nodes=50000)
    g = B.bn_to_graph(bn, levels=True)
    G = g.contents
    V = G.current_num_vertices
    order = B.view_u32(G.levels_to_nodes, 2 * V)[V:]
    assert np.array_equal(np.sort(order), np.arange(V, dtype=np.uint32))
    lib.graph_destroy(g)
    lib.bp_bn_free(bn)


def test_diameter_matches_floyd_warshall(lib):
    d = golden("dog_xml")
    g = graph_for_case(d)
    lib.calculate_diameter(g)
    G = g.contents
    V, E = G.current_num_vertices, G.current_num_edges
    INF = 99999
    dist = np.full((V, V), INF, dtype=np.int64)
    for e in range(E):
        dist[G.edges_src_index[e], G.edges_dest_index[e]] = 1
    for k in range(V):
        dist = np.minimum(dist, dist[:, k:k + 1] + dist[k:k + 1, :])
    finite = dist[dist < INF]
    assert G.diameter == (finite.max() if finite.size else -1)
    lib.graph_destroy(g)


# ---------------------------------------------------------------------------
# error paths (status mode: bp_set_abort_on_error(0))


def test_too_many_states(lib):
    g = lib.create_graph(1, 0)
    lib.graph_add_node(g, 11, b"big")
    assert "MAX_STATES" in B.last_error()
    assert g.contents.current_num_vertices == 0
    lib.graph_destroy(g)


def test_edge_dims_must_match_nodes(lib):
    g = lib.create_graph(2, 1)
    lib.graph_add_node(g, 2, b"a")
    lib.graph_add_node(g, 3, b"b")
    tab = (C.c_float * 100)()
    lib.graph_add_edge(g, 0, 1, 2, 2, tab)
    assert "does not match" in B.last_error()
    assert g.contents.current_num_edges == 0
    lib.graph_destroy(g)


def test_edge_capacity(lib):
    g = lib.create_graph(2, 1)
    lib.graph_add_node(g, 2, b"a")
    lib.graph_add_node(g, 2, b"b")
    tab = (C.c_float * 100)()
    lib.graph_add_edge(g, 0, 1, 2, 2, tab)
    lib.graph_add_edge(g, 1, 0, 2, 2, tab)
    assert "created for 1 edges" in B.last_error()
    lib.graph_destroy(g)


def test_unknown_node_name(lib):
    g = lib.create_graph(1, 0)
    lib.graph_add_node(g, 2, b"a")
    assert lib.find_node_by_name(b"zz", g) == 0xFFFFFFFF
    assert "no node named" in B.last_error()
    lib.graph_destroy(g)


def test_first_node_with_a_name_wins(lib):
    g = lib.create_graph(3, 0)
    for n in (b"a", b"b", b"a"):
        lib.graph_add_node(g, 2, n)
    assert lib.find_node_by_name(b"a", g) == 0
    lib.graph_destroy(g)


def test_unreachable_levels_are_an_error(lib):
    """The reference's level loop never ends when a node cannot be reached
    (src/graph/graph.c:815-822); the drop-in reports it."""
    g = lib.create_graph(5, 12)
    for n in (b"a", b"b", b"c", b"d", b"e"):
        lib.graph_add_node(g, 2, n)
    tab = (C.c_float * 100)(*([0.5] * 100))
    # a complete 4-node digraph (in-degree 3 everywhere) that the only
    # level-0 node (isolated node 0) cannot reach
    for s in range(1, 5):
        for t in range(1, 5):
            if s != t:
                lib.graph_add_edge(g, s, t, 2, 2, tab)
    lib.set_up_src_nodes_to_edges(g)
    lib.set_up_dest_nodes_to_edges(g)
    lib.init_levels_to_nodes(g)
    assert "unreachable" in B.last_error()
    lib.graph_destroy(g)


def test_max_degree_is_lifted(lib):
    g = lib.create_graph(31, 60)
    for i in range(31):
        lib.graph_add_node(g, 2, f"n{i}".encode())
    tab = (C.c_float * 100)(*([0.5] * 100))
    for i in range(1, 31):
        lib.graph_add_edge(g, i, 0, 2, 2, tab)
        lib.graph_add_edge(g, 0, i, 2, 2, tab)
    lib.set_up_src_nodes_to_edges(g)
    lib.set_up_dest_nodes_to_edges(g)
    assert g.contents.max_degree == 30
    lib.graph_destroy(g)


def test_missing_xml_file(lib):
    assert not lib.parse_xml_file(b"/nonexistent/net.xml")
    assert "cannot open" in B.last_error()


def test_malformed_xml(lib, tmp_path):
    p = tmp_path / "bad.xml"
    p.write_text("<BIF><NETWORK><VARIABLE><NAME>a</NAME><OUTCOME>x</OUTCOME></VARIABLE>")
    assert not lib.parse_xml_file(str(p).encode())
    assert "ends inside an element" in B.last_error()


def test_xml_entities_comments_and_cdata(lib, tmp_path):
    p = tmp_path / "ent.xml"
    p.write_text("""<?xml version="1.0"?>
<!DOCTYPE BIF [ <!ELEMENT BIF (NETWORK)*> ]>
<BIF VERSION="0.3"><NETWORK><NAME>n</NAME>
<VARIABLE TYPE="nature"><NAME>a&amp;b</NAME><OUTCOME>t</OUTCOME><OUTCOME>f</OUTCOME></VARIABLE>
<VARIABLE TYPE="nature"><NAME>c</NAME><OUTCOME>t</OUTCOME><OUTCOME><![CDATA[f]]></OUTCOME></VARIABLE>
<!-- a comment -->
<DEFINITION><FOR>a&amp;b</FOR><TABLE>0.3 0.7</TABLE></DEFINITION>
<DEFINITION><FOR>c</FOR><GIVEN>a&amp;b</GIVEN><TABLE>9 9 <!-- x --> 0.1 0.2 0.9 0.8</TABLE></DEFINITION>
</NETWORK></BIF>""")
    g = lib.parse_xml_file(str(p).encode())
    assert g, B.last_error()
    G = g.contents
    assert G.current_num_vertices == 2 and G.current_num_edges == 1  # parent observed: no reverse edge
    assert lib.find_node_by_name(b"a&b", g) == 0
    assert B.view_u32(G.node_num_vars, 2).tolist() == [2, 2]  # the CDATA outcome counts
    # the TABLE's last text node (after the comment) is the one read, child-state-major:
    # J[parent][child] = TABLE[parent + child * 2]
    J = B.view_f32(G.edges_joint_probabilities, 100).reshape(10, 10)
    assert J[:2, :2].tolist() == np.array([[0.1, 0.9], [0.2, 0.8]], dtype=np.float32).tolist()
    lib.graph_destroy(g)
