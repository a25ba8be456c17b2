"""Writes the golden fixtures of tests/golden/ from the compiled reference.

Run in the development container only (it needs the reference sources, built
by `make -C oracle ref` into oracle/_ref/):

    python tests/golden/make_golden.py [case ...]

Inputs are either the reference's own files (src/benchmark_files), its
example network, or networks produced by this repository's seeded generator
(include/bp_model.h); a generated input is written as XMLBIF and read by the
reference's own XMLBIF parser, and its SHA-256 is stored so the tests can
confirm the generator still produces the same file.  Outputs are what the
reference computes (oracle/ref_harness.c): structure, CSR, levels, tree BP,
loopy BP per iteration, loopy_propagate_until, and the same graph core built
with -Dfloat=double.  Large cases keep SHA-256 digests and samples instead of
full arrays.
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "belief-propagation_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

import bp_ctypes as B  # noqa: E402
from ref_files import read_tagged, write_replay  # noqa: E402,F401

OUT = Path(__file__).resolve().parent
SCRATCH = OUT / "_scratch"
REF = ROOT / "oracle" / "_ref"
BENCH = Path("/root/reference/src/benchmark_files")


def harness(kind: str, path, action: list[str], fp64=False, omp=False, env=None) -> dict:
    exe = REF / ("ref_harness64" if fp64 else "ref_harness_omp" if omp else "ref_harness")
    out = SCRATCH / "out.bin"
    cmd = [str(exe), kind, str(path), *action, str(out)]
    e = dict(os.environ)
    if env:
        e.update(env)
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, env=e)
    return read_tagged(out)


def compact(a: np.ndarray, count: int, width: int) -> np.ndarray:
    return a.reshape(count, -1)[:, :width].copy()


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def file_sha(p) -> str:
    return hashlib.sha256(Path(p).read_bytes()).hexdigest()


def struct_arrays(d: dict) -> dict:
    """Structure as the tests compare it: states [V][S] and tables [E][S][S]
    (S = largest state count; entries beyond a node's / edge's own counts are 0)."""
    keep = ["meta", "node_num_vars", "node_states", "observed", "node_names", "edges_src", "edges_dst",
            "edges_x", "edges_y", "edges_J", "src_node_list", "src_edge_list", "dst_node_list",
            "dst_edge_list", "levels_to_nodes", "num_levels"]
    out = {k: d[k] for k in keep if k in d}
    V, E = int(d["meta"][0]), int(d["meta"][1])
    S = int(d["node_num_vars"].max()) if V else 1
    out["node_states"] = compact(d["node_states"], V, S)
    if "levels_to_nodes" in out:
        # slots [num_levels, V) are never written by the reference (uninitialised)
        lv = out["levels_to_nodes"].copy()
        lv[int(d["num_levels"][0]):V] = 0
        out["levels_to_nodes"] = lv
    out["edges_J"] = d["edges_J"].reshape(E, 10, 10)[:, :S, :S].copy()
    return out


def struct_digest(d: dict) -> dict:
    return {f"sha_{k}": np.array(sha(v)) for k, v in struct_arrays(d).items()}


def write_levels(g, path):
    G = g.contents
    V = G.current_num_vertices
    lv = B.view_u32(G.levels_to_nodes, 2 * V)
    with open(path, "wb") as f:
        f.write(np.array([G.num_levels], dtype=np.uint32).tobytes())
        f.write(lv.tobytes())


def save(name: str, arrays: dict):
    np.savez_compressed(OUT / f"{name}.npz", **arrays)
    print(f"  wrote {name}.npz ({(OUT / f'{name}.npz').stat().st_size} bytes)")


def loopy_case(kind, path, iters, S, V, E, prefix=""):
    d = harness(kind, path, ["loopy", str(iters)])
    out = {}
    out[prefix + "init_prev"] = compact(d["init_prev"], E, S)
    # states after every iteration for small graphs, after 1, 2 and the last otherwise
    st_iters = list(range(1, iters + 1)) if V <= 1000 else sorted({1, min(2, iters), iters})
    out[prefix + "state_iters"] = np.array(st_iters, dtype=np.uint32)
    out[prefix + "states"] = np.stack([compact(d[f"states_{i}"], V, S) for i in st_iters])
    # messages after iterations 1, 2 and the last one (all iterations would bloat the fixtures)
    keep = sorted({1, min(2, iters), iters})
    out[prefix + "msg_iters"] = np.array(keep, dtype=np.uint32)
    out[prefix + "cur"] = np.stack([compact(d[f"cur_{i}"], E, S) for i in keep])
    out[prefix + "prev"] = np.stack([compact(d[f"prev_{i}"], E, S) for i in keep])
    out[prefix + "deltas"] = d["deltas"]
    return out


def small_case(name, kind, path, iters=20, until=(1e-3, 200), tree=True, fp64=None, extra=None, full_struct=True,
               loopy=True):
    """Full arrays: structure, levels, loopy per iteration, until, tree, OpenMP, FP64."""
    print(name)
    d = harness(kind, path, ["struct", "levels"])
    arrays = struct_arrays(d) if full_struct else {**struct_digest(d), "meta": d["meta"]}
    V, E = int(d["meta"][0]), int(d["meta"][1])
    nv = d["node_num_vars"]
    S = int(nv.max()) if V else 1
    square = bool(V == 0 or (nv == nv[0]).all())
    arrays["square"] = np.array(square)
    if extra:
        arrays.update(extra)
    if square and loopy:
        arrays.update(loopy_case(kind, path, iters, S, V, E, "loopy_"))
        u = harness(kind, path, ["until", repr(until[0]), str(until[1])])
        arrays["until_conv"] = np.array(until[0], dtype=np.float32)
        arrays["until_max"] = np.array(until[1], dtype=np.uint32)
        arrays["until_iterations"] = u["iterations"]
        arrays["until_states"] = compact(u["states"], V, S)
        arrays["until_cur"] = compact(u["cur"], E, S)
        arrays["until_prev"] = compact(u["prev"], E, S)
        om = harness(kind, path, ["loopyq", str(iters)], omp=True, env={"OMP_NUM_THREADS": "4"})
        arrays["omp_states"] = compact(om["states"], V, S)
        # the reference's edge-centric schedule (defective: only node 0 is
        # normalised), a few iterations
        ed = harness(kind, path, ["edge", "3"])
        arrays["edge3_states"] = compact(ed["states"], V, S)
        arrays["edge3_cur"] = compact(ed["cur"], E, S)
    if square:
        if tree:
            t = harness(kind, path, ["tree"])
            arrays["tree_states"] = compact(t["states"], V, S)
            arrays["tree_messages"] = compact(t["tree_messages"], E, S)
        if fp64 is not None and loopy:
            r64 = fp64
            d64 = harness("replay" if r64 != "example" else "example", r64 if r64 != "example" else "-",
                          ["loopy", str(iters)], fp64=True)
            st_iters = list(range(1, iters + 1)) if V <= 1000 else sorted({1, min(2, iters), iters})
            arrays["fp64_state_iters"] = np.array(st_iters, dtype=np.uint32)
            arrays["fp64_states"] = np.stack([compact(d64[f"states_{i}"], V, S) for i in st_iters])
            arrays["fp64_cur"] = compact(d64[f"cur_{iters}"], E, S)
            arrays["fp64_prev"] = compact(d64[f"prev_{iters}"], E, S)
            arrays["fp64_deltas"] = d64["deltas"]
        if fp64 is not None and tree:
            r64 = fp64
            t64 = harness("replay" if r64 != "example" else "example", r64 if r64 != "example" else "-",
                          ["tree"], fp64=True)
            arrays["fp64_tree_states"] = compact(t64["states"], V, S)
    save(name, arrays)


def generated_graph(preset: str, **kw):
    bn = B.generate(preset, **kw)
    lib = B.load()
    xml = SCRATCH / f"{preset}_{abs(hash(tuple(sorted(kw.items())))) % 100000}.xml"
    assert lib.bp_bn_write_xmlbif(bn, str(xml).encode()) == 0
    g = B.bn_to_graph(bn, levels=False)
    return bn, g, xml


def gen_case(name, preset, iters, until=(1e-3, 200), tree=True, full_struct=True, loopy=True, **kw):
    bn, g, xml = generated_graph(preset, **kw)
    replay = SCRATCH / f"{name}.replay"
    write_replay(g, replay)
    extra = {"xml_sha256": np.array(file_sha(xml)), "preset": np.array(preset),
             "overrides": np.array(repr(sorted(kw.items())))}
    small_case(name, "xml", xml, iters=iters, until=until, tree=tree, fp64=replay, extra=extra,
               full_struct=full_struct, loopy=loopy)


def large_loopy(name, preset, iters, sample=2000, via="xml", **kw):
    """Digests of a full-size run: structure, final states/messages, deltas, a sample.

    via="replay" rebuilds the graph through the reference's graph API instead
    of its XMLBIF parser (whose libxml2 DOM and leaked XPath contexts do not
    fit in memory at a million nodes); the parser itself is pinned by the
    other cases."""
    print(name)
    bn, g, xml = generated_graph(preset, **kw)
    G = g.contents
    V, E = G.current_num_vertices, G.current_num_edges
    src_kind, src_path = "xml", xml
    if via == "replay":
        src_kind, src_path = "replay", SCRATCH / f"{name}.replay"
        write_replay(g, src_path)
    d = harness(src_kind, src_path, ["struct"])
    arrays = struct_digest(d)
    arrays["xml_sha256"] = np.array(file_sha(xml))
    arrays["preset"] = np.array(preset)
    arrays["overrides"] = np.array(repr(sorted(kw.items())))
    arrays["via"] = np.array(via)
    arrays["meta"] = d["meta"]
    nv = d["node_num_vars"]
    S = int(nv.max())
    square = bool((nv == nv[0]).all())
    arrays["square"] = np.array(square)
    if square:
        q = harness(src_kind, src_path, ["loopyq", str(iters)])
        st = compact(q["states"], V, S)
        arrays["iters"] = np.array(iters)
        arrays["deltas"] = q["deltas"]
        arrays["sha_states"] = np.array(sha(st))
        arrays["sha_cur"] = np.array(sha(compact(q["cur"], E, S)))
        arrays["sha_prev"] = np.array(sha(compact(q["prev"], E, S)))
        rng = np.random.default_rng(7)
        idx = np.sort(rng.choice(V, size=min(sample, V), replace=False)).astype(np.uint32)
        arrays["sample_nodes"] = idx
        arrays["sample_states"] = st[idx]
        replay = SCRATCH / f"{name}.replay"
        if via != "replay":
            write_replay(g, replay)
        q64 = harness("replay", replay, ["loopyq", str(iters)], fp64=True)
        st64 = compact(q64["states"], V, S)
        arrays["fp64_sha_states"] = np.array(sha(st64))
        arrays["fp64_sample_states"] = st64[idx]
        arrays["fp64_deltas"] = q64["deltas"]
        arrays["fp32_vs_fp64_maxabs"] = np.array(np.abs(st.astype(np.float64) - st64).max())
    save(name, arrays)


def large_tree(name, preset, sample=2000, **kw):
    """Tree BP on a large generated tree, levels from the reference's own
    init_levels_to_nodes (graph.c:806-827; its membership scan is quadratic,
    about 15 minutes at 1M nodes).  The build's levels are computed too and
    must equal the reference's; both digests are stored."""
    print(name)
    bn, g, xml = generated_graph(preset, **kw)
    lib = B.load()
    lib.set_up_src_nodes_to_edges(g)
    lib.set_up_dest_nodes_to_edges(g)
    lib.init_levels_to_nodes(g)
    G = g.contents
    V, E = G.current_num_vertices, G.current_num_edges
    build_levels = sha(B.view_u32(G.levels_to_nodes, 2 * V).copy())
    replay = SCRATCH / f"{name}.replay"
    write_replay(g, replay)
    t = harness("replay", replay, ["tree", "reflevels"])
    # slots [num_levels, V) are never written by the reference (malloc'd)
    ref_lv = t["levels_to_nodes"].copy()
    ref_lv[int(t["num_levels"][0]):V] = 0
    if build_levels != sha(ref_lv):
        sys.exit(f"{name}: the build's levels differ from the reference's init_levels_to_nodes")
    t["levels_to_nodes"] = ref_lv
    S = 2
    st = compact(t["states"], V, S)
    arrays = {"xml_sha256": np.array(file_sha(xml)), "preset": np.array(preset),
              "overrides": np.array(repr(sorted(kw.items()))),
              "V": np.array(V), "E": np.array(E), "num_levels": t["num_levels"],
              "sha_levels": np.array(sha(t["levels_to_nodes"])),
              "levels_source": np.array("the reference's init_levels_to_nodes (graph.c:806-827)"),
              "sha_states": np.array(sha(st)),
              "sha_tree_messages": np.array(sha(compact(t["tree_messages"], E, S)))}
    rng = np.random.default_rng(11)
    idx = np.sort(rng.choice(V, size=min(sample, V), replace=False)).astype(np.uint32)
    arrays["sample_nodes"] = idx
    arrays["sample_states"] = st[idx]
    # the -Dfloat=double graph core on the same tree and levels
    lv = SCRATCH / f"{name}.levels"
    write_levels(g, lv)
    t64 = harness("replay", replay, ["tree", str(lv)], fp64=True)
    st64 = compact(t64["states"], V, S)
    arrays["fp64_sha_states"] = np.array(sha(st64))
    arrays["fp64_sample_states"] = st64[idx]
    arrays["fp32_vs_fp64_maxabs"] = np.array(np.abs(st.astype(np.float64) - st64).max())
    save(name, arrays)


def bif_cases():
    files = ["dog.bif", "small/asia.bif", "small/cancer.bif", "small/earthquake.bif", "small/sachs.bif",
             "small/survey.bif", "medium/alarm.bif", "medium/child.bif", "medium/insurance.bif",
             "medium/water.bif", "medium/hailfinder.bif", "large/hepar2.bif", "large/win95pts.bif",
             "alarm.bif", "insurance.bif", "Water.bif", "hailfinder.bif", "carpo.bif"]
    for rel in files:
        p = BENCH / rel
        name = "bif_" + rel.replace("/", "_").replace(".bif", "").lower()
        try:
            d = harness("bif", p, ["struct", "levels"])
        except subprocess.CalledProcessError:
            print(f"{name}: the reference rejects it (skipped)")
            continue
        E = int(d["meta"][1])
        nv = d["node_num_vars"]
        square = bool((nv == nv[0]).all())
        if E <= 1500 and square:
            small_case(name, "bif", p, iters=20, tree=True,
                       extra={"source": np.array(rel), "file_sha256": np.array(file_sha(p))})
        elif E <= 3000:
            print(name)
            arrays = struct_arrays(d)
            arrays["square"] = np.array(square)
            arrays["source"] = np.array(rel)
            arrays["file_sha256"] = np.array(file_sha(p))
            save(name, arrays)
        else:
            print(name)
            arrays = struct_digest(d)
            arrays["meta"] = d["meta"]
            arrays["source"] = np.array(rel)
            arrays["file_sha256"] = np.array(file_sha(p))
            save(name, arrays)


CASES = {
    "example": lambda: small_case("example", "example", "-", iters=10, until=(1e-9, 10000), fp64="example"),
    "dog_xml": lambda: small_case("dog_xml", "xml", BENCH / "dog.xml",
                                  extra={"source": np.array("dog.xml")}),
    "bf_1000_2000_1": lambda: small_case("bf_1000_2000_1", "xml", BENCH / "xml/bf_1000_2000_1.xml",
                                         extra={"source": np.array("xml/bf_1000_2000_1.xml")}),
    "bf_structs": lambda: [save("struct_" + p.stem, {**struct_arrays(harness("xml", p, ["struct", "levels"])),
                                                      "source": np.array("xml/" + p.name)})
                           for p in sorted((BENCH / "xml").glob("bf_*.xml")) if p.stem != "bf_1000_2000_1"],
    "c1": lambda: gen_case("c1", "c1", 20),
    "c1_mixed": lambda: gen_case("c1_mixed", "c1", 20, states_hi=4, seed=1002),
    "lattice16": lambda: gen_case("lattice16", "c3", 50, rows=16, cols=16),
    "tree10k": lambda: gen_case("tree10k", "tree", 5, full_struct=False, loopy=False, num_nodes=10000),
    "c5_small": lambda: gen_case("c5_small", "c5", 30, full_struct=False, num_nodes=2000),
    "c4net_small": lambda: gen_case("c4net_small", "c4net", 30, full_struct=False, num_nodes=1500,
                                    states_hi=2, states_lo=2),
    "bif": bif_cases,
    "c2b": lambda: large_loopy("c2b", "c2b", 20),
    "c2": lambda: large_loopy("c2", "c2", 20),
    "c3": lambda: large_loopy("c3", "c3", 50, via="replay"),
    "tree1m": lambda: large_tree("tree1m", "tree"),
}


def main(argv):
    SCRATCH.mkdir(exist_ok=True)
    if not (REF / "ref_harness").exists():
        sys.exit("oracle/_ref is not built: make -C oracle ref (needs the reference sources)")
    names = argv or list(CASES)
    for n in names:
        CASES[n]()


if __name__ == "__main__":
    main(sys.argv[1:])
