"""The CSV benchmark driver (belief-propagation_amd/bp_benchmark), the
reference's src/c_benchmark/main.c over the library."""
from __future__ import annotations

import csv
import subprocess
from pathlib import Path

import pytest

import bp_ctypes as B
from conftest import INPUTS, ROOT

DRIVER = ROOT / "belief-propagation_amd" / "bp_benchmark"
HEADER = ["File Name", "Propagation Type", "Number of Nodes", "Number of Edges", "Diameter", "Number of Iterations",
          "BP Run Time(s)"]


def test_driver_is_built_and_prints_usage():
    assert DRIVER.exists(), "run __graft_entry__.build()"
    r = subprocess.run([str(DRIVER), "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "usage" in r.stderr


@pytest.mark.gpu
def test_driver_rows_match_the_library(lib, tmp_path):
    out = tmp_path / "bench.csv"
    files = [str(INPUTS / "dog.xml"), str(INPUTS / "asia.bif")]
    print("driver on", files, flush=True)
    r = subprocess.run([str(DRIVER), "-o", str(out)] + files, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rows = list(csv.reader(out.open()))
    assert rows[0] == HEADER
    body = rows[1:]
    # xml: regular, loopy, loopy-edge; bif: regular, loopy
    assert [row[1] for row in body] == ["regular", "loopy", "loopy-edge", "regular", "loopy"]
    for row in body:
        path = row[0]
        g = lib.parse_xml_file(path.encode()) if path.endswith(".xml") else lib.parse_bif_file(path.encode())
        G = g.contents
        assert int(row[2]) == G.current_num_vertices and int(row[3]) == G.current_num_edges
        if row[1] == "loopy":
            lib.set_up_src_nodes_to_edges(g)
            lib.set_up_dest_nodes_to_edges(g)
            d = B.Device(g, 0)
            d.init()
            assert int(row[5]) == d.until(1e-3, 200)
            d.free()
        assert float(row[6]) >= 0.0
        lib.graph_destroy(g)
