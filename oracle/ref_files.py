"""TEST INFRASTRUCTURE: files exchanged with the reference harness
(oracle/ref_harness.c, built from the reference's own sources into
oracle/_ref/).  Used by the fixture script (tests/golden/make_golden.py) and
by the CPU-baseline leg of bench.py; the library never loads this module.

- write_replay: a graph as the harness's "replay" source, which rebuilds it
  through the reference's graph_add_node / graph_set_node_state /
  graph_add_edge in the original order (ref_harness.c, replay source).
- read_tagged: the tagged binary file the harness writes (ref_harness.c,
  put()): per record a u32 name length, the name, one type byte ('u' u32,
  'f' f32, 'd' f64, 'b' u8), a u64 element count and the raw little-endian
  data.
"""
from __future__ import annotations

import struct

import numpy as np

import bp_ctypes as B

_DT = {"u": np.uint32, "f": np.float32, "d": np.float64, "b": np.uint8}


def read_tagged(path) -> dict:
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        (n,) = struct.unpack_from("<I", data, pos)
        pos += 4
        name = data[pos:pos + n].decode()
        pos += n
        t = chr(data[pos])
        pos += 1
        (count,) = struct.unpack_from("<Q", data, pos)
        pos += 8
        dt = np.dtype(_DT[t])
        out[name] = np.frombuffer(data, dtype=dt, count=count, offset=pos).copy()
        pos += count * dt.itemsize
    return out


def write_replay(g, path) -> None:
    """Header u32 V, E; per node u32 state count, u8 observed, its states
    (f32) and its 50-byte name; per edge u32 src, dst, x, y and its x * y
    table (f32, row-major)."""
    G = g.contents
    V, E = G.current_num_vertices, G.current_num_edges
    nv = B.view_u32(G.node_num_vars, V)
    st = B.view_f32(G.node_states, V * 10).reshape(V, 10)
    obs = B.view_bytes(G.observed_nodes, V)
    names = B.view_bytes(G.node_names, V * 50).reshape(V, 50)
    src, dst = B.view_u32(G.edges_src_index, E), B.view_u32(G.edges_dest_index, E)
    xs, ys = B.view_u32(G.edges_x_dim, E), B.view_u32(G.edges_y_dim, E)
    J = B.view_f32(G.edges_joint_probabilities, E * 100).reshape(E, 10, 10)
    with open(path, "wb") as f:
        f.write(np.array([V, E], dtype=np.uint32).tobytes())
        if V and (nv == nv[0]).all():
            n = int(nv[0])
            rec = np.zeros(V, dtype=[("n", "<u4"), ("obs", "u1"), ("st", "<f4", (n,)), ("name", "u1", (50,))])
            rec["n"], rec["obs"], rec["st"], rec["name"] = nv, obs, st[:, :n], names
            f.write(rec.tobytes())
            erec = np.zeros(E, dtype=[("hdr", "<u4", (4,)), ("J", "<f4", (n * n,))])
            erec["hdr"] = np.stack([src, dst, xs, ys], axis=1)
            erec["J"] = J[:, :n, :n].reshape(E, n * n)
            f.write(erec.tobytes())
            return
        for v in range(V):
            n = int(nv[v])
            f.write(np.array([n], dtype=np.uint32).tobytes())
            f.write(bytes([int(obs[v])]))
            f.write(st[v, :n].astype(np.float32).tobytes())
            f.write(names[v].tobytes())
        for e in range(E):
            x, y = int(xs[e]), int(ys[e])
            f.write(np.array([src[e], dst[e], x, y], dtype=np.uint32).tobytes())
            f.write(J[e, :x, :y].astype(np.float32).tobytes())
