"""Hand-worked belief propagation on networks whose neighbouring nodes have
different state counts.

The reference leaves this case undefined: a 2-state source's buffer is read
up to the destination's state count (src/graph/graph.c:856 fills n entries,
:899 reads y of them) and a message of x entries is combined over the
destination's n entries (graph.c:103-104 zeroes x of them, :455 reads n).
The rule this library defines is zero-extension: **every read beyond a
state count yields 0.0**, whatever the padded memory holds.

The expected values below are worked out by hand from that rule, in exact
rational arithmetic (fractions.Fraction), step by step for each network.
They do not call the oracle or the library.  The graphs are built through
the reference's graph API, and then every padded slot (states beyond a
node's count, table entries beyond x by y, message entries beyond x) is
filled with 7.0, so an implementation that reads padding, pads with 1.0 or
normalises over the padded row width gets visibly different numbers.

Two networks:

- `chain`: a(2 states) - b(3) - c(3), both directions (loopy BP: init and
  one iteration; graph.c:725-751, 1154-1175, 985-1024);
- `star`: b(3) with leaves a(2), c(3), d(2) (tree BP; graph.c:376-424,
  461-538, 540-596; driver src/c_benchmark/main.c:205-215).
"""
from __future__ import annotations

import ctypes as C
from fractions import Fraction as F

import numpy as np

import bp_ctypes as B

MAX_STATES = 10
POISON = 7.0
# a result of at most a few dozen f32 operations against the exact value
RTOL = 2e-6


def N(v):
    """Normalise by the sequential sum; a sum <= 0 divides by 1 (graph.c:440-447)."""
    s = sum(v, F(0))
    if s <= 0:
        s = F(1)
    return [x / s for x in v]


def times(*vs):
    """Entrywise product of equally long vectors."""
    out = [F(1)] * len(vs[0])
    for v in vs:
        assert len(v) == len(out)
        out = [a * b for a, b in zip(out, v)]
    return out


def Jb(J, b):
    """m[i] = sum_j J[i][j] * b[j]: the message orientation of graph.c:899."""
    return [sum((J[i][j] * b[j] for j in range(len(J[0]))), F(0)) for i in range(len(J))]


h = F(1, 2)
q = F(1, 4)
# tables as the caller passes them (row i: source state, column j)
J_ab = [[F(1), F(2), F(4)], [h, F(1), F(8)]]            # 2 x 3
J_ba = [[F(1), h], [F(2), F(1)], [F(4), F(8)]]          # 3 x 2
J_bc = [[F(1), F(2), h], [q, F(1), F(2)], [F(2), h, F(1)]]
J_cb = [[F(1), q, F(2)], [F(2), F(1), h], [h, F(2), F(1)]]
J_db = [[F(2), F(1), F(1)], [F(1), F(2), F(4)]]          # 2 x 3
J_bd = [[F(2), F(1)], [F(1), F(2)], [F(1), F(4)]]        # 3 x 2

S_a = [q, 3 * q]
S_c = [h, q, q]
S_d = [3 * q, q]


def _build(nodes, edges, levels: bool):
    """nodes: (count, state or None); edges: (src, dst, table).  Builds
    through the graph API, sets up both CSR views (and levels), then fills
    every padded slot with POISON."""
    lib = B.load()
    g = lib.create_graph(len(nodes), len(edges))
    buf = (C.c_float * MAX_STATES)()
    for v, (n, st) in enumerate(nodes):
        name = f"n{v}".encode()
        if st is None:
            lib.graph_add_node(g, n, name)
        else:
            for k in range(n):
                buf[k] = float(st[k])
            lib.graph_add_and_set_node_state(g, n, name, buf)
    tab = (C.c_float * (MAX_STATES * MAX_STATES))()
    for s, d, J in edges:
        for i in range(len(J)):
            for j in range(len(J[0])):
                tab[i * MAX_STATES + j] = float(J[i][j])
        lib.graph_add_edge(g, s, d, len(J), len(J[0]), tab)
    lib.set_up_src_nodes_to_edges(g)
    lib.set_up_dest_nodes_to_edges(g)
    if levels:
        lib.init_levels_to_nodes(g)
    G = g.contents
    V, E = G.current_num_vertices, G.current_num_edges
    nv = B.view_u32(G.node_num_vars, V)
    st = B.view_f32(G.node_states, V * MAX_STATES).reshape(V, MAX_STATES)
    for v in range(V):
        st[v, nv[v]:] = POISON
    xs, ys = B.view_u32(G.edges_x_dim, E), B.view_u32(G.edges_y_dim, E)
    Jv = B.view_f32(G.edges_joint_probabilities, E * MAX_STATES * MAX_STATES).reshape(E, MAX_STATES, MAX_STATES)
    for arr in (G.edges_messages, G.last_edges_messages):
        m = B.view_f32(arr, E * MAX_STATES).reshape(E, MAX_STATES)
        for e in range(E):
            m[e, xs[e]:] = POISON
    for e in range(E):
        Jv[e, xs[e]:, :] = POISON
        Jv[e, :, ys[e]:] = POISON
    return g


def chain_graph():
    """a(2) - b(3) - c(3); edges e0 a->b, e1 b->a, e2 b->c, e3 c->b."""
    nodes = [(2, S_a), (3, None), (3, S_c)]
    edges = [(0, 1, J_ab), (1, 0, J_ba), (1, 2, J_bc), (2, 1, J_cb)]
    return _build(nodes, edges, levels=False)


def chain_expected():
    """Init messages, the messages and beliefs after one loopy iteration."""
    one3 = [F(1)] * 3
    # init_previous_edge: P[e] = N(J_e . states[src]) (graph.c:736-750)
    P0 = N(Jb(J_ab, S_a + [0]))          # a has 2 states; column 2 reads 0
    P1 = N(Jb(J_ba, one3[:2]))           # y = 2: b's third entry is not read
    P2 = N(Jb(J_bc, one3))
    P3 = N(Jb(J_cb, S_c))
    assert P0 == [F(2, 3), F(1, 3)] and P1 == [F(1, 11), F(2, 11), F(8, 11)]
    assert P2 == [F(14, 41), F(13, 41), F(14, 41)] and P3 == [F(17, 55), F(22, 55), F(16, 55)]
    # one iteration, send phase: buf = state (x) previous in-messages
    buf_a = times(S_a, P1[:2])           # a reads the first 2 of P1's 3 entries
    buf_b = times(one3, P0 + [0], P3)    # P0 has 2 entries; k = 2 reads 0
    buf_c = times(S_c, P2)
    assert buf_b[2] == 0
    C0 = N(Jb(J_ab, buf_a + [0]))
    C1 = N(Jb(J_ba, buf_b[:2]))
    C2 = N(Jb(J_bc, buf_b))              # uses buf_b[2] = 0
    C3 = N(Jb(J_cb, buf_c))
    assert C0 == [F(2, 3), F(1, 3)] and C1 == [F(15, 149), F(30, 149), F(104, 149)]
    assert C2 == [F(52, 125), F(61, 375), F(158, 375)]
    # marginalise: states (x)= product of current in-messages, then N
    # (graph.c:1008-1023)
    s_a = N(times(S_a, C1[:2]))
    s_b = N(times(one3, C0 + [0], C3))
    s_c = N(times(S_c, C2))
    assert s_b[2] == 0
    return {"init": [P0, P1, P2, P3], "msgs": [C0, C1, C2, C3], "states": [s_a, s_b, s_c]}


NAN = float("nan")


def chain_graph_inf():
    """chain_graph() with J_ab[0][0] = +inf (a network file can hold an
    infinite weight): a's message to b then has an infinite sum."""
    g = chain_graph()
    G = g.contents
    E = G.current_num_edges
    Jv = B.view_f32(G.edges_joint_probabilities, E * MAX_STATES * MAX_STATES).reshape(E, MAX_STATES, MAX_STATES)
    Jv[0, 0, 0] = np.inf
    return g


def chain_inf_expected():
    """The chain with J_ab[0][0] = +inf, by the same rules and two more:
    N() of a vector with an infinite sum gives NaN for the infinite entry and
    0 for a finite one (IEEE division), and a product skips a NaN factor
    (graph.c:450-459).  The zero extension does not look at a message's
    values: the entry past a NaN message's x still reads 0."""
    one3 = [F(1)] * 3
    P0 = [NAN, F(0)]                     # J_ab . (S_a, 0) = (inf, 7/8)
    P1 = N(Jb(J_ba, one3[:2]))
    P2 = N(Jb(J_bc, one3))
    P3 = N(Jb(J_cb, S_c))
    buf_a = times(S_a, P1[:2])
    buf_c = times(S_c, P2)
    # b's buffer: P0[0] is NaN (skipped), P0[1] = 0, entry 2 reads 0
    buf_b = [P3[0], F(0), F(0)]
    C0 = [NAN, F(0)]                     # J_ab . (buf_a, 0): row 0 infinite again
    C1 = N(Jb(J_ba, buf_b[:2]))
    C2 = N(Jb(J_bc, buf_b))
    C3 = N(Jb(J_cb, buf_c))
    assert C1 == [F(1, 7), F(2, 7), F(4, 7)] and C2 == [F(4, 13), F(1, 13), F(8, 13)]
    s_a = N(times(S_a, C1[:2]))
    s_b = N([C3[0], F(0), F(0)])         # C0[0] skipped, C0[1] = 0, entry 2 reads 0
    s_c = N(times(S_c, C2))
    assert s_b == [F(1), F(0), F(0)] and s_a == [F(1, 7), F(6, 7)]
    return {"init": [P0, P1, P2, P3], "msgs": [C0, C1, C2, C3], "states": [s_a, s_b, s_c]}


def star_graph():
    """b(3) with leaves a(2), c(3), d(2): nodes 0=a, 1=b, 2=c, 3=d; edges
    e0 a->b, e1 b->a, e2 c->b, e3 b->c, e4 d->b, e5 b->d."""
    nodes = [(2, S_a), (3, None), (3, S_c), (2, S_d)]
    edges = [(0, 1, J_ab), (1, 0, J_ba), (2, 1, J_cb), (1, 2, J_bc), (3, 1, J_db), (1, 3, J_bd)]
    return _build(nodes, edges, levels=True)


def star_expected():
    """Tree BP: level 0 = {a, c, d} (in-degree <= 2), level 1 = {b}."""
    # start: level-0 nodes send N(J . state) (graph.c:388-423)
    E0 = N(Jb(J_ab, S_a + [0]))
    E2 = N(Jb(J_cb, S_c))
    E4 = N(Jb(J_db, S_d + [0]))
    assert E4 == [F(7, 12), F(5, 12)]
    # no up-sweep level (2 levels); after reset, b: buf = 1 (x) in-messages,
    # 2-wide messages extended by 0 (graph.c:472-496)
    buf = times([F(1)] * 3, E0 + [0], E2, E4 + [0])
    assert buf == [F(238, 1980), F(110, 1980), F(0)]
    # b sends to its unvisited neighbours a, c, d (graph.c:503-521)
    E1 = N(Jb(J_ba, buf[:2]))
    E3 = N(Jb(J_bc, buf))
    E5 = N(Jb(J_bd, buf[:2]))
    assert E3 == [F(916, 2317), F(339, 2317), F(1062, 2317)]
    # marginalise replaces each state by the product of its in-messages
    # (graph.c:553-585)
    s_a = N(E1[:2])
    s_b = N(times([F(1)] * 3, E0 + [0], E2, E4 + [0]))
    s_c = N(E3)
    s_d = N(E5[:2])
    assert s_a == [F(1, 3), F(2, 3)] and s_b == [F(119, 174), F(55, 174), F(0)]
    return {"msgs": [E0, E1, E2, E3, E4, E5], "states": [s_a, s_b, s_c, s_d], "num_levels": 2,
            "levels": [0, 2, 3, 1]}


def star_graph_inf():
    """star_graph() with J_ab[0][0] = +inf: a's message to b is (NaN, 0)."""
    g = star_graph()
    G = g.contents
    E = G.current_num_edges
    Jv = B.view_f32(G.edges_joint_probabilities, E * MAX_STATES * MAX_STATES).reshape(E, MAX_STATES, MAX_STATES)
    Jv[0, 0, 0] = np.inf
    return g


def star_inf_expected():
    """Tree BP on the star with J_ab[0][0] = +inf, by the rules of
    chain_inf_expected: b's buffer skips E0's NaN entry, multiplies by its 0
    and by the 0 read past its two entries."""
    E0 = [NAN, F(0)]                     # J_ab . (S_a, 0) = (inf, 7/8)
    E2 = N(Jb(J_cb, S_c))
    E4 = N(Jb(J_db, S_d + [0]))
    v = E2[0] * E4[0]
    buf = [v, F(0), F(0)]
    E1 = N(Jb(J_ba, buf[:2]))
    E3 = N(Jb(J_bc, buf))
    E5 = N(Jb(J_bd, buf[:2]))
    assert E1 == [F(1, 7), F(2, 7), F(4, 7)] and E3 == [F(4, 13), F(1, 13), F(8, 13)]
    assert E5 == [F(1, 2), F(1, 4), F(1, 4)]
    s_a = N(E1[:2])
    s_b = N(buf)
    s_c = N(E3)
    s_d = N(E5[:2])
    assert s_b == [F(1), F(0), F(0)] and s_d == [F(2, 3), F(1, 3)]
    return {"msgs": [E0, E1, E2, E3, E4, E5], "states": [s_a, s_b, s_c, s_d], "num_levels": 2,
            "levels": [0, 2, 3, 1]}


def rows(flat_rows: np.ndarray, counts) -> list:
    return [[float(x) for x in flat_rows[i, :n]] for i, n in enumerate(counts)]


def mismatches(got: list, want: list, what: str) -> list[str]:
    """Entries of got (float rows) that differ from the exact rows want by
    more than RTOL relative (an exact 0 must be 0, an expected NaN NaN)."""
    out = []
    for i, (gr, wr) in enumerate(zip(got, want)):
        if len(gr) != len(wr):
            out.append(f"{what}[{i}]: {len(gr)} entries, expected {len(wr)}")
            continue
        for k, (a, b) in enumerate(zip(gr, wr)):
            b = float(b)
            if b != b:
                if a == a:
                    out.append(f"{what}[{i}][{k}] = {a!r}, expected NaN")
                continue
            if not (abs(a - b) <= RTOL * abs(b)):
                out.append(f"{what}[{i}][{k}] = {a!r}, expected {b!r}")
    return out


def check_chain_init(g, exp) -> list[str]:
    G = g.contents
    x = B.view_u32(G.edges_x_dim, G.current_num_edges)
    return mismatches(rows(B.messages(g, "prev"), x), exp["init"], "init message")


def check_chain_iteration(g, exp) -> list[str]:
    """After one iteration: the new messages sit behind previous_edge_messages
    (the pointer swap of graph.c:1179-1182)."""
    G = g.contents
    x = B.view_u32(G.edges_x_dim, G.current_num_edges)
    nv = B.view_u32(G.node_num_vars, G.current_num_vertices)
    return (mismatches(rows(B.messages(g, "prev"), x), exp["msgs"], "message")
            + mismatches(rows(B.states(g), nv), exp["states"], "belief"))


def check_star(g, exp) -> list[str]:
    G = g.contents
    x = B.view_u32(G.edges_x_dim, G.current_num_edges)
    nv = B.view_u32(G.node_num_vars, G.current_num_vertices)
    return (mismatches(rows(B.messages(g, "tree"), x), exp["msgs"], "tree message")
            + mismatches(rows(B.states(g), nv), exp["states"], "belief"))
