"""TEST INFRASTRUCTURE (the oracle): never imported by the product.

The differentiable form's reference: the reference's unrolled loopy
recursion in float64 on the CPU, differentiated by torch autograd, together
with the map from CPT parameters to what the reference's XMLBIF reader builds
from them.

Recursion, as the reference computes it (src/graph/graph.c):
  init_previous_edge (:725-751)     m_0[e] = N(J_e . st_0[src e])
  first combine buffer (:851-882)   buf_0[v] = st_0[v] (.) prod_{e in in(v)} m_0[e]
  iteration t (:1121-1183)          m_t[e] = N(J_e . buf_{t-1}[src e])
                                    st_t[v] = N(st_{t-1}[v] (.) prod m_t)  (N(st_{t-1}) without in-edges, :1008)
                                    buf_t[v] = st_t[v] (.) prod m_t
with N(x) = x / sum(x), and x unchanged when sum(x) <= 0 (:440-447), and
J_e[i][j], i the source state (:896-903).

Parameter map (src/bnf-xml-parser/xml-expression.c:290-419): a GIVEN-less
definition's TABLE is its node's initial belief (the last one for a node
wins); for a definition with GIVENs, the parents are taken last GIVEN
first, and for each one the forward edge parent -> child gets
J[j][k] = sum of TABLE[k * slice + idx] over the configurations idx in which
that parent is in state j ((idx / offset) % n_parent, offset = product of the
state counts of the GIVENs after it); the reverse edge child -> parent gets
the transpose and exists only when the parent is not observed.

How this oracle is pinned (tests/test_diff_oracle.py):
  - the map, applied to a network's tables, gives the library's parsed
    tables to within the f32 rounding of their sums (the parser is pinned
    bit for bit to the reference);
  - the forward equals the C FP64 oracle (bp_oracle.c, itself bitwise equal
    to the reference built with -Dfloat=double) to 1e-12;
  - the gradient equals central finite differences of the float64 loss.
"""
from __future__ import annotations

import numpy as np
import torch


class DiffStructure:
    """Structure of a network for the differentiable recursion.

    bn: bp_ctypes.bn_arrays(bn); graph: bp_ctypes.graph_arrays(g) of the
    graph the reader built from it; num_param_defs: the leading definitions
    that are parameters.
    """

    def __init__(self, bn: dict, graph: dict, num_param_defs: int):
        ns = bn["num_states"].astype(np.int64)
        V = len(ns)
        S = int(ns.max())
        if int(ns.min()) != S:
            raise ValueError("uniform state counts only")
        src = graph["edges_src"].astype(np.int64)
        dst = graph["edges_dst"].astype(np.int64)
        E = len(src)
        observed = graph["observed"].astype(bool)
        self.V, self.E, self.S = V, E, S
        self.src = torch.from_numpy(src)
        self.dst = torch.from_numpy(dst)

        gb = bn["given_begin"].astype(np.int64)
        gv = bn["given"].astype(np.int64)
        tb = bn["table_begin"].astype(np.int64)
        dfor = bn["def_for"].astype(np.int64)
        npd = num_param_defs
        self.num_params = int(tb[npd])

        # edges of each (definition, parent rank), rank 0 = last GIVEN
        npar = gb[1:npd + 1] - gb[:npd]
        maxr = int(npar.max()) if npd else 0
        tgt, srcp = [], []
        # sequential edge numbering over (definition, rank)
        d_idx = np.repeat(np.arange(npd), npar)
        r_idx = np.concatenate([np.arange(n) for n in npar]) if npd else np.zeros(0, np.int64)
        parent = gv[gb[d_idx] + (npar[d_idx] - 1 - r_idx)]
        has_rev = ~observed[parent]
        counts = 1 + has_rev.astype(np.int64)
        first = np.concatenate([[0], np.cumsum(counts)[:-1]])
        if counts.sum() != E:
            raise ValueError(f"{counts.sum()} edges from the definitions, graph has {E}")
        e_fwd = first
        e_rev = np.where(has_rev, first + 1, -1)
        if not (np.array_equal(src[e_fwd], parent) and np.array_equal(dst[e_fwd], dfor[d_idx])):
            raise ValueError("forward edges do not follow the reader's order")
        rv = e_rev[has_rev]
        if not (np.array_equal(src[rv], dfor[d_idx][has_rev]) and np.array_equal(dst[rv], parent[has_rev])):
            raise ValueError("reverse edges do not follow the reader's order")
        # offset of rank r: product of the state counts of ranks < r
        offset = np.ones(len(d_idx), dtype=np.int64)
        pos = 0
        for d_n in npar:
            o = 1
            for r in range(d_n):
                offset[pos + r] = o
                o *= ns[parent[pos + r]]
            pos += d_n
        slice_ = (tb[1:npd + 1] - tb[:npd]) // ns[dfor[:npd]]
        SS = S * S
        for r in range(maxr):
            sel = np.nonzero(r_idx == r)[0]
            defs = d_idx[sel]
            lens = tb[defs + 1] - tb[defs]
            p = np.repeat(tb[defs], lens) + np.concatenate([np.arange(n) for n in lens])
            rep = np.repeat(np.arange(len(sel)), lens)
            local = p - tb[defs][rep]
            sl = slice_[defs][rep]
            k = local // sl
            idx = local % sl
            npr = ns[parent[sel]][rep]
            j = (idx // offset[sel][rep]) % npr
            tgt.append(e_fwd[sel][rep] * SS + j * S + k)
            srcp.append(p)
            hr = has_rev[sel][rep]
            tgt.append(e_rev[sel][rep][hr] * SS + k[hr] * S + j[hr])
            srcp.append(p[hr])
        self.map_tgt = torch.from_numpy(np.concatenate(tgt)) if tgt else torch.zeros(0, dtype=torch.int64)
        self.map_src = torch.from_numpy(np.concatenate(srcp)) if srcp else torch.zeros(0, dtype=torch.int64)

        # initial beliefs: the last GIVEN-less definition of each node
        state_def = np.full(V, -1, dtype=np.int64)
        for d in range(len(dfor)):
            if gb[d + 1] == gb[d]:
                state_def[dfor[d]] = d
        prior = np.nonzero((state_def >= 0) & (state_def < npd))[0]
        self.prior_nodes = torch.from_numpy(prior)
        self.prior_params = torch.from_numpy(tb[state_def[prior]][:, None] + np.arange(S)[None, :])
        self.base_st0 = torch.from_numpy(graph["node_states"].reshape(V, -1)[:, :S].astype(np.float64))

        # in-edges of each node, padded with an index to a row of ones
        order = np.argsort(dst, kind="stable")
        deg = np.bincount(dst, minlength=V)
        D = max(int(deg.max()) if V else 0, 1)
        pad = np.full((V, D), E, dtype=np.int64)
        starts = np.concatenate([[0], np.cumsum(deg)[:-1]])
        col = np.arange(E) - np.repeat(starts, deg)
        pad[dst[order], col] = order
        self.in_pad = torch.from_numpy(pad)
        self.has_in = torch.from_numpy(deg > 0)[:, None]

    def tables(self, params: torch.Tensor) -> torch.Tensor:
        """[E, S, S] tables from the parameter vector."""
        flat = torch.zeros(self.E * self.S * self.S, dtype=params.dtype)
        flat = flat.index_add(0, self.map_tgt, params[self.map_src])
        return flat.view(self.E, self.S, self.S)

    def initial(self, params: torch.Tensor) -> torch.Tensor:
        st0 = self.base_st0.to(params.dtype).clone()
        st0[self.prior_nodes] = params[self.prior_params]
        return st0


def _normalise(x: torch.Tensor) -> torch.Tensor:
    s = x.sum(-1, keepdim=True)
    return x / torch.where(s > 0, s, torch.ones_like(s))


def unrolled(st: DiffStructure, J: torch.Tensor, st0: torch.Tensor, K: int) -> torch.Tensor:
    """Beliefs after init + K iterations (float64 if the inputs are)."""
    ones = torch.ones(1, st.S, dtype=J.dtype)

    def send(b):
        return _normalise(torch.einsum("eij,ej->ei", J, b[st.src]))

    def prod_in(m):
        return torch.cat([m, ones])[st.in_pad].prod(1)

    m = send(st0)
    T = prod_in(m)
    buf = st0 * T
    s = st0
    for _ in range(K):
        m = send(buf)
        T = prod_in(m)
        s = _normalise(torch.where(st.has_in, s * T, s))
        buf = s * T
    return s


def cross_entropy(beliefs: torch.Tensor, query: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """Mean cross-entropy of the query nodes' beliefs against one-hot targets."""
    return -torch.log(beliefs[query, target]).mean()


def loss_and_grad(st: DiffStructure, params: np.ndarray, K: int, query, target):
    """float64 loss and dL/dparams through the map and the recursion."""
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    b = unrolled(st, st.tables(p), st.initial(p), K)
    loss = cross_entropy(b, torch.as_tensor(query), torch.as_tensor(target))
    loss.backward()
    return float(loss.detach()), p.grad.numpy().copy()


def loss_only(st: DiffStructure, params: np.ndarray, K: int, query, target) -> float:
    with torch.no_grad():
        p = torch.tensor(params, dtype=torch.float64)
        b = unrolled(st, st.tables(p), st.initial(p), K)
        return float(cross_entropy(b, torch.as_tensor(query), torch.as_tensor(target)))


def beliefs(st: DiffStructure, params: np.ndarray, K: int) -> np.ndarray:
    with torch.no_grad():
        p = torch.tensor(params, dtype=torch.float64)
        return unrolled(st, st.tables(p), st.initial(p), K).numpy()
