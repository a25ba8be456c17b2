"""The differentiable form on the GPU (include/bp_diff.h through
belief_network_wrapper), against the plain path and the float64 oracle.

- tables rebuilt on the device from the CPTs are bit-identical to the
  reader's (so to the reference's);
- the unrolled forward's beliefs are bit-identical to bp_loopy_iterate's;
- the adjoint's gradient does not depend on the checkpoint interval (bit for
  bit), matches the float64 oracle's gradient (oracle/bp_diff_oracle.py,
  itself pinned to finite differences) to the f32 rounding of the recursion,
  and matches central finite differences of the float64 loss along random
  directions;
- the autograd Function composes with torch (softmax of log-parameters,
  cross-entropy) and its parameter gradient matches the oracle's.
The loss floors beliefs at 1e-30 inside the log: the reference's beliefs
accumulate multiplicatively every iteration (graph.c:1008-1012), and after
30 iterations a few percent of them fall below the f32 range.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import bp_ctypes as B
import bp_diff_oracle as D
from belief_network_wrapper import DifferentiableNetwork

pytestmark = pytest.mark.gpu

FLOOR = 1e-30
# gradient tolerances against the float64 oracle per unrolled depth K
# (measured on the 3000-node network: relative l2 1.8e-7, 1.1e-6, 2.7e-4 and
# largest entry error over largest entry 8.0e-8, 9.1e-7, 2.8e-4 at K = 1, 6, 30)
GRAD_REL = {1: 1e-6, 6: 5e-6, 30: 1e-3}
GRAD_MAX = {1: 5e-7, 6: 5e-6, 30: 1e-3}


def setup(preset="c5", **kw):
    bn = B.generate(preset, **kw)
    A = B.bn_arrays(bn)
    g = B.bn_to_graph(bn)
    G = B.graph_arrays(g)
    net = DifferentiableNetwork(bn, device=0)
    return bn, A, g, G, net


def queries(G, S, n, seed=11):
    rng = np.random.default_rng(seed)
    cand = np.nonzero(G["observed"] == 0)[0]
    q = rng.choice(cand, min(n, len(cand)), replace=False)
    return q, rng.integers(0, S, len(q))


def ce(b, q, t):
    return -torch.log(torch.clamp(b[q, t], min=FLOOR)).mean()


def teardown(lib, bn, g, net):
    net.close()
    lib.graph_destroy(g)
    lib.bp_bn_free(bn)


@pytest.mark.parametrize("preset,kw", [("c5", dict(num_nodes=3000)), ("c1", dict())])
def test_tables_and_forward_bitwise(lib, preset, kw):
    bn, A, g, G, net = setup(preset, **kw)
    print(f"{preset} V={net.V} E={net.E} params={net.num_params}", flush=True)
    P = torch.tensor(A["table"][: net.num_params], device="cuda")
    K = 30
    b = net.beliefs(P, K, 5)
    torch.cuda.synchronize()
    E, S = net.E, net.S
    assert np.array_equal(net.tables_host().view(np.uint32),
                          G["edges_J"].reshape(E, 10, 10)[:, :S, :S].copy().view(np.uint32))
    dev = B.Device(g, 0)
    dev.init()
    dev.iterate(K)
    dev.download(g)
    ref = B.states(g)[:, :S]
    assert np.array_equal(b.cpu().numpy().view(np.uint32), ref.copy().view(np.uint32))
    dev.free()
    teardown(lib, bn, g, net)


def test_gradient_independent_of_checkpoint_interval(lib):
    bn, A, g, G, net = setup("c5", num_nodes=3000)
    P = torch.tensor(A["table"][: net.num_params], device="cuda")
    q, t = queries(G, net.S, 300)
    K = 17
    grads = []
    for c in (1, 3, 5, 17, 40):
        p = P.clone().requires_grad_(True)
        loss = ce(net.beliefs(p, K, c), torch.as_tensor(q, device="cuda"), torch.as_tensor(t, device="cuda"))
        loss.backward()
        grads.append(p.grad.cpu().numpy())
    for gr in grads[1:]:
        assert np.array_equal(gr.view(np.uint32), grads[0].view(np.uint32))
    teardown(lib, bn, g, net)


@pytest.mark.parametrize("K", [0, 1, 6, 30])
def test_gradient_matches_fp64_oracle(lib, K):
    bn, A, g, G, net = setup("c5", num_nodes=3000)
    st = D.DiffStructure(A, G, len(A["num_states"]))
    Pn = A["table"][: net.num_params]
    q, t = queries(G, net.S, 1000)
    p = torch.tensor(Pn, device="cuda", requires_grad=True)
    loss = ce(net.beliefs(p, K, 5), torch.as_tensor(q, device="cuda"), torch.as_tensor(t, device="cuda"))
    loss.backward()
    g32 = p.grad.double().cpu().numpy()
    # float64 oracle with the same floor
    p64 = torch.tensor(Pn.astype(np.float64), requires_grad=True)
    b64 = D.unrolled(st, st.tables(p64), st.initial(p64), K)
    l64 = -torch.log(torch.clamp(b64[q, t], min=FLOOR)).mean()
    l64.backward()
    g64 = p64.grad.numpy()
    assert abs(float(loss.detach()) - float(l64.detach())) <= 1e-4 * max(1.0, abs(float(l64.detach())))
    if K == 0:
        # no iteration: the unobserved query nodes keep their default
        # belief (1.0 per state), which no parameter reaches
        assert not g64.any() and not g32.any()
        teardown(lib, bn, g, net)
        return
    rel = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    cos = g32 @ g64 / (np.linalg.norm(g32) * np.linalg.norm(g64))
    worst = np.abs(g32 - g64).max() / np.abs(g64).max()
    print(f"K={K}: loss {float(loss.detach()):.6f} vs {float(l64.detach()):.6f}; gradient rel l2 {rel:.3e}, "
          f"max entry error / max entry {worst:.3e}, cosine {cos:.12f}")
    # a few times what the f32 recursion itself departs from float64: the
    # error grows with the unrolled depth through the multiplicative belief
    # accumulation
    assert rel < GRAD_REL[K]
    assert worst < GRAD_MAX[K]
    assert cos > 1 - 1e-6
    teardown(lib, bn, g, net)


def test_gradient_matches_finite_differences(lib):
    bn, A, g, G, net = setup("c5", num_nodes=3000)
    st = D.DiffStructure(A, G, len(A["num_states"]))
    Pn = A["table"][: net.num_params]
    q, t = queries(G, net.S, 500, seed=3)
    K = 10
    p = torch.tensor(Pn, device="cuda", requires_grad=True)
    ce(net.beliefs(p, K, 5), torch.as_tensor(q, device="cuda"), torch.as_tensor(t, device="cuda")).backward()
    gr = p.grad.double().cpu().numpy()
    rng = np.random.default_rng(9)
    P64 = Pn.astype(np.float64)

    def loss64(x):
        with torch.no_grad():
            xt = torch.tensor(x)
            b = D.unrolled(st, st.tables(xt), st.initial(xt), K)
            return float(-torch.log(torch.clamp(b[q, t], min=FLOOR)).mean())

    for _ in range(3):
        v = rng.standard_normal(len(P64))
        eps = 1e-6
        fd = (loss64(P64 + eps * v) - loss64(P64 - eps * v)) / (2 * eps)
        print(f"directional derivative: fd {fd:.8f}, adjoint {gr @ v:.8f}")
        assert abs(fd - gr @ v) <= 2e-3 * max(1.0, abs(fd))
    teardown(lib, bn, g, net)


def test_autograd_through_softmax_parameters(lib):
    bn, A, g, G, net = setup("c5", num_nodes=2000)
    st = D.DiffStructure(A, G, len(A["num_states"]))
    gen = torch.Generator().manual_seed(4)
    theta = torch.randn(net.num_params, generator=gen, dtype=torch.float32)
    q, t = queries(G, net.S, 400, seed=4)
    K = 8
    th = theta.cuda().requires_grad_(True)
    loss = ce(net.beliefs(net.softmax_tables(th), K, 5), torch.as_tensor(q, device="cuda"),
              torch.as_tensor(t, device="cuda"))
    loss.backward()
    g32 = th.grad.double().cpu().numpy()
    th64 = theta.double().requires_grad_(True)
    rows = net._rows.cpu()
    inv = net._inv.cpu()
    p64 = torch.softmax(th64[rows], dim=1).reshape(-1)[inv]
    b64 = D.unrolled(st, st.tables(p64), st.initial(p64), K)
    (-torch.log(torch.clamp(b64[q, t], min=FLOOR)).mean()).backward()
    g64 = th64.grad.numpy()
    rel = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print(f"theta gradient rel l2 {rel:.3e}")
    assert rel < 2e-3
    teardown(lib, bn, g, net)


def test_fused_softmax_matches_torch_softmax(lib):
    bn, A, g, G, net = setup("c5", num_nodes=2000)
    gen = torch.Generator().manual_seed(8)
    theta = torch.randn(net.num_params, generator=gen, dtype=torch.float32).cuda()
    q, t = queries(G, net.S, 300, seed=8)
    qd, td = torch.as_tensor(q, device="cuda"), torch.as_tensor(t, device="cuda")
    K = 9
    th1 = theta.clone().requires_grad_(True)
    b1 = net.beliefs(net.softmax_tables(th1), K, 3)
    ce(b1, qd, td).backward()
    J1 = net.tables_host().astype(np.float64)
    th2 = theta.clone().requires_grad_(True)
    b2 = net.beliefs_from_log(th2, K, 3)
    ce(b2, qd, td).backward()
    J2 = net.tables_host().astype(np.float64)
    # expf and the row sums may round differently from torch's softmax: the
    # tables agree to f32 rounding, and the recursion carries that along
    assert np.abs(J1 - J2).max() <= 1e-5 * np.abs(J1).max()
    assert float((b1 - b2).abs().max()) < 1e-4
    g1, g2 = th1.grad.double(), th2.grad.double()
    rel = float((g1 - g2).norm() / g1.norm())
    print(f"fused softmax: beliefs max |diff| {float((b1 - b2).abs().max()):.2e}, gradient rel l2 {rel:.2e}")
    assert rel < 1e-3
    teardown(lib, bn, g, net)


def test_full_scale_c5(lib):
    """200k nodes, K = 30, checkpoints every 5: gradient against the float64
    oracle by directional finite differences; forward/backward times."""
    bn, A, g, G, net = setup("c5")
    print(f"c5: V={net.V} E={net.E} params={net.num_params}", flush=True)
    st = D.DiffStructure(A, G, len(A["num_states"]))
    gen = torch.Generator().manual_seed(6006)
    theta = torch.randn(net.num_params, generator=gen, dtype=torch.float32)
    q, t = queries(G, net.S, 1000, seed=6006)
    K = 30
    th = theta.cuda().requires_grad_(True)
    qd, td = torch.as_tensor(q, device="cuda"), torch.as_tensor(t, device="cuda")
    loss = ce(net.beliefs_from_log(th, K, 5), qd, td)
    loss.backward()
    gr = th.grad.double().cpu().numpy()
    rows, inv = net._rows.cpu(), net._inv.cpu()
    th64 = theta.double()

    def loss64(x):
        with torch.no_grad():
            p = torch.softmax(x[rows], dim=1).reshape(-1)[inv]
            b = D.unrolled(st, st.tables(p), st.initial(p), K)
            return float(-torch.log(torch.clamp(b[q, t], min=FLOOR)).mean())

    l64 = loss64(th64)
    print(f"loss f32 {float(loss.detach()):.6f} fp64 {l64:.6f}", flush=True)
    # every belief after K iterations against the float64 recursion
    with torch.no_grad():
        p64 = torch.softmax(th64[rows], dim=1).reshape(-1)[inv]
        b64 = D.unrolled(st, st.tables(p64), st.initial(p64), K).numpy()
        b32 = net.beliefs_from_log(theta.cuda(), K, 5).double().cpu().numpy()
    fp64_maxabs = float(np.abs(b32 - b64).max())
    print(f"c5 beliefs after K={K}: max |f32 - fp64| over all {net.V} nodes {fp64_maxabs:.3e}", flush=True)
    assert fp64_maxabs < 3e-3  # measured 7.0e-4
    assert abs(float(loss.detach()) - l64) <= 1e-3 * max(1.0, abs(l64))
    rng = np.random.default_rng(1)
    for _ in range(2):
        v = torch.tensor(rng.standard_normal(net.num_params))
        eps = 1e-5
        fd = (loss64(th64 + eps * v) - loss64(th64 - eps * v)) / (2 * eps)
        print(f"directional derivative: fd {fd:.8f}, adjoint {gr @ v.numpy():.8f}", flush=True)
        assert abs(fd - gr @ v.numpy()) <= 5e-3 * max(1.0, abs(fd))
    teardown(lib, bn, g, net)
