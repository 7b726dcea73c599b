"""The hand-written adjoint (torch.autograd.Function) of the FP64 solve with
respect to the per-node coefficient b of -Lap u - kh^2 b u = 0:

  * against the oracle's adjoint gradient (itself pinned by finite
    differences, tests/test_oracle_adjoint.py) on 8 x 8 leaves, p=16, with
    the leaf factors kept and recomputed in chunks (same gradient);
  * the adjoint workload: 34 x 34 leaves, p=16 (N = 259,896), 10 ppw,
    b = 1 + 0.1 * (32 seeded Fourier modes, |k| <= 8), 1024 seeded interior
    observation nodes, data from a perturbed b, leaf factors recomputed
    under a memory budget: the gradient against central finite differences
    (step 1e-6) and the CPU oracle's FP64 gradient, and its cost against
    one forward factor + solve."""
import numpy as np
import pytest
import torch

from oracle import hps_oracle as O

pytestmark = pytest.mark.gpu


def _problem(n0, p, ppw=10, **kw):
    from slablu_amd.src_disc import built_in_funcs as bif
    from slablu_amd.src_solver.adjoint import HelmholtzAdjoint
    n = n0 * (p - 2)
    kh = O.kh_from_ppw(n, ppw)
    a = 1 / (2 * n0)
    uu = lambda xx: bif.uu_dir_func_greens(xx, kh, center=(1.5, 1.5))
    prob = HelmholtzAdjoint(torch.tensor([[0, 1.0], [0, 1.0]]), kh, a, p, uu, buf_constant=1.0, **kw)
    b = bif.fourier_field(prob.xx)
    return prob, b, kh, a


def _grad(prob, b, obs, d):
    from slablu_amd.src_solver.adjoint import observed_loss
    bt = b.clone().cuda().requires_grad_(True)
    u = prob.solve(bt)
    L = observed_loss(u, obs.cuda(), d.cuda())
    L.backward()
    return bt.grad.detach().cpu(), float(L.detach()), u.detach().cpu()


def _oracle_grad(n0, p, kh, a, b, obs, d):
    hps = O.HPSOracle(O.helmholtz_pdo(kh), np.array([[0, 1.0], [0, 1.0]]), a, p)
    part = O.SlabPartition(hps.n0, hps.n1, p, 1.0)
    uu = lambda xx: O.uu_dir_func_greens(xx, kh, center=(1.5, 1.5))
    return O.adjoint_gradient(hps, part, kh, b.numpy(), uu, obs.numpy(), d.numpy())


def test_adjoint_gradient_vs_oracle_small():
    n0, p = 8, 16
    prob, b, kh, a = _problem(n0, p)
    print("case adjoint 8x8 p=16", flush=True)
    obs = prob.observation_nodes(64, seed=0)
    d = 0.01 * torch.randn(64, generator=torch.Generator().manual_seed(1), dtype=torch.float64)
    grad, L, u = _grad(prob, b, obs, d)
    go, Lo, uo = _oracle_grad(n0, p, kh, a, b, obs, d)
    eu = np.linalg.norm(u[:, 0].numpy() - uo) / np.linalg.norm(uo)
    eg = np.abs(grad.numpy() - go).max() / np.abs(go).max()
    print("adjoint small: u err", eu, "loss", L, Lo, "grad err", eg, flush=True)
    assert eu < 1e-10
    assert abs(L - Lo) < 1e-9 * Lo
    assert eg < 1e-8
    # only leaf-interior nodes carry a gradient
    off = np.setdiff1d(np.arange(grad.numel()), prob.interior_nodes().numpy())
    assert torch.count_nonzero(grad[off]) == 0
    # leaf factors recomputed in chunks of 7 leaves instead of kept: the same gradient
    prob2, _, _, _ = _problem(n0, p, leaf_memory_budget=1, leaf_chunk=7)
    grad2, L2, _ = _grad(prob2, b, obs, d)
    assert prob2.dom.hps.leaf["M"] is None and prob2.dom.hps.leaf["chunk"] == 1
    prob3, _, _, _ = _problem(n0, p, leaf_memory_budget=7 * 196 * 252 * 8)
    grad3, L3, _ = _grad(prob3, b, obs, d)
    assert prob3.dom.hps.leaf["M"] is None and prob3.dom.hps.leaf["chunk"] == 7
    for gg, LL in ((grad2, L2), (grad3, L3)):
        assert abs(LL - L) <= 1e-14 * L
        assert (gg - grad).abs().max() <= 1e-13 * grad.abs().max()


def test_adjoint_workload_34x34_p16_recompute():
    n0, p = 34, 16
    # a quarter of the leaf factors fits: recomputed in 4 chunks
    per_leaf = 196 * (196 + 56) * 8
    prob, b, kh, a = _problem(n0, p, leaf_memory_budget=n0 * n0 * per_leaf // 4)
    assert O.dof_count(p, n0, n0) == 259896
    print("case adjoint workload 34x34 p=16", flush=True)
    obs = prob.observation_nodes(1024, seed=0)
    from slablu_amd.src_disc import built_in_funcs as bif
    b_pert = b + 0.02 * (bif.fourier_field(prob.xx, seed=1) - 1.0)
    with torch.no_grad():
        d = prob.solve(b_pert.cuda())[obs.cuda(), 0].cpu()
    assert not prob.dom.hps.leaf["stored"]
    grad, L, u = _grad(prob, b, obs, d)
    # cost: a second forward + backward after the warm-up
    grad_b, _, _ = _grad(prob, b, obs, d)
    tf, tb = prob.timings["forward_s"], prob.timings["backward_s"]
    print("adjoint workload: loss", L, "forward %.1f ms backward %.1f ms ratio %.2f" % (1e3 * tf, 1e3 * tb, tb / tf),
          flush=True)
    assert torch.equal(grad, grad_b)
    assert tb <= 2.0 * tf
    # central finite differences at the four largest gradient entries
    from slablu_amd.src_solver.adjoint import observed_loss
    h = 1e-6
    top = torch.topk(grad.abs(), 4).indices
    for i in top.tolist():
        Ls = []
        for s in (1.0, -1.0):
            bp = b.clone()
            bp[i] += s * h
            with torch.no_grad():
                Ls.append(float(observed_loss(prob.solve(bp.cuda()), obs.cuda(), d.cuda())))
        fd = (Ls[0] - Ls[1]) / (2 * h)
        print("node", i, "adjoint", float(grad[i]), "FD", fd, flush=True)
        assert abs(float(grad[i]) - fd) <= 1e-4 * abs(fd)
    go, Lo, uo = _oracle_grad(n0, p, kh, a, b, obs, d)
    eg = np.abs(grad.numpy() - go).max() / np.abs(go).max()
    print("adjoint workload vs oracle gradient", eg, "loss", L, Lo, flush=True)
    assert eg < 1e-8
    assert abs(L - Lo) < 1e-9 * Lo
