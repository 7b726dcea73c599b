"""HBS compression of the slab Schur complements on the GPU
(src_solver/hbs_reconstruction.py and slablu_solver.py:1139-1308 of the
reference): the batched Householder bases it rests on against LAPACK, exact recovery
of matrices that are HBS by construction (dense reference formed on the host
from the random factors), and the solver with hbs_rank against its exact
formation of T."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _hip():
    from slablu_amd import hip
    return hip


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("b,m,k", [(3, 1, 1), (2, 37, 20), (4, 250, 103), (2, 64, 64)])
def test_orth_against_lapack_qr(dtype, b, m, k):
    """Householder bases: orthonormal columns, the same span as LAPACK's QR
    (Q Q^T A = A), and the same Q up to column signs."""
    hip = _hip()
    g = torch.Generator().manual_seed(m * 7 + k)
    A0 = torch.randn(b, m, k, generator=g, dtype=torch.float64)
    Q = hip.orth(A0.to("cuda", dtype).contiguous()).double().cpu()
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    I = torch.eye(k, dtype=torch.float64)
    assert float((Q.transpose(1, 2) @ Q - I).abs().max()) < 10 * tol
    assert float((Q @ (Q.transpose(1, 2) @ A0) - A0).norm() / A0.norm()) < 10 * tol
    Qr = torch.linalg.qr(A0)[0]
    sign = torch.sign((Q * Qr).sum(dim=1, keepdim=True))
    assert float((Q * sign - Qr).abs().max()) < 1e3 * tol


def test_orth_of_rank_deficient_blocks():
    """A zero block and a rank-2 block still get k orthonormal columns, the
    rank-2 range among them."""
    hip = _hip()
    g = torch.Generator().manual_seed(3)
    A0 = torch.zeros(2, 30, 6, dtype=torch.float64)
    A0[1] = torch.randn(30, 2, generator=g, dtype=torch.float64) @ torch.randn(2, 6, generator=g,
                                                                                dtype=torch.float64)
    Q = hip.orth(A0.cuda().contiguous()).cpu()
    assert torch.isfinite(Q).all()
    assert float((Q.transpose(1, 2) @ Q - torch.eye(6, dtype=torch.float64)).abs().max()) < 1e-12
    assert float((Q[1] @ (Q[1].T @ A0[1]) - A0[1]).norm()) < 1e-12 * float(A0[1].norm())


def _random_hbs(b, m, bs, k, seed):
    """Dense (b, m bs, m bs) matrices that are HBS with rank k on the
    binary tree of m leaves: random orthonormal bases and diagonal blocks
    on every node, assembled top-down on the host."""
    rng = np.random.default_rng(seed)
    L = int(np.log2(m))
    out = np.empty((b, m * bs, m * bs))
    for t in range(b):
        A = rng.standard_normal((2 * k, 2 * k))                       # root
        for lev in range(L - 1, -1, -1):
            nodes = m >> lev
            rows = bs if lev == 0 else 2 * k
            U = [np.linalg.qr(rng.standard_normal((rows, k)))[0] for _ in range(nodes)]
            V = [np.linalg.qr(rng.standard_normal((rows, k)))[0] for _ in range(nodes)]
            Ub = np.zeros((nodes * rows, nodes * k))
            Vb = np.zeros((nodes * rows, nodes * k))
            Db = np.zeros((nodes * rows, nodes * rows))
            for i in range(nodes):
                Ub[i * rows:(i + 1) * rows, i * k:(i + 1) * k] = U[i]
                Vb[i * rows:(i + 1) * rows, i * k:(i + 1) * k] = V[i]
                Db[i * rows:(i + 1) * rows, i * rows:(i + 1) * rows] = rng.standard_normal((rows, rows))
            A = Ub @ A @ Vb.T + Db
        out[t] = A
    return out


@pytest.mark.parametrize("b,m,bs,k", [(3, 8, 16, 8), (2, 4, 24, 12), (1, 16, 20, 10)])
def test_hbs_exact_recovery(b, m, bs, k):
    """A matrix that is HBS with ranks k is recovered exactly (to rounding)
    from bs + k samples per side: dense form, products and transposed
    products against the matrix itself."""
    from slablu_amd.src_solver.hbs_reconstruction import HBS_matrix
    A = _random_hbs(b, m, bs, k, seed=b * 1000 + m)
    n = m * bs
    hbs = HBS_matrix(b, n, k)
    assert (hbs.bs, hbs.m, hbs.n_pad) == (bs, m, 0)
    ns = bs + k
    rng = np.random.default_rng(7)
    Om = rng.random((n, ns))
    Ps = rng.random((n, ns))
    Ad = torch.tensor(A, device="cuda")
    Omega = torch.tensor(Om, device="cuda")[None]
    Psi = torch.tensor(Ps, device="cuda")[None]
    Y = Ad @ Omega
    Z = Ad.transpose(1, 2) @ Psi
    hbs.reconstruct(Omega, Psi, Y, Z)
    D = hbs.todense().cpu().numpy()
    err = np.linalg.norm(D - A) / np.linalg.norm(A)
    print("case hbs exact recovery", (b, m, bs, k), "rel err", err, "nbytes", hbs.nbytes(), flush=True)
    assert err < 1e-9
    q = torch.tensor(rng.standard_normal((b, n, 3)), device="cuda")
    assert float(torch.linalg.norm(hbs.matvec(q) - Ad @ q) / torch.linalg.norm(Ad @ q)) < 1e-9
    assert float(torch.linalg.norm(hbs.matvec(q, transpose=True) - Ad.transpose(1, 2) @ q)
                 / torch.linalg.norm(Ad.transpose(1, 2) @ q)) < 1e-9
    # matrices b_start..b_end-1 alone
    if b > 1:
        assert float(torch.linalg.norm(hbs.matvec(q[1:], b_start=1) - Ad[1:] @ q[1:])) < 1e-9 * float(
            torch.linalg.norm(Ad[1:] @ q[1:]))


def _solver(n, p, ppw, precision, hbs_rank, bufc=1.0):
    from slablu_amd.src_disc import pdo, built_in_funcs as bif
    from slablu_amd.src_solver.slablu_solver import Domain_Solver
    from oracle import hps_oracle as O
    kh = O.kh_from_ppw(n, ppw)
    op = pdo.PDO_2d(pdo.ones, pdo.ones, c=lambda xx: bif.bfield_constant(xx, kh))
    a = 1 / (2 * (n / (p - 2)))
    dom = Domain_Solver(torch.tensor([[0, 1.0], [0, 1.0]]), op, kh, a, p=p, buf_constant=bufc,
                        precision=precision, hbs_rank=hbs_rank)
    uu = lambda xx: bif.uu_dir_func_greens(xx, kh, center=(1.5, 1.5))
    return dom, uu


def test_hbs_slab_schur_complements_against_exact():
    """n = 256, p = 10 (32 x 32 leaves, 256 interface points): the HBS
    recovery of every slab Schur complement against the same operator
    applied to the identity, and the error falls with the rank."""
    errs = []
    for rank in (24, 40):
        dom, uu = _solver(256, 10, 10, "fp64", rank)
        dom.build(verbose=False)
        Np, L = dom.Npan, dom.Npan_loc * 8
        eye = torch.zeros(Np, 2 * L, 2 * L, dtype=torch.float64, device="cuda")
        eye[:] = torch.eye(2 * L, dtype=torch.float64, device="cuda")
        K = dom._slab_schur_apply(eye)
        Kh = dom._Kf_hbs()
        err = float(torch.linalg.norm(Kh - K) / torch.linalg.norm(K))
        print("case hbs slab schur rank", rank, "blocks", dom.Tslab_hbs.m, "x", dom.Tslab_hbs.bs,
              "rel err", err, flush=True)
        errs.append(err)
    assert errs[1] < errs[0] and errs[1] < 1e-4


def test_hbs_mixed_refined_matches_exact_fp64():
    """Mixed precision with HBS-compressed T_slab: the FP64-residual FGMRES
    still converges to the exact FP64 solution (the compression only
    weakens the preconditioner)."""
    ref, uu = _solver(256, 10, 10, "fp64", None)
    ref.build(verbose=False)
    sol_ref = ref.solve(uu)[0]
    dom, uu = _solver(256, 10, 10, "mixed", 32)
    dom.build(verbose=False)
    sol = dom.solve(uu)[0]
    err = float(torch.linalg.norm(sol - sol_ref) / torch.linalg.norm(sol_ref))
    print("case hbs mixed rank 32: vs exact fp64", err, "steps", len(dom.refine_history), flush=True)
    assert dom.refine_converged
    assert err < 1e-8


def test_hbs_rank_rules():
    dom, _ = _solver(256, 10, 10, "fp64", "auto")
    f = dom.hps.p - 2
    nw = (dom.kh / (2 * np.pi)) * (dom.buf_pans * f / (dom.Npan_loc * f))
    assert dom.hbs_rank_used() == int(4 * nw) + int(1.5 * dom.hps.p) + 30
    with pytest.raises(ValueError):
        _solver(256, 10, 10, "fp64", 0)
    big, _ = _solver(256, 10, 10, "fp64", 200)
    with pytest.raises(ValueError, match="samples"):
        big.build(verbose=False)
