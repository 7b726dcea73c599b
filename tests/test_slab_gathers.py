"""CPU checks of the block gathers that turn the leaf DtN stack into the
slab-level block structures, against dense submatrices of the oracle's
assembled face operator A (no GPU needed: pure index work)."""
import numpy as np
import pytest
import torch

from oracle import hps_oracle as O


def _setup(n0, n1, p, kh, bufc):
    from slablu_amd.src_disc import pdo
    from slablu_amd.src_solver.slablu_solver import Domain_Solver
    a = 1.0 / (2 * n1)
    box = torch.tensor([[0, n0 * 2 * a], [0, 1.0]])
    op = pdo.PDO_2d(pdo.ones, pdo.ones, c=lambda xx: -(kh ** 2) * torch.ones(xx.shape[0], dtype=xx.dtype))
    dom = Domain_Solver(box, op, kh, a, p=p, buf_constant=bufc, device="cpu")
    hps = O.HPSOracle(O.helmholtz_pdo(kh), box.numpy(), a, p)
    assert (hps.n0, hps.n1) == (dom.hps.n0, dom.hps.n1)
    DtN = hps.dtns()
    A = hps.sparse_mat(DtN).toarray()
    dom.wdt = torch.float64
    dom.Dt7 = dom._dt7(torch.tensor(DtN))
    return dom, hps, A


def _face(dom, s, c, j, vertical):
    n1, buf, f = dom.Npan_loc, dom.buf_pans, dom.hps.p - 2
    fid = (2 * n1 + 1) * (s * buf + c) + (j if vertical else n1 + j)
    return np.arange(f) + fid * f


CASES = [(12, 12, 10, 30.0, 1.0), (18, 12, 10, 25.0, 1.0), (8, 6, 16, 20.0, 2.0)]


@pytest.mark.parametrize("n0,n1,p,kh,bufc", CASES)
def test_index_sets_match_oracle(n0, n1, p, kh, bufc):
    dom, hps, A = _setup(n0, n1, p, kh, bufc)
    part = O.SlabPartition(n0, n1, p, bufc)
    assert (dom.buf_pans, dom.Npan) == (part.buf, part.Npan)
    for name in ("I_L", "I_R", "I_D", "I_U", "I_C", "I_slabC", "I_slabX", "I_Ctot", "I_Xtot"):
        assert np.array_equal(getattr(dom, name).numpy(), getattr(part, name)), name
    assert np.array_equal(dom.hps.I_unique.numpy(), hps.I_unique)
    assert np.array_equal(dom.hps.Iext.numpy(), hps.iext_box())
    assert np.array_equal(dom.hps.xx_active.numpy(), hps.xx_active)
    assert np.array_equal(dom.hps.grid_xx.numpy(), hps.grid_xx)
    assert np.array_equal(dom.hps.H.Interp_mat, hps.Interp_mat)


@pytest.mark.parametrize("n0,n1,p,kh,bufc", CASES)
def test_elim_rem_interface_blocks(n0, n1, p, kh, bufc):
    dom, hps, A = _setup(n0, n1, p, kh, bufc)
    Np, buf, f = dom.Npan, dom.buf_pans, p - 2
    m = buf - 1
    diag, sub, sup = (x.numpy() for x in dom.elim_blocks())
    for s in range(Np):
        for j in range(n1):
            I = np.concatenate([_face(dom, s, c, j, True) for c in range(1, buf)])
            d, sb, sp = O.extract_btri(A[np.ix_(I, I)], m, f)
            k = s * n1 + j
            assert np.array_equal(diag[k], d) and np.array_equal(sub[k], sb) and np.array_equal(sup[k], sp)
    rd, ru, rl = (x.numpy() for x in dom.rem_blocks())
    for s in range(Np):
        I = np.concatenate([np.concatenate([_face(dom, s, c, r, False) for c in range(buf)]) for r in range(1, n1)])
        d, sb, sp = O.extract_btri(A[np.ix_(I, I)], n1 - 1, buf * f)
        assert np.array_equal(rd[s], d) and np.array_equal(rl[s], sb) and np.array_equal(ru[s], sp)
        # A_rr has no couplings outside the block tridiagonal band
        M = A[np.ix_(I, I)].copy()
        B = buf * f
        for i in range(n1 - 1):
            for jb in range(max(0, i - 1), min(n1 - 1, i + 2)):
                M[i * B:(i + 1) * B, jb * B:(jb + 1) * B] = 0
        assert np.abs(M).max() == 0
    Aii = dom.interface_blocks().numpy()
    for s in range(Np - 1):
        I = np.concatenate([_face(dom, s + 1, 0, j, True) for j in range(n1)])
        assert np.array_equal(Aii[s], A[np.ix_(I, I)])


@pytest.mark.parametrize("n0,n1,p,kh,bufc", CASES)
def test_row_couplings(n0, n1, p, kh, bufc):
    dom, hps, A = _setup(n0, n1, p, kh, bufc)
    Np, buf, f = dom.Npan, dom.buf_pans, p - 2
    dom.device = torch.device("cpu")
    dom._couplings()
    for s in range(Np):
        for j in range(n1):
            k = s * n1 + j
            Ie = np.concatenate([_face(dom, s, c, j, True) for c in range(1, buf)])
            Ir = np.concatenate([np.concatenate([_face(dom, s, c, j + sg, False) for c in range(buf)])
                                 for sg in (0, 1)])
            Ix = np.concatenate((_face(dom, s, 0, j, True), _face(dom, s, buf, j, True)))
            assert np.array_equal(dom.cR[k].numpy(), A[np.ix_(Ie, Ir)])
            assert np.array_equal(dom.cC[k].numpy(), A[np.ix_(Ir, Ie)])
            assert np.array_equal(dom.cE[k].numpy(), A[np.ix_(Ie, Ix)])
            assert np.array_equal(dom.cAXe[k].numpy(), A[np.ix_(Ix, Ie)])
            # the leaf row's own couplings of the interfaces with its rem faces
            Xr = A[np.ix_(Ir, Ix)].copy()
            AXr = A[np.ix_(Ix, Ir)].copy()
            assert np.array_equal(dom.cXr[k].numpy(), Xr)
            assert np.array_equal(dom.cAXr[k].numpy(), AXr)


def test_refinement_stop_test_uses_the_geometric_tail():
    """The stop test on the relative face changes c_k: with a contraction
    rho = c_k / c_{k-1} below 0.1 the error left is c_k rho / (1 - rho);
    a slow or first step is judged by c_k itself."""
    import types
    from slablu_amd.src_solver.slablu_solver import Domain_Solver
    s = types.SimpleNamespace(refine_tol=1e-8)
    conv = lambda c, prev: Domain_Solver._converged(s, c, prev)
    assert not conv(2.7e-7, None)                 # no history: c itself
    assert conv(2.7e-7, 1.8e-5)                   # rho 0.015: tail 4.1e-9
    assert not conv(2.7e-7, 1.0e-6)               # rho 0.27: too slow to trust the tail
    assert not conv(5e-7, 2e-5)                   # rho 0.025: tail 1.28e-8 > tol
    assert conv(9e-9, 1e-8)                       # slow but already below tol


def test_refinement_residual_guard():
    """A step passes the stop test only if it still halved the FP64 residual
    (or the residual is at the 1e-12 floor): a stagnating step, whose tiny
    update gives a tiny change, is not taken as converged."""
    from slablu_amd.src_solver.slablu_solver import Domain_Solver
    prog = Domain_Solver._progressed
    assert prog(1e-6, 1e-4, 1.0)                  # contracting
    assert not prog(0.9e-4, 1e-4, 1.0)            # stagnating
    assert not prog(1e-6, None, 1.0)              # no previous step
    assert prog(5e-13, 5e-13, 1.0)                # at the FP64 floor
    assert prog(0.0, None, 0.0)                   # zero right-hand side
