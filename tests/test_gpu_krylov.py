"""Batched Krylov products (slab_mdot / slab_maxpy) against FP64 torch
reductions on the host, and the mixed-precision refinement with several
right-hand sides (batched Gram-Schmidt) against single-column solves."""
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nc,k,n", [(1, 1, 5), (3, 5, 10007), (64, 12, 70001), (2, 32, 4096), (2, 40, 4096),
                                    (3, 70, 1001)])
def test_mdot_maxpy(nc, k, n):
    from slablu_amd import hip
    g = torch.Generator().manual_seed(nc * 31 + k)
    # strided views: a basis inside a wider buffer, w inside another
    Vb = torch.randn(nc, k + 2, n + 3, generator=g, dtype=torch.float64)
    wb = torch.randn(nc, n + 5, generator=g, dtype=torch.float64)
    coef = torch.randn(nc, k, generator=g, dtype=torch.float64)
    V, w = Vb[:, 1:k + 1, 2:n + 2], wb[:, 3:n + 3]
    Vd, wd = Vb.cuda()[:, 1:k + 1, 2:n + 2], wb.cuda()[:, 3:n + 3]
    out = hip.mdot(Vd, wd).cpu()
    ref = torch.einsum("cin,cn->ci", V, w)
    assert (out - ref).abs().max() <= 1e-12 * ref.abs().max()
    hip.maxpy(Vd, coef.cuda(), wd, alpha=-0.5)
    ref_w = w - 0.5 * torch.einsum("cin,ci->cn", V, coef)
    torch.cuda.synchronize()
    assert (wd.cpu() - ref_w).abs().max() <= 1e-12 * ref_w.abs().max()
    # the padding around the views is untouched
    wb2 = wb.clone()
    wb2[:, 3:n + 3] = ref_w
    assert (wb.cuda().cpu()[:, :3] == wb2[:, :3]).all()


def test_mixed_multi_rhs_columns_match_single_solves():
    from oracle import hps_oracle as O
    from slablu_amd.src_disc import pdo, built_in_funcs as bif
    from slablu_amd.src_solver.slablu_solver import Domain_Solver
    n, p, ppw = 128, 10, 10
    kh = O.kh_from_ppw(n, ppw)
    op = pdo.PDO_2d(pdo.ones, pdo.ones, c=lambda xx: bif.bfield_lshaped_waveguide(xx, kh))
    a = 1 / (2 * (n / (p - 2)))
    dom = Domain_Solver(torch.tensor([[0, 1.0], [0, 1.0]]), op, kh, a, p=p, buf_constant=0.6, precision="mixed")
    dom.build(verbose=False)
    centers = [(1.3, 0.2), (-0.4, 1.1), (0.5, -0.6), (1.5, 1.5)]
    solm = dom.solve(lambda xx: bif.uu_dir_func_greens_multi(xx, kh, centers))[0]
    assert dom.refine_converged
    for k, c in enumerate(centers):
        s1 = dom.solve(lambda xx: bif.uu_dir_func_greens(xx, kh, center=c))[0]
        err = float(torch.linalg.norm(solm[:, k] - s1[:, 0]) / torch.linalg.norm(s1))
        print("column", k, "vs single solve", err, flush=True)
        assert err < 1e-8
