"""Block-tridiagonal chains eliminated piecewise around separator blocks
(SeparatedBlockTridiagonal) against dense solves of the same chains, plain
and transposed, LU and inverse-factor forms, whole batches and sub-ranges;
and the slab solver with separated rem chains against the single-chain
elimination on the same problem."""
import pytest
import torch

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-12, torch.float32: 2e-5}


def _dense(diag, sub, sup):
    b, m, bs, _ = diag.shape
    M = torch.zeros(b, m * bs, m * bs, dtype=torch.float64)
    for i in range(m):
        M[:, i * bs:(i + 1) * bs, i * bs:(i + 1) * bs] = diag[:, i]
        if i < m - 1:
            M[:, (i + 1) * bs:(i + 2) * bs, i * bs:(i + 1) * bs] = sub[:, i]
            M[:, i * bs:(i + 1) * bs, (i + 1) * bs:(i + 2) * bs] = sup[:, i]
    return M


def _rel(a, b):
    a = a.double().cpu()
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("b,m,bs,d", [(3, 49, 20, 7), (2, 19, 24, 4), (2, 13, 100, 5), (1, 12, 16, 3),
                                      (2, 13, 154, 4)])
@pytest.mark.parametrize("inv", [False, True])
@pytest.mark.parametrize("nrhs", [1, 5, 40, 140])
def test_separated_chains_vs_dense(dtype, b, m, bs, d, inv, nrhs):
    from slablu_amd.src_solver.block_tridiag_separated import SeparatedBlockTridiagonal
    print("case separated", dtype, b, m, bs, d, inv, nrhs, flush=True)
    g = torch.Generator().manual_seed(m * 100 + bs + nrhs)
    # shifted past the random part's spectral radius (about sqrt(bs)): the
    # diagonal blocks stay well conditioned, so the FP32 bar measures the
    # elimination, not the conditioning (with a shift of 3, FP32 block
    # elimination of a 154-wide chain already loses 2e-3 on the CPU)
    shift = 3 + bs ** 0.5
    diag0 = torch.randn(b, m, bs, bs, generator=g, dtype=torch.float64) + shift * torch.eye(bs, dtype=torch.float64)
    sub0 = torch.randn(b, m - 1, bs, bs, generator=g, dtype=torch.float64) * 0.3
    sup0 = torch.randn(b, m - 1, bs, bs, generator=g, dtype=torch.float64) * 0.3
    f0 = torch.randn(b, m * bs, nrhs, generator=g, dtype=torch.float64)
    S = SeparatedBlockTridiagonal(b, m, bs, d, inv_factors=inv)
    S.build_sweep(*(t.to("cuda", dtype).contiguous() for t in (diag0, sub0, sup0)))
    M = _dense(diag0, sub0, sup0)
    for tr in (False, True):
        ref = torch.linalg.solve(M.transpose(1, 2) if tr else M, f0)
        x = S.lusolve_sweep(f0.to("cuda", dtype).clone(), transpose=tr)
        torch.cuda.synchronize()
        print("transpose", tr, "error", _rel(x, ref), flush=True)
        assert _rel(x, ref) < 200 * TOL[dtype], tr
        if b > 1:
            # the last b-1 chains alone, from a strided (non-contiguous) view
            buf = torch.zeros(b, (m + 2) * bs, nrhs, dtype=dtype, device="cuda")
            fv = buf[:, bs:(m + 1) * bs]
            fv.copy_(f0.to("cuda", dtype))
            S.lusolve_sweep(fv[1:], b_start=1, b_end=b, transpose=tr)
            torch.cuda.synchronize()
            assert _rel(fv[1:], ref[1:]) < 200 * TOL[dtype], tr
            assert torch.count_nonzero(buf[:, :bs]) == 0 and torch.count_nonzero(buf[:, (m + 1) * bs:]) == 0
    assert int(S.info.abs().max()) == 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_slab_solver_separated_vs_single_chain(precision):
    """The full two-level solve with separated rem chains (49 blocks: 7 pieces
    of 6 around 7 separators at n=1000's shape is too big for a quick test;
    here 20 x 20 leaves at p=22, rem chains of 19 blocks) against one twisted
    chain per slab, and the transposed slab solves."""
    from slablu_amd.src_disc import pdo, built_in_funcs as bif
    from slablu_amd.src_solver.slablu_solver import Domain_Solver
    from slablu_amd.src_solver.block_tridiag_separated import SeparatedBlockTridiagonal
    n, p, ppw = 400, 22, 10
    kh = (int(n / ppw) + 0.03) * 2 * 3.141592653589793 + 1.8
    op = pdo.PDO_2d(pdo.ones, pdo.ones, c=lambda xx: bif.bfield_lshaped_waveguide(xx, kh))
    a = 1 / (2 * (n / (p - 2)))
    uu = lambda xx: bif.uu_dir_func_greens(xx, kh, center=(1.5, 1.5))
    sols = {}
    for sep in (True, False):
        dom = Domain_Solver(torch.tensor([[0, 1.0], [0, 1.0]]), op, kh, a, p=p, buf_constant=1.0,
                            precision=precision)
        dom.separated_rem = sep
        print("case slab solver separated" if sep else "case slab solver single chain", precision, flush=True)
        dom.build(verbose=False)
        assert isinstance(dom.Ared_solver, SeparatedBlockTridiagonal) == sep
        sols[sep] = dom.solve(uu)[0]
        if precision == "fp64":
            res = (dom.test_slab_solve(), dom.test_slab_solve(transpose=True), dom.test_solve_reduced(),
                   dom.test_solve_reduced(transpose=True))
            print("residuals (slab, slab^T, two-level, two-level^T)", res, flush=True)
            assert max(res) < 1e-10
        else:
            assert dom.refine_converged
    err = float(torch.linalg.norm(sols[True] - sols[False]) / torch.linalg.norm(sols[False]))
    print("separated vs single chain", precision, err, flush=True)
    assert err < (1e-10 if precision == "fp64" else 1e-8)


def test_wide_fp64_rem_blocks_vs_oracle():
    """FP64 rem blocks of 154 (11 leaf columns per slab at p = 16), wider than
    the one-launch chain kernels hold: the per-step batched factorization,
    single chain and separated, against the oracle's sparse solve of the same
    problem (22 x 22 leaves, waveguide coefficient, 10 ppw)."""
    import numpy as np
    from oracle import hps_oracle as O
    from slablu_amd import hip
    from slablu_amd.src_disc import pdo, built_in_funcs as bif
    from slablu_amd.src_solver.slablu_solver import Domain_Solver
    n0, p = 22, 16
    n = n0 * (p - 2)
    hps, part, kh = O.setup_problem(n, p, ppw=10, bfield=O.bfield_lshaped_waveguide)
    ref = O.full_solve(hps, part, lambda xx: O.uu_dir_func_greens(xx, kh, center=(1.5, 1.5)), known_sol=False,
                       A=hps.sparse_mat())["sol_tot"]
    op = pdo.PDO_2d(pdo.ones, pdo.ones, c=lambda xx: bif.bfield_lshaped_waveguide(xx, kh))
    uu = lambda xx: bif.uu_dir_func_greens(xx, kh, center=(1.5, 1.5))
    assert not hip.btri_fused(torch.float64, 154)
    for sep in (False, True):
        dom = Domain_Solver(torch.tensor([[0, 1.0], [0, 1.0]]), op, kh, 1 / (2 * n0), p=p, buf_constant=3.1)
        dom.separated_rem = sep
        assert (dom.buf_pans, dom.rem_bs, dom.rem_nblocks) == (11, 154, 21)
        print("case wide fp64 rem blocks, separated" if sep else "case wide fp64 rem blocks, single chain",
              flush=True)
        dom.build(verbose=False)
        from slablu_amd.src_solver.block_tridiag_separated import SeparatedBlockTridiagonal
        assert isinstance(dom.Ared_solver, SeparatedBlockTridiagonal) == sep
        sol = dom.solve(uu)[0].cpu().numpy()
        err = np.linalg.norm(sol - ref) / np.linalg.norm(ref)
        res = dom.test_solve_reduced(nrhs=4)
        print("wide fp64 rem blocks", "separated" if sep else "single chain", "error vs oracle", err,
              "two-level residual", res, flush=True)
        assert err < 1e-9 and res < 1e-10
