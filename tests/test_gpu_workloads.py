"""The benchmark workloads on the GPU against the CPU oracle.

  32 x 32 leaves, p=10, 10 ppw, buf_constant 0.3 (16 slabs), source at
      (1.5, 1.5): FP64 and mixed against the oracle's exact discrete solution;
  constant-coefficient Helmholtz, p=22, 10 ppw, source at (1.5, 1.5), at
      n=400 (20 x 20 leaves) against the oracle;
  L-shaped waveguide, p=22, 10 ppw, mixed + refined, at n=400 against the
      oracle (whose coefficient is the brute-force lattice sum);
  both p=22 workloads at full size (n=1000, 50 x 50 leaves, N=1,102,000):
      residuals of the full discrete system, mixed against FP64, converged
      refinement, no zero pivots;
plus the failure reports: a singular leaf is counted by build(), and a
refinement cut off before its tolerance is flagged."""
import numpy as np
import pytest
import torch

from oracle import hps_oracle as O

pytestmark = pytest.mark.gpu

SRC = (1.5, 1.5)


def _solver(n, p, ppw, field="constant", precision="fp64", bufc=1.0, center=SRC, **kw):
    from slablu_amd.src_disc import pdo, built_in_funcs as bif
    from slablu_amd.src_solver.slablu_solver import Domain_Solver
    kh = O.kh_from_ppw(n, ppw)
    bf = bif.bfield_constant if field == "constant" else bif.bfield_lshaped_waveguide
    op = pdo.PDO_2d(pdo.ones, pdo.ones, c=lambda xx: bf(xx, kh))
    a = 1 / (2 * (n / (p - 2)))
    dom = Domain_Solver(torch.tensor([[0, 1.0], [0, 1.0]]), op, kh, a, p=p, buf_constant=bufc, precision=precision,
                        **kw)
    uu_dir = lambda xx: bif.uu_dir_func_greens(xx, kh, center=center)
    return dom, uu_dir, kh


def _oracle(n, p, ppw, field="constant", bufc=1.0, center=SRC):
    bf = O.bfield_constant if field == "constant" else O.bfield_lshaped_waveguide
    hps, part, kh = O.setup_problem(n, p, ppw=ppw, bfield=bf, buf_constant=bufc)
    r = O.full_solve(hps, part, lambda xx: O.uu_dir_func_greens(xx, kh, center=center))
    return hps, part, r


def _err(sol, ref):
    return float(np.linalg.norm(sol.numpy() - ref) / np.linalg.norm(ref))


def test_cfg1_32x32_p10_16slabs_vs_oracle():
    n, p, ppw, bufc = 256, 10, 10, 0.3
    hps, part, r = _oracle(n, p, ppw, bufc=bufc)
    assert (part.buf, part.Npan) == (2, 16)
    assert O.dof_count(p, hps.n0, hps.n1) == 82432
    for precision, tol in (("fp64", 1e-10), ("mixed", 1e-8)):
        dom, uu_dir, kh = _solver(n, p, ppw, precision=precision, bufc=bufc)
        print("case cfg1", precision, "Npan", dom.Npan, flush=True)
        assert (dom.buf_pans, dom.Npan, dom.hps.n0, dom.hps.n1) == (2, 16, 32, 32)
        info = dom.build(verbose=False)
        assert all(v == 0 for v in info["zero_pivots"].values())
        sol, res, true_err, resloc, toc = dom.solve(uu_dir, known_sol=True)
        err = _err(sol, r["sol_tot"])
        print("cfg1", precision, "rel err vs oracle", err, "true_err", true_err, r["true_err"], flush=True)
        assert err < tol
        assert abs(true_err - r["true_err"]) < 10 * tol * max(1.0, r["true_err"])
        if precision == "mixed":
            assert dom.refine_converged


def test_cfg2_constant_p22_10ppw_n400_fp64_vs_oracle():
    n, p, ppw = 400, 22, 10
    dom, uu_dir, kh = _solver(n, p, ppw)
    print("case cfg2 n=400", flush=True)
    info = dom.build(verbose=False)
    assert all(v == 0 for v in info["zero_pivots"].values())
    sol, res, true_err, resloc, toc = dom.solve(uu_dir, known_sol=True)
    hps, part, r = _oracle(n, p, ppw)
    err = _err(sol, r["sol_tot"])
    print("cfg2 n=400 fp64 rel err vs oracle", err, "true_err", true_err, r["true_err"], "res", res, flush=True)
    assert err < 1e-10
    assert abs(true_err - r["true_err"]) < 1e-11
    assert res < 1e-11


def test_cfg3_waveguide_p22_10ppw_n400_vs_oracle():
    """The waveguide coefficient on both sides: the product's field on the
    device, the oracle's brute-force lattice sum on the host; FP64 to 1e-10
    and mixed + refined to 1e-8 against the oracle."""
    n, p, ppw = 400, 22, 10
    hps, part, r = _oracle(n, p, ppw, field="waveguide")
    for precision, tol in (("fp64", 1e-10), ("mixed", 1e-8)):
        dom, uu_dir, kh = _solver(n, p, ppw, field="waveguide", precision=precision)
        print("case cfg3 n=400", precision, flush=True)
        dom.build(verbose=False)
        sol, res, true_err, resloc, toc = dom.solve(uu_dir, known_sol=False)
        err = _err(sol, r["sol_tot"])
        print("cfg3 n=400", precision, "rel err vs oracle", err, "refinement", dom.refine_history, flush=True)
        assert err < tol
        if precision == "mixed":
            assert dom.refine_converged and dom.solve_info["refine_converged"]


@pytest.fixture(scope="module")
def full_size_fp64_waveguide():
    dom, uu_dir, kh = _solver(1000, 22, 10, field="waveguide")
    dom.build(verbose=False)
    sol, res, _, resloc, _ = dom.solve(uu_dir)
    out = dict(sol=sol, res=res, resloc=resloc)
    del dom
    torch.cuda.empty_cache()
    return out


def test_full_size_cfg2_constant_fp64_properties():
    """n=1000, p=22, 10 ppw, constant coefficient (N = 1,102,000): reduced
    and leaf residuals of the FP64 solve, the slab self-check, the known
    solution (the reference's own test bound), no zero pivots."""
    dom, uu_dir, kh = _solver(1000, 22, 10)
    print("case full-size cfg2 fp64", flush=True)
    assert O.dof_count(22, dom.hps.n0, dom.hps.n1) == 1102000 and dom.Npan == 10
    info = dom.build(verbose=True)
    assert all(v == 0 for v in info["zero_pivots"].values())
    sol, res, true_err, resloc, toc = dom.solve(uu_dir, known_sol=True)
    slab = dom.test_slab_solve()
    uu_true = uu_dir(dom.hps.xx_tot.clone()).cpu()
    print("full-size cfg2: res", res, "resloc", resloc, "true_err", true_err, "slab self-check", slab, flush=True)
    assert res < 1e-10
    assert resloc < 1e-8 * float(torch.linalg.norm(uu_true)) * kh ** 2
    assert slab < 1e-10
    assert true_err < 2e-5


def test_full_size_cfg3_waveguide_mixed_vs_fp64(full_size_fp64_waveguide):
    """n=1000 waveguide: the mixed + refined solution within 1e-8 of the FP64
    one, the refinement converged by its own test, FP64 residuals small."""
    ref = full_size_fp64_waveguide
    dom, uu_dir, kh = _solver(1000, 22, 10, field="waveguide", precision="mixed")
    print("case full-size cfg3 mixed", flush=True)
    info = dom.build(verbose=True)
    assert all(v == 0 for v in info["zero_pivots"].values())
    sol, res, _, resloc, _ = dom.solve(uu_dir)
    err = float(torch.linalg.norm(sol - ref["sol"]) / torch.linalg.norm(ref["sol"]))
    print("full-size cfg3: mixed vs fp64", err, "refinement", dom.refine_history, dom.refine_residuals,
          "fp64 res", ref["res"], "mixed face residual", res, flush=True)
    assert dom.refine_converged
    assert err < 1e-8
    assert ref["res"] < 1e-10


def test_zero_pivot_of_a_singular_leaf_is_reported():
    """c11 = c22 = 0 and no c inside the first leaf: that leaf's interior
    block is exactly zero.  build() counts exactly one leaf with a zero pivot
    and warns (or raises with check_pivots='raise')."""
    from slablu_amd.src_disc import pdo
    from slablu_amd.src_solver.slablu_solver import Domain_Solver
    n, p = 64, 10
    a = 1 / (2 * (n / (p - 2)))

    def coef(xx):
        xx = torch.as_tensor(xx)
        inside = (xx[:, 0] < 2 * a) & (xx[:, 1] < 2 * a)
        return torch.where(inside, torch.zeros_like(xx[:, 0]), torch.ones_like(xx[:, 0]))

    op = pdo.PDO_2d(coef, coef)
    for mode in ("warn", "raise"):
        dom = Domain_Solver(torch.tensor([[0, 1.0], [0, 1.0]]), op, 0.0, a, p=p, buf_constant=1.0,
                            check_pivots=mode)
        print("case singular leaf", mode, flush=True)
        if mode == "warn":
            with pytest.warns(RuntimeWarning, match="zero pivots"):
                info = dom.build(verbose=False)
            assert info["zero_pivots"]["leaf"] == 1
        else:
            with pytest.raises(FloatingPointError):
                dom.build(verbose=False)


def test_unconverged_refinement_is_flagged():
    n, p, ppw = 128, 10, 10
    dom, uu_dir, kh = _solver(n, p, ppw, precision="mixed", refine_maxit=1)
    print("case refinement cut off", flush=True)
    dom.build(verbose=False)
    with pytest.warns(RuntimeWarning, match="refinement stopped"):
        dom.solve(uu_dir)
    assert dom.refine_converged is False and dom.solve_info["refine_steps"] == 1
    dom.refine_maxit = 12
    dom.solve(uu_dir)
    assert dom.refine_converged is True
    assert dom.solve_info["refine_residuals"][-1] < dom.solve_info["refine_residuals"][0]
