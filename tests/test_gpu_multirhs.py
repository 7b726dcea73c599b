"""The batched multi-right-hand-side form: 64 Green's-function columns from
sources drawn uniformly from the annulus 1.2 < |x - (0.5, 0.5)| < 2 (seeded),
the waveguide coefficient, 10 ppw, p = 16.

  * the FP64 solve replayed as one captured HIP graph equals the eager solve
    (replays with eager work between them, a rebuild drops the graph);
  * 34 x 34 leaves with 64 columns, FP64 (graph) and mixed + refined against
    the oracle's SuperLU solve of all 64 columns;
  * the full workload, 132 x 132 leaves (N = 3,906,672): autotuned sweep
    tiles recorded (every candidate checked against the built-in tile), the
    FP64 solve's backward error, the residual of all 64 columns, and the
    mixed + refined solve at least as accurate as FP64."""
import numpy as np
import pytest
import torch

from oracle import hps_oracle as O

pytestmark = pytest.mark.gpu


def _dom(n0, p, precision, ppw=10, bufc=1.0):
    from slablu_amd.src_disc import pdo, built_in_funcs as bif
    from slablu_amd.src_solver.slablu_solver import Domain_Solver
    n = n0 * (p - 2)
    kh = O.kh_from_ppw(n, ppw)
    op = pdo.PDO_2d(pdo.ones, pdo.ones, c=lambda xx: bif.bfield_lshaped_waveguide(xx, kh))
    dom = Domain_Solver(torch.tensor([[0, 1.0], [0, 1.0]]), op, kh, 1 / (2 * n0), p=p, buf_constant=bufc,
                        precision=precision)
    srcs = bif.annulus_sources(64, seed=0)
    uu = lambda xx: bif.uu_dir_func_greens_multi(xx, kh, srcs)
    return dom, uu, kh, srcs


def test_annulus_sources():
    from slablu_amd.src_disc import built_in_funcs as bif
    s = np.array(bif.annulus_sources(5000, seed=3))
    r = np.hypot(s[:, 0] - 0.5, s[:, 1] - 0.5)
    assert r.min() > 1.2 and r.max() < 2.0
    # uniform by area: the fraction inside radius rho is (rho^2 - 1.44) / 2.56
    assert abs(np.mean(r < 1.6) - (1.6 ** 2 - 1.44) / 2.56) < 0.03
    assert bif.annulus_sources(4, seed=0) == bif.annulus_sources(4, seed=0)


def test_graph_solve_matches_eager():
    dom, uu, kh, _ = _dom(8, 16, "fp64")
    dom.build(verbose=False)
    print("case graph solve 8x8 p=16, 64 columns", flush=True)
    eager = dom.solve(uu)[0].to(torch.float64)
    for _ in range(3):
        g = dom.solve_graph(uu).cpu()
        torch.cuda.synchronize()
        _ = dom.solve(uu)                 # eager work between replays
        assert torch.equal(g, eager)
    assert len(dom._solve_graphs) == 1
    dom.build(verbose=False)
    assert dom._solve_graphs == {}
    assert torch.equal(dom.solve_graph(uu).cpu(), eager)


def test_multirhs_34x34_vs_oracle():
    n0, p = 34, 16
    hps, part, kh = O.setup_problem(n0 * (p - 2), p, ppw=10, bfield=O.bfield_lshaped_waveguide)
    from slablu_amd.src_disc import built_in_funcs as bif
    srcs = bif.annulus_sources(64, seed=0)
    A = hps.sparse_mat()
    uu_o = lambda xx: np.concatenate([O.uu_dir_func_greens(xx, kh, center=c) for c in srcs], axis=1)
    r = O.full_solve(hps, part, uu_o, known_sol=False, A=A)
    ref = r["sol_tot"]
    for precision, tol in (("fp64", 1e-10), ("mixed", 1e-8)):
        dom, uu, _, _ = _dom(n0, p, precision)
        print("case multi-rhs 34x34", precision, flush=True)
        dom.build(verbose=False)
        sol = dom.solve_graph(uu).cpu().numpy() if precision == "fp64" else dom.solve(uu)[0].numpy()
        err = np.linalg.norm(sol - ref, axis=0) / np.linalg.norm(ref, axis=0)
        print("multi-rhs", precision, "max column error vs oracle", err.max(), flush=True)
        assert err.max() < tol
        if precision == "mixed":
            assert dom.refine_converged


def test_multirhs_workload_132x132():
    from slablu_amd import autotune
    n0, p = 132, 16
    dom, uu, kh, _ = _dom(n0, p, "fp64")
    assert O.dof_count(p, n0, n0) == 3906672
    print("case multi-rhs 132x132 fp64", flush=True)
    dom.build(verbose=False)
    print("rem chains", type(dom.Ared_solver).__name__, dom.Npan, dom.rem_nblocks, dom.rem_bs, flush=True)
    be0 = dom.backward_error(nrhs=4)
    tuning = autotune.tune_sweeps(dom, 64)
    print("tuning", tuning, flush=True)
    assert set(tuning) == {"elim", "rem"} and all("nr" in v for v in tuning.values())
    assert all(not v["rejected"] for v in tuning.values())
    s64 = dom.solve_graph(uu)
    be = dom.backward_error(nrhs=4, normA=be0["norm_A"])
    print("fp64 two-level solve of random data, before / after tuning", be0, be, flush=True)
    # kh^2 = 1.34e6 lies above several interior Dirichlet eigenvalues of a
    # 0.0076-wide leaf (the first is about 3.4e5), and the bump lattice makes
    # some leaves nearly resonant: ||A_CC|| is 5.7e5 here against 4e4 for the
    # constant coefficient, and block elimination with pivoting inside the
    # blocks shows a backward error near 1e4 unit round-offs (3e-14 for the
    # constant coefficient at the same size)
    assert be0["backward_error"] < 1e-11 and be["backward_error"] < 1e-11
    _, res64, _, _, _ = dom.solve(uu)
    print("fp64 relative residual of the 64 Green's-function columns", res64, flush=True)
    assert res64 < 1e-6
    s64 = s64.cpu()
    del dom
    torch.cuda.empty_cache()
    dom, uu, kh, _ = _dom(n0, p, "mixed")
    print("case multi-rhs 132x132 mixed", flush=True)
    dom.build(verbose=False)
    sm, resm, _, _, _ = dom.solve(uu)
    err = (torch.linalg.norm(sm - s64, dim=0) / torch.linalg.norm(s64, dim=0)).max().item()
    print("mixed vs fp64, max over 64 columns", err, "mixed residual", resm, "Krylov directions", dom.refine_kmax,
          "refinement", dom.refine_history, flush=True)
    # the FP32 two-level preconditioner inherits the near-resonant leaves and
    # slabs: FGMRES needs about 30 directions here (3 to 5 at n = 1000,
    # p = 22), which the restart length sized from free HBM provides
    assert dom.refine_converged
    # the solutions agree to the FP64 direct solve's own accuracy, with
    # residuals of the same size
    assert resm < 1e-6
    assert err < 1e-6
