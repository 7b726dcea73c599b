"""End-to-end parity of the GPU SlabLU path (Domain_Solver) with the CPU
oracle (exact discrete solution through SuperLU) and with the reference's
own figures (tests/golden/reference_figures.json)."""
import json
import os

import numpy as np
import pytest
import torch

from oracle import hps_oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _fig(tag):
    with open(os.path.join(GOLD, "reference_figures.json")) as fh:
        return json.load(fh)[tag]


def _solver(n, p, ppw, precision="fp64", bufc=1.0, box=(1.0, 1.0), center=(-0.1, 0.5)):
    from slablu_amd.src_disc import pdo, built_in_funcs as bif
    from slablu_amd.src_solver.slablu_solver import Domain_Solver
    kh = O.kh_from_ppw(n, ppw)
    op = pdo.PDO_2d(pdo.ones, pdo.ones, c=lambda xx: bif.bfield_constant(xx, kh))
    a = 1 / (2 * (n / (p - 2)))
    dom = Domain_Solver(torch.tensor([[0, box[0]], [0, box[1]]]), op, kh, a, p=p, buf_constant=bufc,
                        precision=precision)
    uu_dir = lambda xx: bif.uu_dir_func_greens(xx, kh, center=center)
    return dom, kh, a, uu_dir


def _oracle(n, p, ppw, bufc=1.0, box=(1.0, 1.0), center=(-0.1, 0.5)):
    kh = O.kh_from_ppw(n, ppw)
    a = 1 / (2 * (n / (p - 2)))
    hps = O.HPSOracle(O.helmholtz_pdo(kh), np.array([[0, box[0]], [0, box[1]]]), a, p)
    part = O.SlabPartition(hps.n0, hps.n1, p, bufc)
    r = O.full_solve(hps, part, lambda xx: O.uu_dir_func_greens(xx, kh, center=center))
    return hps, part, r


@pytest.mark.parametrize("tag", ["small_superLU", "p16_superLU"])
def test_fp64_vs_reference_solution(tag):
    F = _fig(tag)
    sols = np.load(os.path.join(GOLD, "reference_solutions.npz"))
    dom, kh, a, uu_dir = _solver(F["n"], F["p"], F["ppw"])
    print("case", tag, flush=True)
    dom.build(verbose=False)
    sol, res, true_err, resloc, toc = dom.solve(uu_dir, known_sol=True)
    ref = sols[tag]
    err = np.linalg.norm(sol.numpy().ravel() - ref) / np.linalg.norm(ref)
    print(tag, "rel err vs reference solution", err, "true_err", true_err, F["trueres"], flush=True)
    assert err < 1e-10
    assert abs(true_err - F["trueres"]) < 1e-11
    assert res < 1e-11


def test_dtn_parity_test2d():
    n, p, ppw = 400, 22, 40
    dom, kh, a, uu_dir = _solver(n, p, ppw)
    dom.hps.factor_leaves()
    kh = O.kh_from_ppw(n, ppw)
    hps = O.HPSOracle(O.helmholtz_pdo(kh), np.array([[0, 1.0], [0, 1.0]]), a, p)
    boxes = np.array([0, 1, 57, 211, 399])
    ref = hps.dtns(boxes)
    got = dom.hps.leaf["DtN"][torch.as_tensor(boxes)].cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("DtN rel err", err, flush=True)
    assert err < 1e-12
    assert int(dom.hps.leaf["info"].abs().max()) == 0


def test_fp64_test2d_square_end_to_end():
    """tests/test_2d.py square case (n=400, p=22, ppw=40): solution vs the
    oracle's exact discrete solution and known-solution error vs the
    reference's figures."""
    n, p, ppw = 400, 22, 40
    dom, kh, a, uu_dir = _solver(n, p, ppw)
    info = dom.build(verbose=True)
    sol, res, true_err, resloc, toc = dom.solve(uu_dir, known_sol=True)
    sol, res, true_err, resloc, toc = dom.solve(uu_dir, known_sol=True)
    hps, part, r = _oracle(n, p, ppw)
    err = np.linalg.norm(sol.numpy() - r["sol_tot"]) / np.linalg.norm(r["sol_tot"])
    print("test2d: rel err vs oracle", err, "true_err", true_err, "res", res, "resloc", resloc,
          "build", info["toc_build"], "solve", toc, flush=True)
    assert err < 1e-10
    assert true_err < 2e-5
    assert abs(true_err - _fig("test2d_square_superLU")["trueres"]) < 1e-11
    assert abs(true_err - r["true_err"]) < 1e-11
    assert res < 1e-11
    assert resloc < 10 * r["resloc"] + 1e-6
    assert dom.test_slab_solve() < 1e-11


def test_fp64_nonsquare_and_center():
    """Rectangular domain, Dirichlet source at (1.5, 1.5)."""
    n, p, ppw = 80, 10, 10
    dom, kh, a, uu_dir = _solver(n, p, ppw, box=(1.5, 1.0), center=(1.5, 1.5))
    dom.build(verbose=False)
    sol, res, true_err, resloc, toc = dom.solve(uu_dir, known_sol=True)
    hps, part, r = _oracle(n, p, ppw, box=(1.5, 1.0), center=(1.5, 1.5))
    err = np.linalg.norm(sol.numpy() - r["sol_tot"]) / np.linalg.norm(r["sol_tot"])
    print("nonsquare rel err", err, true_err, r["true_err"], flush=True)
    assert err < 1e-10
    assert abs(true_err - r["true_err"]) < 1e-11


def test_multi_rhs_matches_single():
    from slablu_amd.src_disc import built_in_funcs as bif
    n, p, ppw = 80, 10, 10
    dom, kh, a, _ = _solver(n, p, ppw)
    dom.build(verbose=False)
    centers = [(1.3, 0.2), (-0.4, 1.1), (0.5, -0.6)]
    multi = lambda xx: bif.uu_dir_func_greens_multi(xx, kh, centers)
    solm, resm, tem, _, _ = dom.solve(multi, known_sol=True)
    assert solm.shape[1] == 3
    for k, c in enumerate(centers):
        s1, _, _, _, _ = dom.solve(lambda xx: bif.uu_dir_func_greens(xx, kh, center=c))
        assert torch.linalg.norm(solm[:, k] - s1[:, 0]) / torch.linalg.norm(s1) < 1e-12


def test_mixed_refined_vs_fp64_oracle():
    n, p, ppw = 400, 22, 40
    dom, kh, a, uu_dir = _solver(n, p, ppw, precision="mixed")
    dom.build(verbose=True)
    sol, res, true_err, resloc, toc = dom.solve(uu_dir, known_sol=True)
    hps, part, r = _oracle(n, p, ppw)
    err = np.linalg.norm(sol.numpy() - r["sol_tot"]) / np.linalg.norm(r["sol_tot"])
    print("mixed: rel err vs FP64 oracle", err, "refinement", dom.refine_history, "true_err", true_err, flush=True)
    assert err < 1e-8
    # the known-solution errors of two solutions differ by at most their
    # distance (triangle inequality): within the refinement target of the
    # FP64 one
    assert abs(true_err - r["true_err"]) <= err * 1.0001 + 1e-12


def test_refinement_graph_replays_match_eager():
    """The FGMRES operator step replayed as a captured HIP graph gives the
    eager step's iterates: two solves with graphs (replays with eager work and
    synchronizations between them, the second reusing the captured graph)
    against one without, solution and refinement history identical."""
    n, p, ppw = 200, 22, 20
    dom, kh, a, uu_dir = _solver(n, p, ppw, precision="mixed")
    dom.build(verbose=False)
    dom.refine_graphs = False
    sol0 = dom.solve(uu_dir)[0]
    hist0 = list(dom.refine_history)
    dom.refine_graphs = True
    for _ in range(2):
        sol1 = dom.solve(uu_dir)[0]       # a solve after the first: captured step
        torch.cuda.synchronize()
        assert len(dom._gsteps) == 1
        assert dom.refine_history == hist0
        assert torch.equal(sol1, sol0)
    # a rebuild drops the captured step (it reads the old factors), and the
    # first solve after it runs eagerly
    dom.build(verbose=False)
    assert dom._gsteps == {}
    sol2 = dom.solve(uu_dir)[0]
    assert dom._gsteps == {}
    assert torch.equal(sol2, sol0)


def _periodic(n, p, ppw, precision, bufc=0.6, m=2):
    """x-periodic Helmholtz on the unit square (n/(p-2) leaves per side, so
    the leaves tile the period exactly), Dirichlet data on the bottom and top
    from the exact mode cos(2 pi m x) cos(k_y y)."""
    from slablu_amd.src_disc import pdo, built_in_funcs as bif
    from slablu_amd.src_solver.slablu_solver import Domain_Solver
    kh = O.kh_from_ppw(n, ppw)
    op = pdo.PDO_2d(pdo.ones, pdo.ones, c=lambda xx: bif.bfield_constant(xx, kh))
    a = 1 / (2 * (n / (p - 2)))
    dom = Domain_Solver(torch.tensor([[0, 1.0], [0, 1.0]]), op, kh, a, p=p, buf_constant=bufc,
                        periodic_bc=True, precision=precision)
    uu = lambda xx: torch.as_tensor(O.uu_periodic_mode(np.asarray(xx), kh, m))
    hps, part, _ = O.setup_problem(n, p, kh=kh, periodic=True, buf_constant=bufc)
    r = O.full_solve(hps, part, lambda xx: O.uu_periodic_mode(xx, kh, m))
    return dom, uu, r


@pytest.mark.parametrize("n,p", [(64, 10), (128, 10), (112, 16)])
def test_periodic_fp64_vs_oracle(n, p):
    """Periodic slabs (cyclic interface ring) against the oracle's SuperLU
    solve of the same x-periodic discrete system, and its known-solution
    error."""
    dom, uu, r = _periodic(n, p, 10, "fp64")
    print("periodic case", n, p, "Npan", dom.Npan, flush=True)
    dom.build(verbose=False)
    sol, res, true_err, resloc, toc = dom.solve(uu, known_sol=True)
    err = np.linalg.norm(sol.numpy() - r["sol_tot"]) / np.linalg.norm(r["sol_tot"])
    print("periodic fp64 rel err vs oracle", err, "true_err", true_err, r["true_err"], flush=True)
    assert err < 1e-10
    assert abs(true_err - r["true_err"]) < 1e-10
    assert res < 1e-10


def test_periodic_mixed_vs_oracle():
    dom, uu, r = _periodic(128, 10, 10, "mixed")
    dom.build(verbose=False)
    sol, res, true_err, resloc, toc = dom.solve(uu, known_sol=True)
    err = np.linalg.norm(sol.numpy() - r["sol_tot"]) / np.linalg.norm(r["sol_tot"])
    print("periodic mixed rel err vs oracle", err, "refinement", dom.refine_history, flush=True)
    assert err < 1e-8


def _curved(geom, n, p, ppw, box, periodic, precision="fp64"):
    """Helmholtz on a mapped domain (annulus sector, or the full curvy
    annulus with x-periodic slabs), Dirichlet data = the free-space Green's
    function at the mapped nodes, as the reference's test_2d.py drives it."""
    from slablu_amd.src_disc import pdo, built_in_funcs as bif
    from slablu_amd.src_solver.slablu_solver import Domain_Solver
    kh = O.kh_from_ppw(n, ppw)
    op, zmap, _ = pdo.get_param_map_and_pdo(geom, bif.bfield_constant, kh)
    a = 1 / (2 * (n / (p - 2)))
    dom = Domain_Solver(torch.tensor([[0, box[0]], [0, box[1]]]), op, kh, a, p=p, buf_constant=0.6,
                        periodic_bc=periodic, precision=precision)
    uu = lambda xx: bif.uu_dir_func_greens(zmap(torch.as_tensor(xx)), kh)
    npf = lambda f: None if f is None else (lambda xx: f(torch.as_tensor(xx)).numpy().reshape(-1))
    opo = O.PDO(npf(op.c11), npf(op.c22), c12=npf(op.c12), c1=npf(op.c1), c2=npf(op.c2), c=npf(op.c))
    hps = O.HPSOracle(opo, np.array([[0, box[0]], [0, box[1]]]), a, p, periodic=periodic)
    part = O.SlabPartition(hps.n0, hps.n1, p, 0.6, periodic=periodic)
    r = O.full_solve(hps, part, lambda xx: np.asarray(uu(xx)).reshape(-1, 1))
    return dom, uu, r, hps


@pytest.mark.parametrize("geom,box,periodic", [("annulus", (1.0, 1.0), False), ("sinusoidal", (1.0, 1.0), False),
                                               ("curvy_annulus", (6.0, 1.0), True)])
def test_curved_domain_fp64_vs_oracle(geom, box, periodic):
    """Mapped domains with the c12 cross term (corner-inclusive leaf
    condensation through Interp_mat) against the oracle's SuperLU solve of
    the same discrete system; the curvy annulus is the reference's
    test_helm_poisson_curvyannulus configuration (box 6 x 1, periodic) at a
    smaller n."""
    dom, uu, r, hps = _curved(geom, 64, 10, 40, box, periodic)
    print("curved case", geom, "Npan", dom.Npan, flush=True)
    dom.build(verbose=False)
    DtN = dom.hps.leaf["DtN"].double().cpu().numpy()
    DtN_o = hps.dtns()
    assert np.abs(DtN - DtN_o).max() < 1e-9 * np.abs(DtN_o).max()
    sol, res, true_err, resloc, toc = dom.solve(uu, known_sol=True)
    err = np.linalg.norm(sol.numpy() - r["sol_tot"]) / np.linalg.norm(r["sol_tot"])
    print(geom, "fp64 rel err vs oracle", err, "true_err", true_err, r["true_err"], flush=True)
    assert err < 1e-10
    assert abs(true_err - r["true_err"]) < 1e-10
    assert true_err < 1e-3


def test_curved_domain_mixed_vs_oracle():
    dom, uu, r, _ = _curved("annulus", 64, 10, 40, (1.0, 1.0), False, precision="mixed")
    dom.build(verbose=False)
    sol, res, true_err, resloc, toc = dom.solve(uu, known_sol=True)
    err = np.linalg.norm(sol.numpy() - r["sol_tot"]) / np.linalg.norm(r["sol_tot"])
    print("annulus mixed rel err vs oracle", err, "refinement", dom.refine_history, flush=True)
    assert err < 1e-8


def test_body_load_manufactured_solution():
    """Body loads, FP64 and mixed: u = sin(a x) sin(b y) solves
    -Lap u - kh^2 u = (a^2 + b^2 - kh^2) u; both paths recover it to
    discretization error and agree with each other to the mixed tolerance."""
    from slablu_amd.src_disc import pdo, built_in_funcs as bif
    from slablu_amd.src_solver.slablu_solver import Domain_Solver
    n, p, ppw = 128, 10, 10
    kh = O.kh_from_ppw(n, ppw)
    ca, cb = 9.0, 13.0
    exact = lambda xx: (torch.sin(ca * torch.as_tensor(xx)[:, 0]) * torch.sin(cb * torch.as_tensor(xx)[:, 1]))[:, None]
    body = lambda xx: (ca ** 2 + cb ** 2 - kh ** 2) * exact(xx)
    op = pdo.PDO_2d(pdo.ones, pdo.ones, c=lambda xx: bif.bfield_constant(xx, kh))
    a = 1 / (2 * (n / (p - 2)))
    sols = {}
    for precision in ("fp64", "mixed"):
        dom = Domain_Solver(torch.tensor([[0, 1.0], [0, 1.0]]), op, kh, a, p=p, buf_constant=0.6,
                            precision=precision)
        dom.build(verbose=False)
        sol, res, true_err, resloc, toc = dom.solve(exact, ff_body_func=body, known_sol=True)
        print("body load", precision, "true_err", true_err, flush=True)
        assert true_err < 1e-5
        sols[precision] = sol
    err = torch.linalg.norm(sols["mixed"] - sols["fp64"]) / torch.linalg.norm(sols["fp64"])
    print("body load mixed vs fp64", err.item(), flush=True)
    assert err < 1e-8
