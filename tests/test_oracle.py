"""Pins the CPU oracle against closed forms, invariants and the golden
fixtures written from the reference (tests/golden/make_golden.py)."""
import json
import os

import numpy as np
import pytest
import scipy.special

from oracle import hps_oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _fig():
    with open(os.path.join(GOLD, "reference_figures.json")) as fh:
        return json.load(fh)


# --- 1-D / 2-D spectral operators -------------------------------------------

@pytest.mark.parametrize("p", [6, 9, 15, 21])
def test_cheb_closed_forms(p):
    D, x = O.cheb(p)
    assert np.allclose(x, np.cos(np.pi * np.arange(p + 1) / p))
    assert np.abs(D @ np.ones(p + 1)).max() < 1e-11
    assert np.abs(D @ x - 1).max() < 1e-11
    for k in range(2, p + 1):
        # exact on polynomials of degree <= p
        assert np.abs(D @ x ** k - k * x ** (k - 1)).max() < 1e-7 * k * p ** 2


@pytest.mark.parametrize("p,a", [(10, 1 / 64), (16, 1 / 68), (22, 0.01)])
def test_cheb_2d_derivatives(p, a):
    zz, Ds, D, Dsq = O.cheb_2d(a, p)
    x, y = zz
    assert np.all(np.diff(np.unique(x)) > 0) and abs(x.min() + a) < 1e-15 and abs(x.max() - a) < 1e-15
    u = x ** 3 + 2 * x * y ** 2 - y ** 4 + 0.5 * x * y
    ux = 3 * x ** 2 + 2 * y ** 2 + 0.5 * y
    uy = 4 * x * y - 4 * y ** 3 + 0.5 * x
    uxx, uyy, uxy = 6 * x, 4 * x - 12 * y ** 2, 4 * y + 0.5
    scale = 1 / a ** 2
    for M, ref in [(Ds["D1"], ux), (Ds["D2"], uy), (Ds["D11"], uxx), (Ds["D22"], uyy), (Ds["D12"], uxy)]:
        assert np.abs(M @ u - ref).max() < 1e-9 * scale


@pytest.mark.parametrize("p,a", [(10, 1 / 64), (22, 0.025)])
def test_leaf_index_sets_and_nx(p, a):
    zz, Ds, JJ, hmin, D, Dsq = O.leaf_discretization_2d(a, p)
    x, y = zz
    f = p - 2
    assert np.allclose(x[JJ["Jl"]], -a) and np.allclose(x[JJ["Jr"]], a)
    assert np.allclose(y[JJ["Jd"]], -a) and np.allclose(y[JJ["Ju"]], a)
    assert np.all(np.diff(y[JJ["Jl"]]) > 0) and np.all(np.diff(x[JJ["Jd"]]) > 0)
    assert len(JJ["Jc"]) == f * f and len(JJ["Jx"]) == 4 * f and len(JJ["Jxreorder"]) == 4 * (p - 1)
    assert np.all(np.abs(x[JJ["Jc"]]) < a) and np.all(np.abs(y[JJ["Jc"]]) < a)
    allidx = np.concatenate((JJ["Jc"], JJ["Jx"]))
    assert len(np.unique(allidx)) == len(allidx)
    # Jxreorder: SW corner, left face up, NW corner, top face right, NE, right face down, SE, bottom left
    jr = JJ["Jxreorder"]
    assert (x[jr[0]], y[jr[0]]) == (-a, -a) and (x[jr[p - 1]], y[jr[p - 1]]) == (-a, a)
    assert (x[jr[2 * (p - 1)]], y[jr[2 * (p - 1)]]) == (a, a) and (x[jr[3 * (p - 1)]], y[jr[3 * (p - 1)]]) == (a, -a)
    Nx = O.diff_ops(Ds, JJ)
    u = np.sin(3 * x / a) * 0 + x ** 2 * y + 2 * y ** 3 - x
    ux, uy = 2 * x * y - 1, x ** 2 + 6 * y ** 2
    nrm = np.concatenate((-ux[JJ["Jl"]], ux[JJ["Jr"]], -uy[JJ["Jd"]], uy[JJ["Ju"]]))
    assert np.abs(Nx @ u - nrm).max() < 1e-9 / a


def test_interp_mat_golden():
    im = np.load(os.path.join(GOLD, "interp_mat.npz"))
    keys = [k for k in im.files if not k.startswith("a_")]
    assert len(keys) >= 6
    for key in keys:
        p = int(key.split("_")[0][1:])
        a = float(im["a_" + key])
        M = O.interp_mat(a, p)
        assert M.shape == (4 * (p - 1), 4 * (p - 2))
        assert np.array_equal(M, im[key]), key
        # face rows reproduce the face data (least-squares fit is consistent)
        zz, Ds, JJ, *_ = O.leaf_discretization_2d(a, p)
        x, y = zz
        u = 1 + x / a - 0.5 * (y / a) ** 2
        v = M @ u[JJ["Jx"]]
        assert np.abs(v - u[JJ["Jxreorder"]]).max() < 1e-6


def test_greens_closed_form():
    rng = np.random.default_rng(0)
    xx = rng.uniform(0, 1, (50, 2))
    kh = 37.3
    u = O.uu_dir_func_greens(xx, kh)[:, 0]
    r = np.hypot(xx[:, 0] + 0.1, xx[:, 1] - 0.5)
    assert np.abs(u - (-scipy.special.y0(kh * r) / 4)).max() < 1e-13
    u0 = O.uu_dir_func_greens(xx, 0.0, center=(1.5, 1.5))[:, 0]
    r0 = np.hypot(xx[:, 0] - 1.5, xx[:, 1] - 1.5)
    assert np.abs(u0 - np.log(r0 ** 2) / np.pi).max() < 1e-13


# --- leaf DtN ------------------------------------------------------------------

def _laplace_hps(n, p):
    a = 1 / (2 * (n / (p - 2)))
    ones = lambda xx: np.ones(xx.shape[0])
    return O.HPSOracle(O.PDO(ones, ones), np.array([[0.0, 1.0], [0.0, 1.0]]), a, p)


@pytest.mark.parametrize("p", [10, 16, 22])
def test_dtn_laplace_exact_on_linears(p):
    hps = _laplace_hps(4 * (p - 2), p)
    DtN = hps.dtns(np.arange(3))
    JJ = hps.JJ
    f = p - 2
    for b in range(3):
        xx = hps.grid_xx[b][JJ["Jx"]]
        one = np.ones(4 * f)
        nx = np.concatenate((-np.ones(f), np.ones(f), np.zeros(2 * f)))
        ny = np.concatenate((np.zeros(2 * f), -np.ones(f), np.ones(f)))
        scale = np.abs(DtN[b]).max()
        assert np.abs(DtN[b] @ one).max() < 1e-11 * scale
        assert np.abs(DtN[b] @ xx[:, 0] - nx).max() < 1e-10 * scale * hps.a
        assert np.abs(DtN[b] @ xx[:, 1] - ny).max() < 1e-10 * scale * hps.a


def test_dtn_helmholtz_plane_wave_and_bruteforce():
    p, n, ppw = 22, 400, 40
    hps, part, kh = O.setup_problem(n, p, ppw=ppw)
    boxes = np.array([0, 17, 399])
    DtN = hps.dtns(boxes)
    JJ = hps.JJ
    th = 0.7
    for k, b in enumerate(boxes):
        xx = hps.grid_xx[b]
        ph = kh * (xx[:, 0] * np.cos(th) + xx[:, 1] * np.sin(th))
        u, ux, uy = np.cos(ph), -kh * np.cos(th) * np.sin(ph), -kh * np.sin(th) * np.sin(ph)
        dn = np.concatenate((-ux[JJ["Jl"]], ux[JJ["Jr"]], -uy[JJ["Jd"]], uy[JJ["Ju"]]))
        assert np.abs(DtN[k] @ u[JJ["Jx"]] - dn).max() < 1e-7 * kh
    # brute force: full leaf system with identity rows on all non-interior nodes
    A = hps.aloc(boxes)
    p2 = p * p
    Jc = JJ["Jc"]
    nonc = np.setdiff1d(np.arange(p2), Jc)
    for k in range(len(boxes)):
        sc = np.abs(A[k]).max()  # identity rows at the scale of the PDE rows
        M = np.zeros((p2, p2))
        M[Jc] = A[k][Jc]
        M[nonc, nonc] = sc
        rhs = np.zeros((p2, len(JJ["Jx"])))
        rhs[JJ["Jx"], np.arange(len(JJ["Jx"]))] = sc
        U = np.linalg.solve(M, rhs)
        ref = hps.Nx @ U
        assert np.abs(ref - DtN[k]).max() < 1e-10 * np.abs(DtN[k]).max()


def test_sparse_mat_invariants():
    hps = _laplace_hps(48, 10)
    A = hps.sparse_mat()
    nf = hps.p - 2
    assert A.shape[0] == nf * (hps.n0 * (2 * hps.n1 + 1) + hps.n1)
    assert np.abs(A @ np.ones(A.shape[0])).max() < 1e-9 * abs(A).max()
    # flux of a linear function cancels on every interior face
    x = hps.xx_active[:, 0]
    r = A @ x
    part = O.SlabPartition(hps.n0, hps.n1, hps.p, 1.0)
    assert np.abs(r[part.I_Ctot]).max() < 1e-8 * abs(A).max()


# --- slab partition -------------------------------------------------------------

@pytest.mark.parametrize("n0,p,bufc,buf,Npan,nC,nX,nCtot,nXtot", [
    (50, 22, 1.0, 5, 10, 8900, 9000, 98000, 4000),
    (20, 22, 1.0, 4, 5, None, None, None, None),
    (32, 10, 0.3, 2, 16, None, None, None, None),
    (132, 16, 1.0, 11, 12, None, None, None, None),
    (34, 16, 1.0, 2, 17, None, None, None, None),
])
def test_slab_partition_sizes(n0, p, bufc, buf, Npan, nC, nX, nCtot, nXtot):
    part = O.SlabPartition(n0, n0, p, bufc)
    assert (part.buf, part.Npan) == (buf, Npan)
    f = p - 2
    nface = f * (n0 * (2 * n0 + 1) + n0)
    if nC is not None:
        assert len(part.I_C) == nC and len(part.I_slabX) == nX
        assert len(part.I_Ctot) == nCtot and len(part.I_Xtot) == nXtot
    allidx = np.concatenate((part.I_Ctot, part.I_Xtot))
    assert len(allidx) == nface and np.array_equal(np.sort(allidx), np.arange(nface))


# --- end to end against the reference's own figures ------------------------------

@pytest.mark.parametrize("tag", ["small_superLU", "p16_superLU"])
def test_end_to_end_small_vs_reference(tag):
    F = _fig()[tag]
    sols = np.load(os.path.join(GOLD, "reference_solutions.npz"))
    hps, part, kh = O.setup_problem(F["n"], F["p"], ppw=F["ppw"])
    assert O.dof_count(F["p"], hps.n0, hps.n1) == F["N"]
    assert (part.buf, part.Npan) == (F["buf"], F["Npan"])
    r = O.full_solve(hps, part, lambda xx: O.uu_dir_func_greens(xx, kh))
    assert r["res"] < 1e-12
    # the known-solution error agrees up to the solvers' rounding difference
    # (|delta true_err| <= ||delta sol|| / ||u_true||, ~1e-12 here)
    assert abs(r["true_err"] - F["trueres"]) < 1e-11
    s = sols[tag]
    assert np.linalg.norm(r["sol_tot"].ravel() - s) / np.linalg.norm(s) < 1e-10
    # the two-level slab restatement gives the same discrete solution
    sl = O.SlabLUOracle(r["A"], part).build()
    x = sl.solve(r["ff"])
    assert np.linalg.norm(x - r["x_C"]) / np.linalg.norm(r["x_C"]) < 1e-11


@pytest.mark.slow
def test_end_to_end_test2d_square_vs_reference():
    """tests/test_2d.py square configuration (n=400, p=22, ppw=40)."""
    figs = _fig()
    F = figs["test2d_square_superLU"]
    hps, part, kh = O.setup_problem(F["n"], F["p"], ppw=F["ppw"])
    r = O.full_solve(hps, part, lambda xx: O.uu_dir_func_greens(xx, kh))
    assert r["true_err"] < 2e-5
    assert abs(r["true_err"] - F["trueres"]) < 1e-11
    assert abs(r["true_err"] - figs["test2d_square_slabLU"]["trueres"]) < 1e-11


def test_oracle_periodic_known_solution():
    """x-periodic oracle (aliased seam faces, cyclic slab interfaces): the
    exact mode cos(2 pi m x) cos(k_y y) is recovered to discretization error,
    the same order as the non-periodic solve of the same mode, and A_CC x = f
    holds to rounding."""
    for n, p in ((64, 10), (112, 16)):
        hps, part, kh = O.setup_problem(n, p, ppw=10, periodic=True, buf_constant=0.6)
        mode = lambda xx: O.uu_periodic_mode(xx, kh, 2)
        r = O.full_solve(hps, part, mode)
        hps2, part2, _ = O.setup_problem(n, p, kh=kh, periodic=False, buf_constant=0.6)
        r2 = O.full_solve(hps2, part2, mode)
        assert r["res"] < 1e-12
        assert r["true_err"] < 5 * r2["true_err"] + 1e-12
        assert len(np.intersect1d(part.I_Ctot, part.I_Xtot)) == 0
        assert len(np.unique(part.I_Ctot)) == len(part.I_Ctot)
