"""CPU oracle for the SlabLU factor+solve path (numpy / scipy, FP64).

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this file:
it is imported by `tests/`, by `__graft_entry__.smoke()` and by the
`cpu_baseline` leg of `bench.py`, and only as the checker or the timed CPU
baseline.  It restates the reference algorithm step by step so that every
quantity the GPU path produces (leaf DtN maps, the sparse face operator A,
slab index sets, the exact discrete solution, leaf recovery, known-solution
error) can be compared with an independent FP64 computation.

Reference files cited below are relative to the reference checkout
(annayesy/slabLU).

Pinning (what checks each function; see tests/test_oracle.py):
  cheb, cheb_2d           closed forms: D.1 = 0, D.x = 1, D^2 exact on
                          polynomials of degree <= p-1.
  leaf_discretization_2d  node coordinates of each index set, counts.
  diff_ops (Nx)           Nx.u = outward normal derivative of polynomials.
  interp_mat              golden fixture written from the reference
                          (tests/golden/make_golden.py) + polynomial exactness.
  HPSOracle.dtns          Laplace: DtN.1 = 0 and DtN.(x-trace) = n_x;
                          Helmholtz: DtN.trace(plane wave) = d_n(plane wave)
                          to spectral accuracy; brute-force dense condensation.
  HPSOracle.sparse_mat    Laplace row sums vanish on interior faces; the
                          Green's-function face trace satisfies A_CC u = f to
                          discretization error.
  slab index sets         counts of the sizes table and disjoint-union checks.
  superlu_solve           residual of A_CC x = f; known-solution error against
                          the reference's own superLU-path figure (golden).
  SlabLUOracle            the same solution as superlu_solve (exact algebra).
  leaf_solve              known solution at every node; leaf residual.
  uu_dir_func_greens      scipy.special.y0 closed form -Y0(k r)/4.
"""
from __future__ import annotations

from functools import reduce

import numpy as np
import numpy.polynomial.chebyshev as cheb_py
import scipy.linalg
import scipy.sparse
import scipy.sparse.linalg as spla
from scipy.special import hankel1

# ----------------------------------------------------------------------------
# 1. Leaf spectral discretization  (src_disc/hps_subdomain_disc.py)
# ----------------------------------------------------------------------------


def cheb(p):
    """Chebyshev differentiation matrix on x_k = cos(pi k / p) (descending).

    Follows src_disc/hps_subdomain_disc.py:16-25 (Trefethen's cheb.m)."""
    x = np.cos(np.pi * np.arange(p + 1) / p)
    c = np.concatenate((np.array([2.0]), np.ones(p - 1), np.array([2.0])))
    c = c * np.power(-np.ones(p + 1), np.arange(p + 1))
    X = x.repeat(p + 1).reshape((-1, p + 1))
    dX = X - X.T
    D = np.outer(c, 1.0 / c) / (dX + np.eye(p + 1))
    D = D - np.diag(np.sum(D, axis=1))
    return D, x


def cheb_2d(a, p):
    """Tensor Chebyshev grid on [-a,a]^2 and the five derivative matrices.

    Follows src_disc/hps_subdomain_disc.py:136-152.  Node id = i*p + j with
    x = xvec[i], y = xvec[j]; xvec ascending.  Returns zz (2, p^2) and the
    dict of D11, D22, D12, D1, D2 (each p^2 x p^2) plus the 1-D factors
    D (scaled, descending-node convention) and Dsq = D @ D."""
    D, xvec = cheb(p - 1)
    xvec = a * np.flip(xvec)
    D = (1 / a) * D
    I = np.eye(p)
    Dsq = D @ D
    Ds = dict(
        D11=np.kron(Dsq, I),
        D22=np.kron(I, Dsq),
        D12=np.kron(D, D),
        D1=-np.kron(D, I),
        D2=-np.kron(I, D),
    )
    zz1 = np.repeat(xvec, p)
    zz2 = np.repeat(xvec, p).reshape(-1, p).T.flatten()
    zz = np.vstack((zz1, zz2))
    return zz, Ds, D, Dsq


def leaf_discretization_2d(a, p):
    """Index sets Jl, Jr, Jd, Ju, Jx, Jc, Jxreorder of one leaf.

    Follows src_disc/hps_subdomain_disc.py:105-134 (corners excluded from Jx;
    Jxreorder walks SW->left->NW->top->NE->right->SE->bottom, corners
    included)."""
    zz, Ds, D, Dsq = cheb_2d(a, p)
    hmin = zz[1, 1] - zz[0, 1]
    Jc0 = np.abs(zz[0, :]) < a - 0.5 * hmin
    Jc1 = np.abs(zz[1, :]) < a - 0.5 * hmin
    Jl = np.argwhere(np.logical_and(zz[0, :] < -a + 0.5 * hmin, Jc1)).reshape(p - 2)
    Jr = np.argwhere(np.logical_and(zz[0, :] > +a - 0.5 * hmin, Jc1)).reshape(p - 2)
    Jd = np.argwhere(np.logical_and(zz[1, :] < -a + 0.5 * hmin, Jc0)).reshape(p - 2)
    Ju = np.argwhere(np.logical_and(zz[1, :] > +a - 0.5 * hmin, Jc0)).reshape(p - 2)
    Jc = np.argwhere(np.logical_and(Jc0, Jc1)).reshape((p - 2) ** 2)
    Jx = np.concatenate((Jl, Jr, Jd, Ju))
    Jcorner = np.setdiff1d(np.arange(p ** 2), np.hstack((Jc, Jx)))
    Jl_c = np.hstack((Jcorner[0], Jl))
    Ju_c = np.hstack((Jcorner[1], Ju))
    Jr_c = np.hstack((Jcorner[3], np.flip(Jr, 0)))
    Jd_c = np.hstack((Jcorner[2], np.flip(Jd, 0)))
    Jxreorder = np.hstack((Jl_c, Ju_c, Jr_c, Jd_c))
    JJ = dict(Jl=Jl, Jr=Jr, Jd=Jd, Ju=Ju, Jx=Jx, Jc=Jc, Jxreorder=Jxreorder)
    return zz, Ds, JJ, hmin, D, Dsq


def diff_ops(Ds, JJ):
    """Outward normal-derivative rows Nx = [-D1[Jl]; D1[Jr]; -D2[Jd]; D2[Ju]].

    Follows src_disc/hps_subdomain_disc.py:83-100."""
    return np.concatenate((-Ds["D1"][JJ["Jl"]], +Ds["D1"][JJ["Jr"]],
                           -Ds["D2"][JJ["Jd"]], +Ds["D2"][JJ["Ju"]]))


def _loc_interp(x_cheb, x_nc, q):
    """src_disc/hps_subdomain_disc.py:30-37."""
    Vc = cheb_py.chebvander(x_cheb, q)
    Vnc = cheb_py.chebvander(x_nc, q)
    Interp_loc = np.linalg.lstsq(Vnc.T, Vc.T, rcond=None)[0].T
    err = np.linalg.norm(Interp_loc @ Vnc - Vc)
    cond = np.linalg.cond(Interp_loc)
    return Interp_loc, err, cond


def interp_mat(a, p, zz=None):
    """Face -> corner-inclusive interpolation Interp_mat (4(p-1) x 4(p-2)).

    Follows src_disc/hps_subdomain_disc.py:53-80: least-squares fit over the
    first q+1 Chebyshev polynomials, q = p+5, p+15, ... until the condition
    test passes; four copies block-diagonal, columns permuted from the
    [l, r, d, u] face order to the Jxreorder walk."""
    if zz is None:
        zz = cheb_2d(a, p)[0]
    x_cheb = zz[1, :p - 1] + a
    x_nc = zz[1, 1:p - 1] + a
    cond_min, cond_max, err_tol = 3.25, 3.5, 2e-8
    q = p + 5
    while True:
        Interp_loc, err, cond = _loc_interp(x_cheb, x_nc, q)
        if (cond < cond_min) or ((err > err_tol) and (cond < cond_max)):
            break
        q += 10
    blk = scipy.linalg.block_diag(*np.repeat(np.expand_dims(Interp_loc, 0), 4, axis=0))
    perm = np.hstack((np.arange(p - 2), np.arange(p - 2) + 3 * (p - 2),
                      np.flip(np.arange(p - 2) + 1 * (p - 2), 0),
                      np.flip(np.arange(p - 2) + 2 * (p - 2), 0)))
    perm = np.argsort(perm)
    return blk[:, perm]


# ----------------------------------------------------------------------------
# 2. Built-in coefficient / Dirichlet functions (src_disc/built_in_funcs.py,
#    src_disc/pdo.py), numpy versions: xx (M,2) -> (M,)
# ----------------------------------------------------------------------------


def uu_dir_func_greens(xx, kh, center=(-0.1, 0.5)):
    """Re[(i/4) H0^(1)(kh r)] (kh != 0) or (1/pi) log r^2 (kh == 0).

    Follows src_disc/built_in_funcs.py:59-70.  Returns (M, 1)."""
    dd0 = xx[:, 0] - center[0]
    dd1 = xx[:, 1] - center[1]
    ddsq = dd0 * dd0 + dd1 * dd1
    if kh == 0:
        u = (1 / np.pi) * np.log(ddsq)
    else:
        u = np.real((1j / 4) * hankel1(0, kh * np.sqrt(ddsq)))
    return u[:, None]


def bfield_constant(xx, kh):
    """c(x) = -kh^2 (src_disc/built_in_funcs.py:79-80)."""
    return -(kh ** 2) * np.ones(xx.shape[0])


def lshaped_waveguide_bumps(xx, amp=1.0, width=0.005, spacing=0.04, removed_row=None, removed_col=None):
    """Brute-force sum of the benchmark's bump lattice (a field of the
    benchmark workloads, not of the reference): amp * exp(-|x - x_k|^2 /
    width^2) over every lattice point x_k = spacing * ((i, j) + 1/2), i, j <
    1/spacing, except the points (i, rr) with i >= cc and (cc, j) with j >= rr
    (rr = cc = (1/spacing) // 2 by default), which leave an L-shaped channel.
    Sums over all lattice points with no neighbourhood cut-off."""
    xx = np.asarray(xx, dtype=np.float64)
    n = int(round(1.0 / spacing))
    rr = n // 2 if removed_row is None else removed_row
    cc = n // 2 if removed_col is None else removed_col
    b = np.zeros(xx.shape[0])
    for i in range(n):
        for j in range(n):
            if (j == rr and i >= cc) or (i == cc and j >= rr):
                continue
            cx, cy = (i + 0.5) * spacing, (j + 0.5) * spacing
            b += amp * np.exp(-((xx[:, 0] - cx) ** 2 + (xx[:, 1] - cy) ** 2) / width ** 2)
    return b


def bfield_lshaped_waveguide(xx, kh):
    """c(x) = -kh^2 (1 - bumps(x)), the coefficient of the waveguide workload."""
    return -(kh ** 2) * (1.0 - lshaped_waveguide_bumps(xx))


def kh_from_ppw(n, ppw):
    """argparse_driver.py:86-90."""
    nwaves = int(n / ppw)
    return (nwaves + 0.03) * 2 * np.pi + 1.8


class PDO:
    """Coefficient callables of -c11 u_xx - c22 u_yy - 2 c12 u_xy + c1 u_x
    + c2 u_y + c u (src_disc/pdo.py:6-12); each xx (M,2) -> (M,) or None."""

    def __init__(self, c11, c22, c12=None, c1=None, c2=None, c=None):
        self.c11, self.c22, self.c12 = c11, c22, c12
        self.c1, self.c2, self.c = c1, c2, c


def helmholtz_pdo(kh, bfield=bfield_constant):
    ones = lambda xx: np.ones(xx.shape[0])
    return PDO(ones, ones, c=lambda xx: np.asarray(bfield(xx, kh)).reshape(-1))


# ----------------------------------------------------------------------------
# 3. Multidomain HPS (src_disc/hps_multidomain_disc.py, hps_parallel_leaf_ops)
# ----------------------------------------------------------------------------


class HPSOracle:
    """HPS_Multidomain restated (src_disc/hps_multidomain_disc.py:17-250)."""

    def __init__(self, pdo, domain, a, p, periodic=False):
        self.pdo, self.p, self.a = pdo, p, a
        self.periodic = periodic
        self.domain = np.asarray(domain, dtype=np.float64)
        n = (self.domain[:, 1] - self.domain[:, 0]) / (2 * a)
        self.n = np.round(n).astype(int)
        self.n0, self.n1 = int(self.n[0]), int(self.n[1])
        self.nboxes = self.n0 * self.n1
        (self.zz, self.Ds, self.JJ, self.hmin,
         self.D, self.Dsq) = leaf_discretization_2d(a, p)
        self.Nx = diff_ops(self.Ds, self.JJ)
        self.Interp_mat = interp_mat(a, p, self.zz)
        self.grid_xx = self.get_grid()
        Jx = self.JJ["Jx"]
        grid_ext = self.grid_xx[:, Jx, :].reshape(-1, 2)
        self.I_unique, self.I_copy1, self.I_copy2 = self.get_unique_inds()
        self.xx_active = grid_ext[self.I_unique]
        self.xx_tot = self.grid_xx.reshape(-1, 2)

    def get_grid(self):
        """Leaf node coordinates, box = i*n1 + j (hps_multidomain_disc.py:111-124)."""
        zz = self.zz.T
        p, a = self.p, self.a
        xx = np.zeros((self.nboxes, p ** 2, 2))
        for i in range(self.n0):
            for j in range(self.n1):
                box = i * self.n1 + j
                zl = zz.copy()
                zl[:, 0] += a + 2 * a * i + self.domain[0, 0]
                zl[:, 1] += a + 2 * a * j + self.domain[1, 0]
                xx[box] = zl
        return xx

    def get_unique_inds(self):
        """Shared-face de-duplication (hps_multidomain_disc.py:126-182)."""
        f = self.p - 2
        n0, n1 = self.n0, self.n1
        box_ind = np.arange(n0 * n1 * 4 * f).reshape(n0, n1, 4 * f)
        n_uni = f * (n0 * (2 * n1 + 1) + n1)
        n_rep = f * (n0 * (n1 - 1) + n1 * (n0 - 1))
        I_unique = np.zeros(n_uni, dtype=np.int64)
        I_copy1 = np.zeros(n_rep, dtype=np.int64)
        I_copy2 = np.zeros(n_rep, dtype=np.int64)
        ou = 0
        oc = 0
        I_unique[ou:ou + n1 * f] = box_ind[0, :, :f].flatten()
        ou += n1 * f
        for i in range(n0):
            if i > 0:
                lrep = box_ind[i, :, :f].flatten()
                rrep = box_ind[i - 1, :, f:2 * f].flatten()
                I_unique[ou:ou + n1 * f] = lrep
                ou += n1 * f
                I_copy1[oc:oc + n1 * f] = lrep
                I_copy2[oc:oc + n1 * f] = rrep
                oc += n1 * f
            I_unique[ou:ou + f] = box_ind[i, 0, 2 * f:3 * f]
            ou += f
            urep = box_ind[i, :n1 - 1, 3 * f:].flatten()
            drep = box_ind[i, 1:, 2 * f:3 * f].flatten()
            I_unique[ou:ou + (n1 - 1) * f] = drep
            ou += (n1 - 1) * f
            I_copy1[oc:oc + (n1 - 1) * f] = drep
            I_copy2[oc:oc + (n1 - 1) * f] = urep
            oc += (n1 - 1) * f
            I_unique[ou:ou + f] = box_ind[i, n1 - 1, 3 * f:]
            ou += f
        I_unique[ou:ou + n1 * f] = box_ind[n0 - 1, :, f:2 * f].flatten()
        return I_unique, I_copy1, I_copy2

    # -- leaf operators ------------------------------------------------------
    def aloc(self, boxes):
        """Dense collocation operator per leaf, rows scaled by the
        coefficient at the row node (hps_parallel_leaf_ops.py:10-37).  The
        accumulation order (c11, c22, c12, c1, c2, c) follows the reference."""
        p = self.p
        xx = self.grid_xx[boxes].reshape(-1, 2)
        nb = len(boxes)
        A = np.zeros((nb, p ** 2, p ** 2))
        pdo = self.pdo

        def acc(func, Dm, c=1.0):
            fv = np.asarray(func(xx), dtype=np.float64).reshape(nb, p ** 2)[:, :, None]
            fv = fv * c
            A[...] += fv * Dm[None]

        acc(pdo.c11, self.Ds["D11"], -1.0)
        acc(pdo.c22, self.Ds["D22"], -1.0)
        if pdo.c12 is not None:
            acc(pdo.c12, self.Ds["D12"], -2.0)
        if pdo.c1 is not None:
            acc(pdo.c1, self.Ds["D1"])
        if pdo.c2 is not None:
            acc(pdo.c2, self.Ds["D2"])
        c_nodes = getattr(self, "c_nodes", None)
        if c_nodes is not None:
            # per-node values of c (nboxes, p^2) in place of the callable
            A[...] += np.asarray(c_nodes)[boxes][:, :, None] * np.eye(p ** 2)[None]
        elif pdo.c is not None:
            acc(pdo.c, np.eye(p ** 2))
        return A

    def dtns(self, boxes=None, chunk=256):
        """Leaf DtN maps (hps_parallel_leaf_ops.py:50-72, mode 'build')."""
        if boxes is None:
            boxes = np.arange(self.nboxes)
        Jc, Jx, Jxr = self.JJ["Jc"], self.JJ["Jx"], self.JJ["Jxreorder"]
        Jtot = np.hstack((Jc, Jx))
        nx = len(Jx)
        out = np.zeros((len(boxes), nx, nx))
        for s in range(0, len(boxes), chunk):
            bb = boxes[s:s + chunk]
            A = self.aloc(bb)
            Acc = A[:, Jc][:, :, Jc]
            if self.pdo.c12 is None:
                S = -np.linalg.solve(Acc, A[:, Jc][:, :, Jx])
            else:
                S = -np.linalg.solve(Acc, A[:, Jc][:, :, Jxr]) @ self.Interp_mat[None]
            Sfull = np.concatenate((S, np.broadcast_to(np.eye(nx), (len(bb), nx, nx))), axis=1)
            out[s:s + chunk] = self.Nx[:, Jtot][None] @ Sfull
        return out

    def iext_box(self):
        """Global face-DOF ids of each leaf's [l,r,d,u] exterior
        (hps_multidomain_disc.py:63-81)."""
        f = self.p - 2
        n0, n1 = self.n0, self.n1
        I = np.tile(np.arange(f), (n0 * n1, 4))
        for i in range(n0):
            off = (2 * n1 + 1) * i
            for j in range(n1):
                b = i * n1 + j
                I[b, :f] += (off + j) * f
                # periodic in x: the last column's right faces are the
                # domain's left boundary faces (slablu_solver.py:91-104)
                rface = j if (self.periodic and i == n0 - 1) else off + 2 * n1 + 1 + j
                I[b, f:2 * f] += rface * f
                I[b, 2 * f:3 * f] += (off + n1 + j) * f
                I[b, 3 * f:] += (off + n1 + j + 1) * f
        return I

    def sparse_mat(self, DtN=None):
        """Scatter-add of leaf DtNs into scipy CSR (hps_multidomain_disc.py:53-109)."""
        if DtN is None:
            DtN = self.dtns()
        I = self.iext_box()
        nb, ne = I.shape
        rows = np.repeat(I, ne, axis=1).reshape(nb, ne, ne)
        cols = np.transpose(rows, (0, 2, 1))
        f = self.p - 2
        nface = f * (self.n0 * (2 * self.n1 + 1) + self.n1)
        A = scipy.sparse.coo_matrix((DtN.flatten(), (rows.flatten(), cols.flatten())), shape=(nface, nface)).tocsr()
        return A

    def leaf_solve(self, uu_sol, DtN_unused=None, ff_body_func=None, chunk=256):
        """Leaf recovery from the unique face solution
        (hps_multidomain_disc.py:227-243; hps_parallel_leaf_ops.py:73-92).
        Returns (flat solution on all leaf nodes (nboxes*p^2, nrhs), leaf
        residual norm)."""
        p = self.p
        f = p - 2
        ne = 4 * f
        nrhs = uu_sol.shape[-1]
        bnd = np.zeros((self.nboxes * ne, nrhs))
        bnd[self.I_unique] = uu_sol
        bnd[self.I_copy2] = bnd[self.I_copy1]
        bnd = bnd.reshape(self.nboxes, ne, nrhs)
        Jc, Jx, Jxr = self.JJ["Jc"], self.JJ["Jx"], self.JJ["Jxreorder"]
        out = np.zeros((self.nboxes, p ** 2, 2 * nrhs))
        for s in range(0, self.nboxes, chunk):
            bb = np.arange(s, min(s + chunk, self.nboxes))
            A = self.aloc(bb)
            Acc = A[:, Jc][:, :, Jc]
            fb = np.zeros((len(bb), len(Jc), nrhs))
            if ff_body_func is not None:
                fb = ff_body_func(self.grid_xx[bb].reshape(-1, 2)).reshape(len(bb), p ** 2, nrhs)[:, Jc]
            u = np.zeros((len(bb), p ** 2, 2 * nrhs))
            u[:, Jxr, :nrhs] = self.Interp_mat[None] @ bnd[bb]
            if self.pdo.c12 is None:
                u[:, Jc, :nrhs] = np.linalg.solve(Acc, fb - A[:, Jc][:, :, Jx] @ bnd[bb])
            else:
                u[:, Jc, :nrhs] = np.linalg.solve(Acc, fb - A[:, Jc][:, :, Jxr] @ u[:, Jxr, :nrhs])
            u[:, Jc, nrhs:] = A[:, Jc] @ u[:, :, :nrhs] - fb
            out[bb] = u
        res = float(np.linalg.norm(out[..., nrhs:]))
        return out[..., :nrhs].reshape(-1, nrhs), res


# ----------------------------------------------------------------------------
# 4. Slab partition (src_solver/slablu_solver.py:35-202)
# ----------------------------------------------------------------------------


def get_nearest_div(n, x):
    """Divisor of n nearest to x, ties -> larger (slablu_solver.py:35-46)."""
    factors = list(reduce(list.__add__, ([i, n // i] for i in range(1, int(n ** 0.5) + 1) if n % i == 0)))
    best, dist = n, abs(n - x)
    for f in factors:
        d = abs(f - x)
        if d < dist:
            best, dist = f, d
        elif d == dist:
            best = max(f, best)
    return int(best)


def slab_buf(n0, n1, p, buf_constant):
    """Leaf columns per slab (slablu_solver.py:127-143)."""
    f = p - 2
    npan_max = max(n0, n1)
    n_tmp = npan_max * f - 1
    buf_points = int(n_tmp ** (2 / 3) * buf_constant)
    buf_points = min(400, buf_points)
    buf = max(int(buf_points / f) + 1, 2)
    return get_nearest_div(n0, buf)


class SlabPartition:
    """inds_pans, I_L, I_R, I_D, I_U, I_C(elim, rem), I_slabC, I_slabX,
    I_Ctot, I_Xtot for the non-periodic HPS case
    (slablu_solver.py:61-124, 127-202)."""

    def __init__(self, n0, n1, p, buf_constant, periodic=False):
        f = p - 2
        self.periodic = periodic
        self.f, self.n0, self.n1 = f, n0, n1
        buf = slab_buf(n0, n1, p, buf_constant)
        Npan = n0 // buf
        self.buf, self.Npan = buf, Npan
        nfaces_pan = (2 * n1 + 1) * buf + n1
        self.inds_pans = np.stack([np.arange(nfaces_pan * f) + (2 * n1 + 1) * buf * j * f
                                   for j in range(Npan)])
        self.elim_nblocks, self.elim_bs = buf - 1, f
        self.rem_nblocks, self.rem_bs = n1 - 1, buf * f
        self.I_L = np.arange(n1 * f)
        self.I_R = np.arange(n1 * f) + (2 * n1 + 1) * buf * f
        I_elim = np.zeros((n1, buf - 1, f), dtype=np.int64)
        I_rem = np.zeros((n1 - 1, buf, f), dtype=np.int64)
        I_D = np.zeros((buf, f), dtype=np.int64)
        I_U = np.zeros((buf, f), dtype=np.int64)
        for b in range(buf):
            bo = (2 * n1 + 1) * b
            I_D[b] = np.arange(f) + (bo + n1) * f
            I_U[b] = np.arange(f) + (bo + 2 * n1) * f
            for j in range(1, n1):
                I_rem[j - 1, b] = np.arange(f) + (bo + n1 + j) * f
            if b > 0:
                for j in range(n1):
                    I_elim[j, b - 1] = np.arange(f) + (bo + j) * f
        self.I_elim_loc = I_elim.reshape(n1, -1)
        self.I_rem_loc = I_rem.reshape(n1 - 1, -1)
        self.I_D = I_D.flatten()
        self.I_U = I_U.flatten()
        self.I_C = np.hstack((I_elim.flatten(), I_rem.flatten()))
        I_Ldir = self.inds_pans[0, self.I_L]
        I_Rdir = self.inds_pans[Npan - 1, self.I_R].copy()
        I_Ddir = self.inds_pans[:, self.I_D].flatten()
        I_Udir = self.inds_pans[:, self.I_U].flatten()
        self.I_seam_copy = I_Rdir       # face ids of the domain's right boundary
        if periodic:
            # the last slab's right faces are slab 0's left faces; every slab's
            # left faces are an interface (slablu_solver.py:91-123)
            if Npan < 2:
                raise ValueError("periodic slabs need at least 2 slabs")
            self.inds_pans[Npan - 1, self.I_R] = self.inds_pans[0, self.I_L]
            self.I_slabC = self.inds_pans[:, self.I_C].flatten()
            self.I_slabX = self.inds_pans[:, self.I_L].flatten()
            self.I_Xtot = np.hstack((I_Ddir, I_Udir))
        else:
            self.I_slabC = self.inds_pans[:, self.I_C].flatten()
            self.I_slabX = self.inds_pans[1:, self.I_L].flatten()
            self.I_Xtot = np.hstack((I_Ldir, I_Rdir, I_Ddir, I_Udir))
        self.I_Ctot = np.hstack((self.I_slabC, self.I_slabX))


def dof_count(p, n0, n1):
    """N = (p-2)(p n0 n1 + n0 + n1)  (argparse_driver.py:163)."""
    return (p - 2) * (p * n0 * n1 + n0 + n1)


# ----------------------------------------------------------------------------
# 5. Exact discrete solve (the reference's `--solver superLU` path,
#    slablu_solver.py:360-385, 489-526, 601-666)
# ----------------------------------------------------------------------------


def get_rhs(A, part, xx_active, uu_dir_func):
    """ff = -A[Ctot, Xtot] u_dir  (slablu_solver.py:489-526, no body load)."""
    uu_dir = uu_dir_func(xx_active[part.I_Xtot])
    ff = -(A[part.I_Ctot][:, part.I_Xtot] @ uu_dir)
    return np.asarray(ff), uu_dir


def superlu_solve(A, part, ff):
    """x = A_CC^{-1} ff with SuperLU (slablu_solver.py:360-385, 601-620);
    returns (x, relative residual)."""
    A_CC = A[part.I_Ctot][:, part.I_Ctot].tocsc()
    LU = spla.splu(A_CC)
    x = LU.solve(ff)
    res = A_CC @ x - ff
    return x, float(np.linalg.norm(res) / np.linalg.norm(ff))


def full_solve(hps, part, uu_dir_func, known_sol=True, A=None, DtN=None, x_C=None):
    """Domain_Solver.solve restated (slablu_solver.py:623-666) with the exact
    SuperLU interior solve.  Returns dict(sol_tot, res, true_err, resloc)."""
    if A is None:
        A = hps.sparse_mat(DtN)
    ff, uu_dir = get_rhs(A, part, hps.xx_active, uu_dir_func)
    if x_C is None:
        x_C, res = superlu_solve(A, part, ff)
    else:
        res = float(np.linalg.norm(A[part.I_Ctot][:, part.I_Ctot] @ x_C - ff) / np.linalg.norm(ff))
    sol_face = np.zeros((A.shape[0], ff.shape[1]))
    sol_face[part.I_Ctot] = x_C
    sol_face[part.I_Xtot] = uu_dir
    if getattr(part, "periodic", False):
        # the right-boundary face slots carry the seam (left-boundary) values
        sol_face[part.I_seam_copy] = sol_face[part.inds_pans[0, part.I_L]]
    sol_tot, resloc = hps.leaf_solve(sol_face)
    true_err = float("nan")
    if known_sol:
        uu_true = uu_dir_func(hps.xx_tot.copy())
        true_err = float(np.linalg.norm(sol_tot - uu_true) / np.linalg.norm(uu_true))
    return dict(sol_tot=sol_tot, sol_face=sol_face, x_C=x_C, res=res, true_err=true_err,
                resloc=resloc, ff=ff, A=A)


# ----------------------------------------------------------------------------
# 6. Two-level slab algorithm restated (block_tridiag_solver_dense.py,
#    block_tridiag_solver_structured.py, slablu_solver.py:697-1308) with the
#    interface Schur complement formed exactly (dense) instead of by HBS.
# ----------------------------------------------------------------------------


def btri_factor(diag, sub, sup):
    """Block LU without inter-block pivoting (block_tridiag_solver_dense.py:152-177).

    diag (b,m,bs,bs), sub/sup (b,m-1,bs,bs).  Returns (LU list, piv list,
    Lfac, Ufac) per batch with LU_i = getrf(D_i - Lfac_i Ufac_i),
    Ufac_i = L_{i-1}^{-1} P_{i-1}^T sup_{i-1}, Lfac_i = sub_{i-1} U_{i-1}^{-1}."""
    b, m = diag.shape[:2]
    LU = np.zeros_like(diag)
    piv = np.zeros(diag.shape[:3], dtype=np.int64)
    Lf = np.zeros_like(sub)
    Uf = np.zeros_like(sup)
    for k in range(b):
        lu, pv = scipy.linalg.lu_factor(diag[k, 0])
        LU[k, 0], piv[k, 0] = lu, pv
        for i in range(1, m):
            U = np.triu(lu)
            Uf[k, i - 1] = scipy.linalg.solve_triangular(lu, _apply_piv(pv, sup[k, i - 1]),
                                                         lower=True, unit_diagonal=True)
            Lf[k, i - 1] = scipy.linalg.solve_triangular(U, sub[k, i - 1].T, trans='T', lower=False).T
            lu, pv = scipy.linalg.lu_factor(diag[k, i] - Lf[k, i - 1] @ Uf[k, i - 1])
            LU[k, i], piv[k, i] = lu, pv
    return LU, piv, Lf, Uf


def _apply_piv(piv, B):
    """Row interchanges of LAPACK getrf (0-based scipy pivots) applied to B."""
    B = np.array(B, copy=True)
    for i, pi in enumerate(piv):
        if pi != i:
            B[[i, pi]] = B[[pi, i]]
    return B


def _apply_piv_T(piv, B):
    B = np.array(B, copy=True)
    for i in range(len(piv) - 1, -1, -1):
        pi = piv[i]
        if pi != i:
            B[[i, pi]] = B[[pi, i]]
    return B


def btri_solve(fac, f, transpose=False):
    """Forward/backward block sweeps (block_tridiag_solver_dense.py:179-222).
    f (b, m*bs, nrhs) -> solution of the same shape."""
    LU, piv, Lf, Uf = fac
    b, m, bs = LU.shape[:3]
    x = np.array(f, dtype=np.float64).reshape(b, m, bs, -1)
    for k in range(b):
        if not transpose:
            x[k, 0] = scipy.linalg.solve_triangular(LU[k, 0], _apply_piv(piv[k, 0], x[k, 0]), lower=True, unit_diagonal=True)
            for i in range(1, m):
                t = x[k, i] - Lf[k, i - 1] @ x[k, i - 1]
                x[k, i] = scipy.linalg.solve_triangular(LU[k, i], _apply_piv(piv[k, i], t), lower=True, unit_diagonal=True)
            x[k, m - 1] = scipy.linalg.solve_triangular(LU[k, m - 1], x[k, m - 1], lower=False)
            for i in range(m - 2, -1, -1):
                t = x[k, i] - Uf[k, i] @ x[k, i + 1]
                x[k, i] = scipy.linalg.solve_triangular(LU[k, i], t, lower=False)
        else:
            x[k, 0] = scipy.linalg.solve_triangular(LU[k, 0], x[k, 0], trans='T', lower=False)
            for i in range(1, m):
                t = x[k, i] - Uf[k, i - 1].T @ x[k, i - 1]
                x[k, i] = scipy.linalg.solve_triangular(LU[k, i], t, trans='T', lower=False)
            x[k, m - 1] = _apply_piv_T(piv[k, m - 1], scipy.linalg.solve_triangular(
                LU[k, m - 1], x[k, m - 1], trans='T', lower=True, unit_diagonal=True))
            for i in range(m - 2, -1, -1):
                t = x[k, i] - Lf[k, i].T @ x[k, i + 1]
                x[k, i] = _apply_piv_T(piv[k, i], scipy.linalg.solve_triangular(
                    LU[k, i], t, trans='T', lower=True, unit_diagonal=True))
    return x.reshape(b, m * bs, -1)


def extract_btri(M, m, bs):
    """Dense (m*bs)^2 block-tridiagonal matrix -> diag, sub, sup blocks (the
    quantities recover_block_tridiag_via_sampling reads off by probing,
    block_tridiag_solver_dense.py:9-77)."""
    diag = np.stack([M[i * bs:(i + 1) * bs, i * bs:(i + 1) * bs] for i in range(m)])
    sub = np.stack([M[(i + 1) * bs:(i + 2) * bs, i * bs:(i + 1) * bs] for i in range(m - 1)]) if m > 1 \
        else np.zeros((0, bs, bs))
    sup = np.stack([M[i * bs:(i + 1) * bs, (i + 1) * bs:(i + 2) * bs] for i in range(m - 1)]) if m > 1 \
        else np.zeros((0, bs, bs))
    return diag, sub, sup


class SlabLUOracle:
    """Two-level SlabLU build + solve restated on the CPU.

    build: Aloc (elim) block-tridiagonal LU per leaf row
    (slablu_solver.py:718-745), Ared (rem Schur complement) block-tridiagonal
    LU per slab (slablu_solver.py:768-882), slab Schur complements T_slab on
    [L;R] formed densely (the quantity get_TXX_slabind samples,
    slablu_solver.py:1066-1211), interface block-tridiagonal Schur recursion
    (block_tridiag_solver_structured.py:34-76).  solve: slablu_solver.py:528-599.
    """

    def __init__(self, A, part):
        self.A = A.tocsr()
        self.part = part

    def _slab_sets(self, s):
        P = self.part
        xoff = P.elim_nblocks * P.elim_bs * P.n1
        I_C = P.inds_pans[s, P.I_C]
        I_elim = P.inds_pans[s, P.I_C[:xoff]]
        I_rem = P.inds_pans[s, P.I_C[xoff:]]
        I_L = P.inds_pans[s, P.I_L]
        I_R = P.inds_pans[s, P.I_R]
        return I_C, I_elim, I_rem, I_L, I_R

    def build(self):
        P = self.part
        self.elim_fac, self.rem_fac, self.Tslab = [], [], []
        self.elim_blocks, self.rem_blocks = [], []
        for s in range(P.Npan):
            self.build_slab(s)
        return self.build_interface()

    def build_slab(self, s):
        """Elim LU, rem Schur complement + LU and dense T_slab of slab s."""
        P, A = self.part, self.A
        m_e, bs_e = P.elim_nblocks, P.elim_bs
        m_r, bs_r = P.rem_nblocks, P.rem_bs
        if not hasattr(self, "elim_fac"):
            self.elim_fac, self.rem_fac, self.Tslab = [], [], []
            self.elim_blocks, self.rem_blocks = [], []
        I_C, I_elim, I_rem, I_L, I_R = self._slab_sets(s)
        Aee = A[I_elim][:, I_elim].toarray()
        d, sb, sp = [], [], []
        for j in range(P.n1):
            blk = Aee[j * m_e * bs_e:(j + 1) * m_e * bs_e, j * m_e * bs_e:(j + 1) * m_e * bs_e]
            dd, ss, pp = extract_btri(blk, m_e, bs_e)
            d.append(dd); sb.append(ss); sp.append(pp)
        d, sb, sp = np.stack(d), np.stack(sb), np.stack(sp)
        self.elim_blocks.append((d.copy(), sb.copy(), sp.copy()))
        efac = btri_factor(d, sb, sp)
        self.elim_fac.append(efac)
        # rem Schur complement (exact)
        Are = A[I_rem][:, I_elim].toarray()
        Aer = A[I_elim][:, I_rem].toarray()
        Arr = A[I_rem][:, I_rem].toarray()
        Y = btri_solve(efac, Aer.reshape(P.n1, m_e * bs_e, -1)).reshape(len(I_elim), -1)
        Sr = Arr - Are @ Y
        dr, sbr, spr = extract_btri(Sr, m_r, bs_r)
        self.rem_blocks.append((dr.copy(), sbr.copy(), spr.copy()))
        self.rem_fac.append(btri_factor(dr[None], sbr[None], spr[None]))
        # slab Schur complement on X = [L; R]
        I_X = np.hstack((I_L, I_R))
        ACX = A[I_C][:, I_X].toarray()
        V = self.slab_solve(s, ACX)
        AXC = A[I_X][:, I_C].toarray()
        self.Tslab.append(-(AXC @ V))

    def build_interface(self):
        """Interface system (block_tridiag_solver_structured.py:34-76)."""
        P, A = self.part, self.A
        L = P.n1 * P.f
        nI = P.Npan - 1
        self.T_diag = np.zeros((nI, L, L))
        self.T_sub = np.zeros((max(nI - 1, 0), L, L))
        self.T_sup = np.zeros((max(nI - 1, 0), L, L))
        for j in range(nI):
            I_j = P.inds_pans[j + 1, P.I_L]
            self.T_diag[j] = A[I_j][:, I_j].toarray() + self.Tslab[j][L:, L:] + self.Tslab[j + 1][:L, :L]
            if j < nI - 1:
                self.T_sup[j] = self.Tslab[j + 1][:L, L:]
                self.T_sub[j] = self.Tslab[j + 1][L:, :L]
        self.S_lu = []
        for j in range(nI):
            S = self.T_diag[j].copy()
            if j > 0:
                S -= self.T_sub[j - 1] @ scipy.linalg.lu_solve(self.S_lu[j - 1], self.T_sup[j - 1])
            self.S_lu.append(scipy.linalg.lu_factor(S))
        return self

    def slab_solve(self, s, f, transpose=False):
        """A_CC(slab s)^{-1} f via elim -> rem -> elim (slablu_solver.py:901-943)."""
        P, A = self.part, self.A
        I_C, I_elim, I_rem, I_L, I_R = self._slab_sets(s)
        ne = len(I_elim)
        m_e, bs_e = P.elim_nblocks, P.elim_bs
        f = np.asarray(f, dtype=np.float64)
        fe, fr = f[:ne], f[ne:]
        Aer = A[I_elim][:, I_rem]
        Are = A[I_rem][:, I_elim]
        if transpose:
            Aer, Are = Are.T, Aer.T
        t = btri_solve(self.elim_fac[s], fe.reshape(P.n1, m_e * bs_e, -1), transpose).reshape(ne, -1)
        g = fr - Are @ t
        ur = btri_solve(self.rem_fac[s], g[None], transpose)[0]
        ue = btri_solve(self.elim_fac[s], (fe - Aer @ ur).reshape(P.n1, m_e * bs_e, -1), transpose).reshape(ne, -1)
        return np.vstack((ue, ur))

    def tred_solve(self, f):
        """Interface block-tridiagonal solve (block_tridiag_solver_structured.py:157-187)."""
        P = self.part
        L = P.n1 * P.f
        nI = P.Npan - 1
        x = np.array(f, dtype=np.float64).reshape(nI, L, -1)
        # forward: z_j = S_j^{-1}(f_j - sub_{j-1} z_{j-1})
        z = np.zeros_like(x)
        for j in range(nI):
            r = x[j] if j == 0 else x[j] - self.T_sub[j - 1] @ z[j - 1]
            z[j] = scipy.linalg.lu_solve(self.S_lu[j], r)
        u = np.zeros_like(x)
        u[nI - 1] = z[nI - 1]
        for j in range(nI - 2, -1, -1):
            u[j] = z[j] - scipy.linalg.lu_solve(self.S_lu[j], self.T_sup[j] @ u[j + 1])
        return u.reshape(nI * L, -1)

    def solve(self, ff):
        """Two-level solve of A_CC x = ff (slablu_solver.py:543-599)."""
        P, A = self.part, self.A
        nC = len(P.I_C)
        Np = P.Npan
        ff = np.asarray(ff, dtype=np.float64)
        yC = np.vstack([self.slab_solve(s, ff[s * nC:(s + 1) * nC]) for s in range(Np)])
        ffX = ff[Np * nC:] - A[P.I_slabX][:, P.I_slabC] @ yC
        solX = self.tred_solve(ffX)
        rhsC = ff[:Np * nC] - A[P.I_slabC][:, P.I_slabX] @ solX
        solC = np.vstack([self.slab_solve(s, rhsC[s * nC:(s + 1) * nC]) for s in range(Np)])
        return np.vstack((solC, solX))


def setup_problem(n, p, ppw=None, kh=None, domain=((0.0, 1.0), (0.0, 1.0)), bfield=bfield_constant,
                  buf_constant=1.0, periodic=False):
    """Driver parameter derivation (argparse_driver.py:86-163); periodic:
    x-periodic, Dirichlet data on the bottom and top boundaries only."""
    if kh is None:
        kh = kh_from_ppw(n, ppw)
    npan = n / (p - 2)
    a = 1 / (2 * npan)
    hps = HPSOracle(helmholtz_pdo(kh, bfield), np.asarray(domain), a, p, periodic=periodic)
    part = SlabPartition(hps.n0, hps.n1, p, buf_constant, periodic=periodic)
    return hps, part, kh


# ----------------------------------------------------------------------------
# 7. Adjoint of the FP64 solve with respect to a per-node coefficient
#    (the north star's adjoint form; the reference has the transposed slab
#    sweeps it needs, block_tridiag_solver_dense.py:205-222 and
#    slablu_solver.py:900-943, but no gradient)
# ----------------------------------------------------------------------------


def helmholtz_nodes_solve(hps, part, kh, b_nodes, uu_dir_func):
    """-Lap u - kh^2 b u = 0 with b given per leaf node (nboxes, p^2):
    exact discrete solution at every leaf node (nboxes*p^2,)."""
    hps.c_nodes = -(kh ** 2) * np.asarray(b_nodes, dtype=np.float64).reshape(hps.nboxes, hps.p ** 2)
    r = full_solve(hps, part, uu_dir_func, known_sol=False)
    return r["sol_tot"][:, 0], r


def adjoint_gradient(hps, part, kh, b_nodes, uu_dir_func, obs, d):
    """dL/db for L = 1/2 ||u[obs] - d||^2 (obs: flat node indices, d their
    data), u from helmholtz_nodes_solve.  The un-condensed system is
    F = [A_cc u_c + A_cx u_x ; sum_b P_b^T (N_c u_c + N_x u_x)] = 0 over the
    leaf interiors and the face unknowns I_Ctot, with c = -kh^2 b only on the
    diagonal of A_cc (interior rows), so dF/db_i = -kh^2 u_i e_i and
    dL/db_i = kh^2 lam_i u_i at interior nodes, lam from F_u^T lam = dL/du:
      A_CC^T lam_C = g_C - (sum_b P_b^T S_b^T g_c)[I_Ctot],  S = A_cc^{-1} A_cx,
      lam_c = A_cc^{-T} (g_c - N_c^T lam_x)            (per leaf).
    The node gradient reaches the face unknowns through the corner-inclusive
    interpolation Interp_mat (face and corner values of the leaf nodes).
    Returns (gradient (nboxes*p^2,), loss, u)."""
    p = hps.p
    u, r = helmholtz_nodes_solve(hps, part, kh, b_nodes, uu_dir_func)
    A = r["A"]
    nb = hps.nboxes
    Jc, Jx, Jxr = hps.JJ["Jc"], hps.JJ["Jx"], hps.JJ["Jxreorder"]
    res = u[obs] - d
    loss = 0.5 * float(res @ res)
    g = np.zeros(nb * p * p)
    np.add.at(g, obs, res)
    g = g.reshape(nb, p * p)
    g_c = g[:, Jc]
    g_bnd = hps.Interp_mat.T @ g[:, Jxr].T                      # (4(p-2), nb)
    I = hps.iext_box()
    nface = A.shape[0]
    g_face = np.zeros(nface)
    np.add.at(g_face, I.ravel(), g_bnd.T.ravel())
    Aall = hps.aloc(np.arange(nb))
    Acc = Aall[:, Jc][:, :, Jc]
    Acx = Aall[:, Jc][:, :, Jx]
    S = np.linalg.solve(Acc, Acx)                                # (nb, ni, nx)
    z = np.einsum("bix,bi->bx", S, g_c)
    np.add.at(g_face, I.ravel(), -z.ravel())
    A_CC = A[part.I_Ctot][:, part.I_Ctot].tocsc()
    lam_C = spla.splu(A_CC).solve(g_face[part.I_Ctot], trans="T")
    lam_face = np.zeros(nface)
    lam_face[part.I_Ctot] = lam_C
    lam_x = lam_face[I]                                          # (nb, nx)
    N_c = hps.Nx[:, Jc]
    rhs = g_c - lam_x @ N_c
    lam_c = np.linalg.solve(np.transpose(Acc, (0, 2, 1)), rhs[:, :, None])[:, :, 0]
    u_c = u.reshape(nb, p * p)[:, Jc]
    grad = np.zeros((nb, p * p))
    grad[:, Jc] = (kh ** 2) * lam_c * u_c
    return grad.ravel(), loss, u


def uu_periodic_mode(xx, kh, m=2):
    """cos(2 pi m x) cos(k_y y), k_y^2 = kh^2 - (2 pi m)^2: an exact solution
    of -Lap u - kh^2 u = 0 that is 1-periodic in x (a closed form for the
    periodic known-solution check)."""
    xx = np.asarray(xx)
    ky = np.sqrt(kh ** 2 - (2 * np.pi * m) ** 2)
    return (np.cos(2 * np.pi * m * xx[:, 0]) * np.cos(ky * xx[:, 1]))[:, None]
