"""Leaf spectral discretization (host set-up, numpy FP64).

Same quantities and conventions as src_disc/hps_subdomain_disc.py:16-152 of
the reference: Chebyshev nodes on [-a, a]^2 with node id i*p + j (x index
i, y index j, both ascending), derivative matrices built from the scaled
1-D Chebyshev matrix D on the descending grid, face index sets with the
corners excluded, the outward normal-derivative rows Nx and the
face-to-corner-inclusive interpolation Interp_mat.  The GPU kernels consume
the 1-D factors D and Dsq = D @ D instead of the p^2 x p^2 Kronecker
products, which are kept here only for the API.
"""
from collections import namedtuple

import numpy as np
import numpy.polynomial.chebyshev as cheb_py
import scipy.linalg

Pdo_2d = namedtuple("Pdo_2d", ["c11", "c22", "c12", "c1", "c2", "c"])
Ds_2d = namedtuple("Ds_2d", ["D11", "D22", "D12", "D1", "D2"])
JJ_2d = namedtuple("JJ_2d", ["Jl", "Jr", "Jd", "Ju", "Jx", "Jc", "Jxreorder"])


def cheb(p):
    """Differentiation matrix on x_k = cos(pi k / p) (Trefethen, Spectral
    Methods in MATLAB, program cheb)."""
    x = np.cos(np.pi * np.arange(p + 1) / p)
    c = np.concatenate(([2.0], np.ones(p - 1), [2.0])) * (-1.0) ** np.arange(p + 1)
    X = np.tile(x[:, None], (1, p + 1))
    dX = X - X.T
    D = np.outer(c, 1.0 / c) / (dX + np.eye(p + 1))
    D = D - np.diag(D.sum(axis=1))
    return D, x


def get_loc_interp(x_cheb, x_cheb_nocorners, q):
    Vc = cheb_py.chebvander(x_cheb, q)
    Vnc = cheb_py.chebvander(x_cheb_nocorners, q)
    interp = np.linalg.lstsq(Vnc.T, Vc.T, rcond=None)[0].T
    return interp, np.linalg.norm(interp @ Vnc - Vc), np.linalg.cond(interp)


class HPS_Disc:
    """One leaf of half-width a and p Chebyshev points per direction (d = 2)."""

    def __init__(self, a, p, d=2):
        if d != 2:
            raise NotImplementedError("only 2-D leaves are provided")
        self.a, self.p, self.d = a, p, d
        self.zz, self.Ds, self.JJ, self.hmin, self.D, self.Dsq = HPS_Disc.leaf_discretization_2d(a, p)
        self.Nx = HPS_Disc.get_diff_ops(self.Ds, self.JJ, d)
        self._get_interp_mat()

    def _get_interp_mat(self, verbose=False):
        """Least-squares interpolation from the p-2 face points of an edge to
        the p-1 points including its first corner, over Chebyshev
        polynomials of degree q = p+5, p+15, ... (first q whose condition
        test passes); four copies, columns permuted from the [l, r, d, u]
        face order to the Jxreorder walk."""
        p, a = self.p, self.a
        x_cheb = self.zz[1, :p - 1] + a
        x_nc = self.zz[1, 1:p - 1] + a
        cond_min, cond_max, err_tol = 3.25, 3.5, 2e-8
        q = p + 5
        while True:
            loc, err, cond = get_loc_interp(x_cheb, x_nc, q)
            if cond < cond_min or (err > err_tol and cond < cond_max):
                break
            q += 10
        blk = scipy.linalg.block_diag(*([loc] * 4))
        f = p - 2
        perm = np.hstack((np.arange(f), np.arange(f) + 3 * f, np.arange(f)[::-1] + f, np.arange(f)[::-1] + 2 * f))
        self.Interp_mat = blk[:, np.argsort(perm)]
        if verbose:
            print("--Interp_mat required lstsqfit of q=%d, condition number %5.5f with error %5.5e" % (q, cond, err))

    @staticmethod
    def get_diff_ops(Ds, JJ, d=2):
        return np.concatenate((-Ds.D1[JJ.Jl], +Ds.D1[JJ.Jr], -Ds.D2[JJ.Jd], +Ds.D2[JJ.Ju]))

    @staticmethod
    def cheb_2d(a, p):
        D, xvec = cheb(p - 1)
        xvec = a * np.flip(xvec)
        D = (1 / a) * D
        I = np.eye(p)
        Dsq = D @ D
        Ds = Ds_2d(D1=-np.kron(D, I), D2=-np.kron(I, D), D11=np.kron(Dsq, I), D22=np.kron(I, Dsq),
                   D12=np.kron(D, D))
        zz = np.vstack((np.repeat(xvec, p), np.tile(xvec, p)))
        return zz, Ds, D, Dsq

    @staticmethod
    def leaf_discretization_2d(a, p):
        zz, Ds, D, Dsq = HPS_Disc.cheb_2d(a, p)
        f = p - 2
        ii = np.repeat(np.arange(p), p)  # x index of each node
        jj = np.tile(np.arange(p), p)    # y index of each node
        inner_i = (ii > 0) & (ii < p - 1)
        inner_j = (jj > 0) & (jj < p - 1)
        Jl = np.flatnonzero((ii == 0) & inner_j)
        Jr = np.flatnonzero((ii == p - 1) & inner_j)
        Jd = np.flatnonzero((jj == 0) & inner_i)
        Ju = np.flatnonzero((jj == p - 1) & inner_i)
        Jc = np.flatnonzero(inner_i & inner_j)
        Jx = np.concatenate((Jl, Jr, Jd, Ju))
        sw, nw, se, ne = 0, p - 1, (p - 1) * p, p * p - 1
        Jxreorder = np.hstack(([sw], Jl, [nw], Ju, [ne], Jr[::-1], [se], Jd[::-1]))
        assert len(Jc) == f * f and len(Jxreorder) == 4 * (p - 1)
        JJ = JJ_2d(Jl=Jl, Jr=Jr, Jd=Jd, Ju=Ju, Jx=Jx, Jc=Jc, Jxreorder=Jxreorder)
        hmin = zz[1, 1] - zz[0, 1]
        return zz, Ds, JJ, hmin, D, Dsq
