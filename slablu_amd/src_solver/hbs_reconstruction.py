"""HBS (hierarchically block separable) matrices recovered from random
samples, on the GPU (drop-in for src_solver/hbs_reconstruction.py).

An n x n HBS matrix (n padded with an identity block to m * bs, m a power of
two) is, level by level from the leaves (m nodes of bs rows) to the root
(two nodes of k rows):

    A_0 = blockdiag(D_0) + blockdiag(U_0) A_1 blockdiag(V_0)^T
    A_l = blockdiag(D_l) + blockdiag(U_l) A_{l+1} blockdiag(V_l)^T   (2k rows per node)
    A_L = D_root                                                     (2k x 2k)

with orthonormal k-column bases U, V on every node.  It is recovered from
Y = A Omega and Z = A^T Psi for random Omega, Psi of bs + k columns
(hbs_reconstruction.py:414-497): on each node, the null space N of its rows
of Omega cancels the node's own diagonal block, so the range of Y N is the
node's column basis U (likewise V from Z and Psi); the diagonal block comes
from the right pseudo-inverse of the node's samples with the parts inside
the bases removed; the samples are then projected onto the bases
(Omega' = V^T Omega, Y' = U^T (Y - D Omega), ...) and sibling nodes merged
for the next level, 3k samples per node.  The diagonal k x k blocks of every
node's D above the leaves are pushed into its children afterwards
(hbs_reconstruction.py:351-381), which changes the representation, not the
matrix.

All dense work runs on this library's batched kernels in FP64: products
(slab_gemm), null spaces and right pseudo-inverses by LU solves
(slab_getrf / slab_getrs) on [Omega_1 Omega_2] and on the normal
equations, orthonormal bases by Householder QR (slab_orth) where the
reference calls torch.linalg.qr and lstsq.  Every node of every matrix of
a batch is one batch entry of each launch.
"""
import numpy as np
import torch

from .. import hip


def get_nearest_div(n, bs):
    """The smallest divisor of n that is at least bs (n itself when bs >= n,
    one leaf block)."""
    bs = min(max(1, int(bs)), n)
    while n % bs > 0:
        bs += 1
    return bs


def _T(x):
    return x.transpose(-1, -2)


def _mm(A, B, alpha=1.0, out=None, beta=0.0):
    """op-free batched product of 3-D (batch, r, c) views (transposed views
    allowed); A or B may have batch 1 (broadcast)."""
    bsz = max(A.shape[0], B.shape[0])
    if out is None:
        out = torch.empty(bsz, A.shape[1], B.shape[2], dtype=torch.float64, device=A.device)
    hip.gemm(A, B, out, alpha=alpha, beta=beta)
    return out


def _flat(x):
    """(..., r, c) -> (batch, r, c) contiguous."""
    return x.reshape(-1, x.shape[-2], x.shape[-1]).contiguous()


def _lu_solve(G, R):
    """G^{-1} R for batched square G (copied) and R (batch, r, c)."""
    LU = G.contiguous().clone()
    ipiv = torch.empty(LU.shape[0], LU.shape[1], dtype=torch.int32, device=LU.device)
    hip.getrf(LU, ipiv)
    X = R.contiguous().clone()
    hip.getrs(LU, ipiv, X)
    return X


def null_right(O):
    """A basis (batch, s, s - r) of the null space of each full-row-rank
    O (batch, r, s): [-O_1^{-1} O_2; I] with O = [O_1 O_2]."""
    b, r, s = O.shape
    k = s - r
    X = _lu_solve(O[:, :, :r], O[:, :, r:])
    N = torch.empty(b, s, k, dtype=torch.float64, device=O.device)
    N[:, :r] = -X
    N[:, r:] = torch.eye(k, dtype=torch.float64, device=O.device)
    return N


def orth(M):
    """Orthonormal columns spanning range(M) for M (batch, r, k), r >= k
    (Householder QR, slab_orth; k columns even for a rank-deficient M)."""
    return hip.orth(M.contiguous().clone())


def pinv_right(Y, O):
    """Y pinv(O) = Y O^T (O O^T)^{-1} for Y (batch, p, s), O (batch or 1, r, s)
    of full row rank (hbs_reconstruction.py:408-412)."""
    G = _mm(O, _T(O))                              # (., r, r)
    C = _mm(O, _T(Y))                              # (b, r, p) = (Y O^T)^T
    if G.shape[0] != C.shape[0]:
        G = G.expand(C.shape[0], -1, -1)
    return _T(_lu_solve(G, C)).contiguous()        # (G^{-1} O Y^T)^T


class HBS_matrix:
    """b HBS matrices of size n, rank k per node (hbs_reconstruction.py:
    49-170): leaf blocks of bs rows (the divisor of n + n_pad nearest above
    2k), padded with identity rows so that the leaf count is a power of
    two."""

    def __init__(self, b, n, k, n_pad=0):
        self.b, self.n, self.k = b, n, k
        self.bs = get_nearest_div(n + n_pad, 2 * k)
        self.m = (n + n_pad) // self.bs
        self.n_pad = n_pad
        lev = np.log2(self.m)
        if abs(int(lev) - lev) > 1e-14:
            m2 = 2 ** (int(lev) + 1)
            self.n_pad += (m2 - self.m) * self.bs
            self.m = m2
        self.L = None
        self.levels = None     # [(U, V, D)] per level, leaves first: (b, nodes, rows, k|rows)
        self.D_root = None     # (b, 2k, 2k)

    @property
    def N(self):
        return self.n + self.n_pad

    def nbytes(self):
        t = sum(u.numel() + v.numel() + d.numel() for u, v, d in self.levels) + self.D_root.numel()
        return int(t) * 8

    # -- recovery --------------------------------------------------------------
    def reconstruct(self, Omega, Psi, Y, Z):
        """Omega, Psi (1, n, bs + k); Y = A Omega, Z = A^T Psi (b, n, bs + k),
        all FP64 on one device."""
        b, m, bs, k = self.b, self.m, self.bs, self.k
        ns = bs + k
        dev = Omega.device
        f64 = lambda t: t.to(dev, torch.float64)
        Omega, Psi, Y, Z = f64(Omega), f64(Psi), f64(Y), f64(Z)
        if not (Omega.shape[-1] == ns and Y.shape[-1] == ns and Psi.shape[-1] == ns and Z.shape[-1] == ns):
            raise ValueError("HBS reconstruction needs bs + k = %d samples per side" % ns)
        if m < 2:
            raise ValueError("HBS needs at least two leaf blocks (n=%d, bs=%d)" % (self.n, bs))
        if self.n_pad > 0:
            # the padding rows carry an identity block: their samples are the
            # test vectors themselves (hbs_reconstruction.py:74-83)
            g = torch.Generator(device=dev)
            g.manual_seed(12345)
            Op = torch.rand(1, self.n_pad, ns, generator=g, dtype=torch.float64, device=dev)
            Pp = torch.rand(1, self.n_pad, ns, generator=g, dtype=torch.float64, device=dev)
            Omega = torch.cat((Omega, Op), dim=1)
            Psi = torch.cat((Psi, Pp), dim=1)
            Y = torch.cat((Y, Op.expand(b, -1, -1)), dim=1)
            Z = torch.cat((Z, Pp.expand(b, -1, -1)), dim=1)
        # leaf level: Omega / Psi blocks are shared by the b matrices
        Ob = Omega.reshape(m, bs, ns)
        Pb = Psi.reshape(m, bs, ns)
        Yb = Y.reshape(b * m, bs, ns)
        Zb = Z.reshape(b * m, bs, ns)
        shared = True
        levels = []
        nodes, rows = m, bs
        while True:
            if nodes == 1:
                # root: (2k x 3k) samples of the top two nodes merged
                self.D_root = pinv_right(Yb, Ob if not shared else Ob.expand(b, -1, -1)).reshape(b, rows, rows)
                break
            U, V, D = self._node_decomp(Ob, Pb, Yb, Zb, k, shared, b, nodes)
            levels.append((U.reshape(b, nodes, rows, k), V.reshape(b, nodes, rows, k),
                           D.reshape(b, nodes, rows, rows)))
            # project the samples onto the bases and merge sibling nodes
            Obx = Ob.repeat(b, 1, 1) if shared else Ob
            Pbx = Pb.repeat(b, 1, 1) if shared else Pb
            Op = _mm(_T(V), Obx)                           # V^T Omega        (b*nodes, k, ns)
            Yp = _mm(_T(U), _mm(D, Obx, alpha=-1.0, out=Yb.clone(), beta=1.0))   # U^T (Y - D Omega)
            Pp = _mm(_T(U), Pbx)                           # U^T Psi
            Zp = _mm(_T(V), _mm(_T(D), Pbx, alpha=-1.0, out=Zb.clone(), beta=1.0))  # V^T (Z - D^T Psi)
            nodes //= 2
            rows = 2 * k
            Ob = Op[:, :, :3 * k].reshape(b * nodes, rows, 3 * k).contiguous()
            Yb = Yp[:, :, :3 * k].reshape(b * nodes, rows, 3 * k).contiguous()
            Pb = Pp[:, :, :3 * k].reshape(b * nodes, rows, 3 * k).contiguous()
            Zb = Zp[:, :, :3 * k].reshape(b * nodes, rows, 3 * k).contiguous()
            shared = False
        self.levels = levels
        self.L = len(levels)
        self._push_down_diagonals()

    @staticmethod
    def _node_decomp(Ob, Pb, Yb, Zb, k, shared, b, nodes):
        """Bases and diagonal block of every node (hbs_reconstruction.py:
        461-487): U = orth(Y null(Omega)), V = orth(Z null(Psi)),
        D = (I - UU^T) Y pinv(Omega) + UU^T ((I - VV^T) Z pinv(Psi))^T."""
        Pn = null_right(Ob)
        Qn = null_right(Pb)
        if shared:
            Pn = Pn.repeat(b, 1, 1)
            Qn = Qn.repeat(b, 1, 1)
            Obx, Pbx = Ob.repeat(b, 1, 1), Pb.repeat(b, 1, 1)
        else:
            Obx, Pbx = Ob, Pb
        U = orth(_mm(Yb, Pn))
        V = orth(_mm(Zb, Qn))
        Xr = pinv_right(Yb, Obx)
        Xc = pinv_right(Zb, Pbx)
        # (I - U U^T) Xr
        D = _mm(U, _mm(_T(U), Xr), alpha=-1.0, out=Xr.clone(), beta=1.0)
        # + U U^T ((I - V V^T) Xc)^T
        Wc = _mm(V, _mm(_T(V), Xc), alpha=-1.0, out=Xc.clone(), beta=1.0)
        _mm(U, _mm(_T(U), _T(Wc)), out=D, beta=1.0)
        return U, V, D

    def _push_down_diagonals(self):
        """Move the diagonal k x k blocks of every node above the leaves (and
        of the root) into its children: child D += U_c B V_c^T."""
        k = self.k
        U, V, D = self.levels[-1]
        b, nodes = U.shape[0], U.shape[1]
        blk = torch.stack((self.D_root[:, :k, :k], self.D_root[:, k:, k:]), dim=1).contiguous()
        self.D_root[:, :k, :k] = 0
        self.D_root[:, k:, k:] = 0
        self._add_child(U, V, D, blk)
        for l in range(self.L - 1, 0, -1):
            U, V, D = self.levels[l]
            nodes = U.shape[1]
            blk = torch.empty(b, 2 * nodes, k, k, dtype=torch.float64, device=U.device)
            blk[:, 0::2] = D[:, :, :k, :k]
            blk[:, 1::2] = D[:, :, k:, k:]
            D[:, :, :k, :k] = 0
            D[:, :, k:, k:] = 0
            Uc, Vc, Dc = self.levels[l - 1]
            self._add_child(Uc, Vc, Dc, blk)

    @staticmethod
    def _add_child(U, V, D, blk):
        b, nodes, rows, k = U.shape
        U3 = U.reshape(b * nodes, rows, k)
        V3 = V.reshape(b * nodes, rows, k)
        B3 = blk.reshape(b * nodes, k, k)
        D3 = D.view(b * nodes, rows, rows)
        _mm(U3, _mm(B3, _T(V3)), out=D3, beta=1.0)

    # -- products --------------------------------------------------------------
    def matvec(self, q, b_start=None, b_end=None, transpose=False, verbose=False):
        """A q (or A^T q) for q (b_end - b_start, n, nc) (hbs_reconstruction.py:
        134-147): one upward pass through the bases, the root block, one
        downward pass."""
        b0 = 0 if b_start is None else b_start
        b1 = self.b if b_end is None else b_end
        nb, m, bs, k = b1 - b0, self.m, self.bs, self.k
        q = q.to(torch.float64)
        nc = q.shape[-1]
        if self.n_pad > 0:
            q = torch.cat((q, torch.zeros(q.shape[0], self.n_pad, nc, dtype=q.dtype, device=q.device)), dim=1)
        sel = lambda t: t[b0:b1]
        lv = [(sel(U), sel(V), sel(D)) for U, V, D in self.levels]
        if transpose:
            lv = [(V, U, _T(D)) for U, V, D in lv]
            root = _T(sel(self.D_root))
        else:
            root = sel(self.D_root)
        qb = q.reshape(nb * m, bs, nc).contiguous()
        # upward pass
        ups = []
        x = qb
        for l, (U, V, D) in enumerate(lv):
            nodes, rows = U.shape[1], U.shape[2]
            xh = _mm(_T(V.reshape(nb * nodes, rows, k)), x)         # (nb*nodes, k, nc)
            ups.append(x)
            x = xh.reshape(nb * nodes // 2, 2 * k, nc)
        # root
        u = _mm(root.reshape(nb, 2 * k, 2 * k).contiguous() if not transpose else root, x.reshape(nb, 2 * k, nc))
        # downward pass: u holds, per node of level l, the incoming k-vector
        for l in range(self.L - 1, -1, -1):
            U, V, D = lv[l]
            nodes, rows = U.shape[1], U.shape[2]
            ul = u.reshape(nb * nodes, k, nc)
            y = _mm(U.reshape(nb * nodes, rows, k), ul)
            _mm(D.reshape(nb * nodes, rows, rows), ups[l], out=y, beta=1.0)
            u = y
        return u.reshape(nb, m * bs, nc)[:, :self.n]

    def todense(self, device=None, b_start=None, b_end=None, verbose=False):
        """The dense (b_end - b_start, n, n) matrices (FP64, on the device of
        the representation, or `device`)."""
        b0 = 0 if b_start is None else b_start
        b1 = self.b if b_end is None else b_end
        dev = self.D_root.device
        eye = torch.eye(self.n, dtype=torch.float64, device=dev).expand(b1 - b0, -1, -1).contiguous()
        out = self.matvec(eye, b0, b1)
        return out if device is None else out.to(device)


def copy_params_hbs(hbs, b):
    """An HBS_matrix of b zero matrices with hbs's tree (hbs_reconstruction.py:6-18)."""
    new = HBS_matrix(b, hbs.n, hbs.k)
    new.bs, new.m, new.n_pad, new.L = hbs.bs, hbs.m, hbs.n_pad, hbs.L
    new.levels = [tuple(torch.zeros((b,) + tuple(t.shape[1:]), dtype=t.dtype, device=t.device) for t in lv)
                  for lv in hbs.levels]
    new.D_root = torch.zeros((b,) + tuple(hbs.D_root.shape[1:]), dtype=hbs.D_root.dtype, device=hbs.D_root.device)
    return new


def copy_data_hbs(hbs, b_start, b_end):
    """Matrices b_start..b_end-1 of hbs (hbs_reconstruction.py:20-25)."""
    new = copy_params_hbs(hbs, b_end - b_start)
    new.levels = [tuple(t[b_start:b_end].clone() for t in lv) for lv in hbs.levels]
    new.D_root = hbs.D_root[b_start:b_end].clone()
    return new
