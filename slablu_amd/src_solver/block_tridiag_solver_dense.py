"""Batched block-tridiagonal LU on the GPU (drop-in for
src_solver/block_tridiag_solver_dense.py of the reference).

Storage layout is the reference's Tdata: (b, m + 2(m-1) + m//bs + 1, bs, bs)
= [diag (m) | sub (m-1) | sup (m-1) | pivot blocks] (block_tridiag_solver_dense.py
:226-257, 291-330), kept on the device.

form='lu' (default): the reference's block LU, diag_i <- packed LU of
D_i - Lfac_i Ufac_i, sub_i <- Lfac = sub_i U_i^{-1}, sup_i <- Ufac =
L_i^{-1} P_i^T sup_i (block_tridiag_solver_dense.py:152-177), pivots packed
into the pivot blocks as values (1-based).  Chains whose blocks fit in LDS
are factored and swept by one workgroup each (slab_btri_factor /
slab_btri_solve); backward stable, which the FP32 (mixed-precision) path
needs.
form='inv': block-Thomas factorization with explicit diagonal inverses,
diag_i <- S_i^{-1} with S_i = D_i - sub_{i-1} G_{i-1}, sup_i <- G_i =
S_i^{-1} sup_i, sub_i unchanged (slab_btri_inv_factor / slab_btri_inv_solve):
every sweep step is a matrix product; about one digit less accurate.
"""
import numpy as np
import torch

from .. import hip


def alloc_pivdata(b, m, bs, device=None, dtype=torch.float64):
    nblocks = 1 if m < bs else m // bs + 1
    return torch.zeros(b, nblocks, bs, bs, device=device, dtype=dtype)


def pack_pivs(b, m, bs, Tdata, pivs):
    """Pivot indices into the trailing blocks of Tdata (block_tridiag_solver_dense.py:245-257)."""
    off = m + 2 * (m - 1)
    flat = pivs.reshape(b, m * bs).to(Tdata.dtype)
    view = Tdata[:, off:].reshape(b, -1)
    view[:, :m * bs] = flat
    Tdata[:, off:] = view.reshape(Tdata[:, off:].shape)
    return Tdata


def unpack_pivs(b, m, bs, Tdata, device=None):
    off = m + 2 * (m - 1)
    return Tdata[:, off:].reshape(b, -1)[:, :m * bs].reshape(b, m, bs).to(torch.int32)


class Block_Tridiagonal:
    def __init__(self, b, m, bs, form="lu", twisted=False, inv_factors=False):
        """form 'lu': the reference's block LU (packed LU + Lfac/Ufac + pivots
        in Tdata); 'inv': block-Thomas with explicit diagonal inverses.
        twisted (form 'lu'): eliminate from both ends of each chain towards
        block m//2, which halves the dependent steps of the factorization and
        of every sweep; Tdata then holds the twisted factors (same layout).
        inv_factors (twisted): the same elimination, but each diagonal block
        is also kept as its triangular inverses L^{-1} P and U^{-1}, so the
        non-transposed sweeps are block products (matrix cores) instead of
        triangular substitutions."""
        if form not in ("inv", "lu"):
            raise ValueError("form must be 'inv' or 'lu'")
        self.b, self.m, self.bs = b, m, bs
        self.form = form
        self.twisted = bool(twisted) and form == "lu"
        self.inv_factors = bool(inv_factors) and self.twisted
        self.linv = self.uinv = None
        self.pivdata_nblocks = int(m / bs) + 1
        self.storage_blocks = m + 2 * (m - 1) + self.pivdata_nblocks
        self.nbytes = (self.b * self.storage_blocks * self.bs ** 2) * 8
        self.fast_LU = False

    def build_sweep(self, Tdiag, Tsub, Tsup, fast_LU=False):
        b, m, bs = self.b, self.m, self.bs
        assert Tdiag.shape[1] == m and Tsub.shape[1] == m - 1 and Tsup.shape[1] == m - 1
        piv = alloc_pivdata(b, m, bs, device=Tdiag.device, dtype=Tdiag.dtype)
        self.Tdata = torch.cat((Tdiag, Tsub, Tsup, piv), 1)
        self._build()

    def build_sweep_from_data(self, Tdata, fast_LU=False):
        assert Tdata.shape[1] == self.storage_blocks
        self.Tdata = Tdata
        self._build()

    def _views(self, b_start=0, b_end=None):
        m = self.m
        T = self.Tdata[b_start:b_end]
        return T[:, :m], T[:, m:2 * m - 1], T[:, 2 * m - 1:3 * m - 2]

    def _build(self):
        if not self.Tdata.is_cuda:
            raise RuntimeError("Block_Tridiagonal factors on the GPU; pass device tensors")
        if not self.Tdata.is_contiguous():
            self.Tdata = self.Tdata.contiguous()
        self.itemsize = self.Tdata.element_size()
        self.nbytes = self.Tdata.numel() * self.itemsize
        b, m, bs = self.b, self.m, self.bs
        self.info = torch.zeros(b, dtype=torch.int32, device=self.Tdata.device)
        d, l, u = self._views()
        if self.form == "inv":
            hip.btri_inv_factor(d, l, u, info=self.info)
            return
        self.pivs = torch.zeros(b, m, bs, dtype=torch.int32, device=self.Tdata.device)
        if self.twisted and not hip.btri_fused(self.Tdata.dtype, bs):
            self.twisted = False
        hip.btri_factor(d, l, u, self.pivs, info=self.info, twisted=self.twisted)
        pack_pivs(b, m, bs, self.Tdata, self.pivs)
        if self.inv_factors and self.twisted:
            self.uinv = d.contiguous().clone()
            self.linv = torch.empty_like(self.uinv)
            hip.btri_invert_diag(self.uinv, self.pivs, self.linv)
            self.nbytes += 2 * self.uinv.numel() * self.itemsize

    def lusolve_sweep(self, f, b_start=None, b_end=None, transpose=False):
        """Solve (in place on f, (b_end-b_start, m*bs, nrhs)); returns f."""
        b_start = 0 if b_start is None else b_start
        b_end = self.b if b_end is None else b_end
        if f.device != self.Tdata.device or f.dtype != self.Tdata.dtype:
            raise ValueError("rhs must be a %s tensor on %s" % (self.Tdata.dtype, self.Tdata.device))
        d, l, u = self._views(b_start, b_end)
        fv = f.reshape(b_end - b_start, self.m * self.bs, -1)
        if fv.stride(-1) != 1:
            raise ValueError("rhs must have unit column stride")
        if self.form == "inv":
            if transpose:
                raise NotImplementedError("transposed sweeps need form='lu'")
            hip.btri_inv_solve(d, l, u, fv)
        else:
            if self.linv is not None:
                hip.btri_solve(self.uinv[b_start:b_end], l, u, None, fv, transpose=transpose, twisted=True,
                               linv=self.linv[b_start:b_end])
            else:
                hip.btri_solve(d, l, u, self.pivs[b_start:b_end], fv, transpose=transpose, twisted=self.twisted)
        return fv


def recover_block_tridiag_via_sampling(b, m, bs, func, func_args=None, device=torch.device("cuda")):
    """Diag/sub/sup blocks of b block-tridiagonal operators known only
    through their action, from 3*bs coloured probes
    (block_tridiag_solver_dense.py:9-77).  Returns Tdata-layout storage."""
    ncol = min(3, m) * bs
    rhs = torch.zeros(m * bs, ncol, device=device, dtype=torch.float64)
    I = torch.eye(ncol, device=device, dtype=torch.float64)
    for i in range(m):
        c = (i % 3) * bs
        rhs[i * bs:(i + 1) * bs] = I[c:c + bs]
    res = func((func_args, rhs.repeat(b, 1, 1)))
    nb = m + 2 * (m - 1) + int(m / bs) + 1
    A = torch.zeros(b, nb, bs, bs, device=device, dtype=res.dtype)
    for i in range(m):
        c = (i % 3) * bs
        A[:, i] = res[:, i * bs:(i + 1) * bs, c:c + bs]
        if i < m - 1:
            cn = ((i + 1) % 3) * bs
            A[:, m + i] = res[:, (i + 1) * bs:(i + 2) * bs, c:c + bs]
            A[:, 2 * m - 1 + i] = res[:, i * bs:(i + 1) * bs, cn:cn + bs]
    return A
