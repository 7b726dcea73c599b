"""Batched leaf operations on the GPU (replaces src_disc/hps_parallel_leaf_ops.py).

Static condensation of every leaf in one pass (all leaves batched in each
launch):
    M   = [A_cc | A_cx]                      slab_leaf_assemble
    M   = [L\\U | L^{-1} P A_cx]              slab_getrf (nelim = n_i)
    S'  = U^{-1} (L^{-1} P A_cx) = A_cc^{-1} A_cx   slab_trsm (in place)
    DtN = N_x - N_c S'                       slab_gemm
which is DtN = Nx[:, [Jc, Jx]] @ [-A_cc^{-1} A_cx; I] of
src_disc/hps_parallel_leaf_ops.py:50-58.  The factored M is kept: its right
block gives leaf interiors from face data (u_c = -S' u_x) and its LU the
interior solve with a body load, so the solve never re-factors A_cc (the
reference re-forms and re-factors every leaf in its solve,
hps_parallel_leaf_ops.py:73-92).
"""
import numpy as np
import torch

from .. import hip


def get_nearest_div(n, div):
    while n % div > 0:
        div -= 1
    return div


class LeafOps:
    """Device constants of one leaf discretization and the batched leaf
    kernels.  H is an HPS_Disc."""

    def __init__(self, H, device):
        self.p = p = H.p
        self.f = f = p - 2
        self.ni, self.nx = f * f, 4 * f
        self.device = device
        JJ = H.JJ
        Jtot = np.hstack((JJ.Jc, JJ.Jx))
        t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)
        self.D = t(H.D)
        self.Dsq = t(H.Dsq)
        self.Nxtot = t(H.Nx[:, Jtot])
        self.Jc = t(JJ.Jc, torch.long)
        self.Jx = t(JJ.Jx, torch.long)
        self.Jxreo = t(JJ.Jxreorder, torch.long)
        self.Intmap = t(H.Interp_mat)
        # N_x in compressed rows of p entries (each row is a derivative along
        # one grid line), for the fused residual kernel
        Nx = np.asarray(H.Nx, dtype=np.float64)
        self.Nx_nat = t(Nx)
        nx_idx = np.zeros((Nx.shape[0], p), dtype=np.int32)
        nx_val = np.zeros((Nx.shape[0], p), dtype=np.float64)
        for r in range(Nx.shape[0]):
            nz = np.nonzero(Nx[r])[0]
            if len(nz) > p:
                raise ValueError("N_x row %d has %d entries (at most p=%d expected)" % (r, len(nz), p))
            nx_idx[r, :len(nz)] = nz
            nx_val[r, :len(nz)] = Nx[r, nz]
        self.res_tables = dict(Intmap=self.Intmap.contiguous(), Jc=t(JJ.Jc, torch.int32),
                               Jxreo=t(JJ.Jxreorder, torch.int32), D=self.D.contiguous(),
                               Dsq=self.Dsq.contiguous(), nx_idx=t(nx_idx, torch.int32), nx_val=t(nx_val))
        self._typed = {}

    def typed(self, dtype):
        """Nx_c, Nx_x, Intmap, D, Dsq in the working precision."""
        if dtype not in self._typed:
            ni = self.ni
            self._typed[dtype] = dict(
                Nx_c=self.Nxtot[:, :ni].to(dtype).contiguous(),
                Nx_x=self.Nxtot[:, ni:].to(dtype).contiguous(),
                Intmap=self.Intmap.to(dtype).contiguous(),
                NxI=(self.Nx_nat[:, self.Jxreo] @ self.Intmap).to(dtype).contiguous(),
                D=self.D.to(dtype).contiguous(),
                Dsq=self.Dsq.to(dtype).contiguous(),
            )
        return self._typed[dtype]

    # -- build ---------------------------------------------------------------
    def factor(self, coefs, nboxes, dtype=torch.float64, out=None):
        """Condense every leaf.  coefs: dict name -> FP64 (nboxes*p*p) device
        tensor.  Returns (DtN (nboxes, nx, nx), M (nboxes, ni, ni+nx) factored,
        ipiv (nboxes, ni) int32, info (nboxes,) int32).  out: optional
        (DtN, M, ipiv, info) tensors (contiguous, those shapes) written in
        place, e.g. slices of the whole stack for one chunk of leaves."""
        ni, nx = self.ni, self.nx
        if out is not None:
            DtN_o, M_o, ipiv, info = out
            info.zero_()
        else:
            DtN_o = M_o = None
            ipiv = torch.empty(nboxes, ni, dtype=torch.int32, device=self.device)
            info = torch.zeros(nboxes, dtype=torch.int32, device=self.device)
        ty = self.typed(dtype)
        if coefs.get("c12") is None:
            M = M_o if M_o is not None else torch.empty(nboxes, ni, ni + nx, dtype=dtype, device=self.device)
            hip.leaf_assemble(M, self.p, 0, self.D, self.Dsq, coefs, self.Nxtot)
            hip.getrf(M, ipiv, nelim=ni, npiv=ni, info=info)
            hip.trsm(M[:, :, :ni], M[:, :, ni:], lower=False, unit=False)
            if DtN_o is not None:
                DtN = DtN_o
                DtN.copy_(ty["Nx_x"].expand(nboxes, nx, nx))
            else:
                DtN = ty["Nx_x"].expand(nboxes, nx, nx).contiguous()
        else:
            # the cross term couples interior and corner nodes: condense
            # against the corner-inclusive exterior nodes (Jxreorder) and
            # map their values from the face values through Intmap
            # (src_disc/hps_parallel_leaf_ops.py:59-65):
            #   S' = A_cc^{-1} A[Jc, Jxr] Intmap,  DtN = Nx[:, Jxr] Intmap - Nx_c S'
            nxr = self.Jxreo.numel()
            Mr = torch.empty(nboxes, ni, ni + nxr, dtype=dtype, device=self.device)
            hip.leaf_assemble_nodes(Mr, self.p, self.res_tables["Jxreo"], self.D, self.Dsq, coefs)
            hip.getrf(Mr, ipiv, nelim=ni, npiv=ni, info=info)
            hip.trsm(Mr[:, :, :ni], Mr[:, :, ni:], lower=False, unit=False)
            M = M_o if M_o is not None else torch.empty(nboxes, ni, ni + nx, dtype=dtype, device=self.device)
            M[:, :, :ni] = Mr[:, :, :ni]
            hip.gemm(Mr[:, :, ni:], ty["Intmap"], M[:, :, ni:])
            del Mr
            if DtN_o is not None:
                DtN = DtN_o
                DtN.copy_(ty["NxI"].expand(nboxes, nx, nx))
            else:
                DtN = ty["NxI"].expand(nboxes, nx, nx).contiguous()
        hip.gemm(ty["Nx_c"], M[:, :, ni:], DtN, alpha=-1.0, beta=1.0)
        return DtN, M, ipiv, info

    # -- solve ---------------------------------------------------------------
    def interior(self, M, ipiv, u_bnd, f_body=None, perm=None):
        """Leaf interiors u_c = A_cc^{-1}(f_c - A_cx u_x) (nboxes, ni, nrhs);
        perm: the composed pivot permutation (hip.piv_to_perm) if kept."""
        ni = self.ni
        nb, _, nrhs = u_bnd.shape
        if f_body is None:
            uc = torch.empty(nb, ni, nrhs, dtype=M.dtype, device=M.device)
            hip.gemm(M[:, :, ni:], u_bnd, uc, alpha=-1.0, beta=0.0)
            return uc
        uc = f_body.to(M.dtype).contiguous().clone()
        hip.getrs(M[:, :, :ni], ipiv, uc, perm=perm)
        hip.gemm(M[:, :, ni:], u_bnd, uc, alpha=-1.0, beta=1.0)
        return uc

    def assemble_nodes(self, u_bnd, uc):
        """All p^2 node values per leaf: corners and faces from the face data
        through Interp_mat (as src_disc/hps_parallel_leaf_ops.py:84 does),
        interior from uc."""
        nb, _, nrhs = u_bnd.shape
        ty = self.typed(u_bnd.dtype)
        uu = torch.zeros(nb, self.p * self.p, nrhs, dtype=u_bnd.dtype, device=u_bnd.device)
        ureo = torch.empty(nb, ty["Intmap"].shape[0], nrhs, dtype=u_bnd.dtype, device=u_bnd.device)
        hip.gemm(ty["Intmap"], u_bnd, ureo)
        uu[:, self.Jxreo] = ureo
        uu[:, self.Jc] = uc
        return uu

    def apply_aloc(self, uu, coefs):
        """(A_loc u) at every node of every leaf, matrix-free through the 1-D
        factors: uu (nb, p^2, nrhs) -> (nb, p^2, nrhs), FP64 or FP32 as uu."""
        p = self.p
        nb, _, nrhs = uu.shape
        ty = self.typed(uu.dtype)
        U = uu.contiguous().view(nb, p, p * nrhs)
        # x-direction: sum_j1 K[i1, j1] U[j1, i2, :]
        def xdir(K):
            out = torch.empty_like(U)
            hip.gemm(K, U, out)
            return out.view(nb, p * p, nrhs)

        Uy = uu.contiguous().view(nb * p, p, nrhs)

        def ydir(K):
            out = torch.empty_like(Uy)
            hip.gemm(K, Uy, out)
            return out.view(nb, p * p, nrhs)

        def cf(name):
            c = coefs.get(name)
            return None if c is None else c.view(nb, p * p, 1).to(uu.dtype)

        res = -cf("c11") * xdir(ty["Dsq"])
        res = res - cf("c22") * ydir(ty["Dsq"])
        if coefs.get("c12") is not None:
            t = torch.empty_like(Uy)
            hip.gemm(ty["D"], xdir(ty["D"]).view(nb * p, p, nrhs), t)
            res = res - 2.0 * cf("c12") * t.view(nb, p * p, nrhs)
        if coefs.get("c1") is not None:
            res = res - cf("c1") * xdir(ty["D"])
        if coefs.get("c2") is not None:
            res = res - cf("c2") * ydir(ty["D"])
        if coefs.get("c") is not None:
            res = res + cf("c") * uu
        return res

    def residual(self, uc, ub, coefs, uu=None):
        """FP64 residual pieces at an iterate given by interior values uc
        (nb, ni, nrhs) and per-leaf face values ub (nb, nx, nrhs), in one
        launch: (rI = -(A_loc u) at the interior nodes, N_x u at the face
        nodes); uu, if given, receives all node values."""
        return hip.leaf_residual(uc.contiguous(), ub.contiguous(), self.res_tables, coefs, self.p, uu=uu)

    def flux(self, uu):
        """Outward normal derivatives Nx u on the 4(p-2) face nodes of every
        leaf: (nb, p^2, nrhs) -> (nb, nx, nrhs)."""
        nb, _, nrhs = uu.shape
        ty = self.typed(uu.dtype)
        Nx = self.Nxtot.to(uu.dtype)
        Jtot = torch.cat((self.Jc, self.Jx))
        out = torch.empty(nb, self.nx, nrhs, dtype=uu.dtype, device=uu.device)
        hip.gemm(Nx, uu[:, Jtot].contiguous(), out)
        return out


def coef_arrays(pdo, xx_flat):
    """Evaluate the PDO coefficient callables on the node coordinates (FP64
    device tensors of shape (M,))."""
    out = {}
    for name in ("c11", "c22", "c12", "c1", "c2", "c"):
        fn = getattr(pdo, name)
        if fn is None:
            out[name] = None
            continue
        v = fn(xx_flat)
        v = torch.as_tensor(v, device=xx_flat.device).to(torch.float64).reshape(-1)
        if v.numel() == 1:
            v = v.expand(xx_flat.shape[0])
        out[name] = v.contiguous()
    if out["c11"] is None or out["c22"] is None:
        raise ValueError("PDO_2d needs c11 and c22")
    return out
