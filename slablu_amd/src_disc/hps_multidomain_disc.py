"""Multidomain HPS discretization, GPU-resident (drop-in for
src_disc/hps_multidomain_disc.py of the reference).

The face operator A = sum_boxes P_b^T DtN_b P_b is kept in its factored
form: the stack of leaf DtN maps (nboxes, 4(p-2), 4(p-2)) on the device plus
the map Iext (nboxes, 4(p-2)) from each leaf's [l, r, d, u] face slots to
global face-DOF ids (hps_multidomain_disc.py:63-81).  A x is applied
matrix-free (gather, batched DtN product, scatter-add of the <= 2
contributions per face DOF); the scipy CSR matrix of the reference is only
built when `sparse_mat` is asked for it.
"""
import time

import numpy as np
import scipy.sparse
import torch

from .. import hip
from . import hps_subdomain_disc
from .hps_parallel_leaf_ops import LeafOps, coef_arrays


def _default_device():
    if not torch.cuda.is_available():
        raise RuntimeError("slablu_amd runs on an AMD GPU (HIP); no GPU is visible")
    return torch.device("cuda")


class HPS_Multidomain:
    def __init__(self, pdo, domain, a, p, device=None, periodic=False):
        self.pdo = pdo
        self.periodic = periodic
        self.domain = torch.as_tensor(domain, dtype=torch.float64)
        self.p, self.a = p, a
        n = (self.domain[:, 1] - self.domain[:, 0]) / (2 * a)
        self.n = torch.round(n).int()
        self.n0, self.n1 = int(self.n[0]), int(self.n[1])
        self.nboxes = self.n0 * self.n1
        self.H = hps_subdomain_disc.HPS_Disc(a, p, d=2)
        self.grid_xx = self.get_grid()
        Jx = torch.as_tensor(self.H.JJ.Jx)
        self.I_unique, self.I_copy1, self.I_copy2 = self.get_unique_inds()
        grid_ext = self.grid_xx[:, Jx, :].reshape(-1, 2)
        self.xx_active = grid_ext[self.I_unique]
        self.xx_tot = self.grid_xx.reshape(-1, 2)
        self.Iext = self.iext_box()
        self.nface = (p - 2) * (self.n0 * (2 * self.n1 + 1) + self.n1)
        self.device = device
        self._dev = None
        self.leaf = None
        # optional (start, stop) torch.cuda.Event pair recorded around the
        # leaf condensation on the current stream (bench.py roofline)
        self.leaf_event_pair = None

    # -- geometry (host) -------------------------------------------------------
    def get_grid(self):
        """Node coordinates per leaf, box = i*n1 + j (hps_multidomain_disc.py:111-124)."""
        zz = torch.as_tensor(self.H.zz.T.copy())
        a = self.a
        i = np.arange(self.n0, dtype=np.float64)
        j = np.arange(self.n1, dtype=np.float64)
        offx = torch.as_tensor((a + (2 * a) * i)) + self.domain[0, 0]
        offy = torch.as_tensor((a + (2 * a) * j)) + self.domain[1, 0]
        xx = torch.empty(self.n0, self.n1, self.p ** 2, 2, dtype=torch.float64)
        xx[..., 0] = zz[:, 0][None, None, :] + offx[:, None, None]
        xx[..., 1] = zz[:, 1][None, None, :] + offy[None, :, None]
        return xx.reshape(self.nboxes, self.p ** 2, 2)

    def get_unique_inds(self):
        """Unique face slots and the shared-face copies (hps_multidomain_disc.py:126-182)."""
        f = self.p - 2
        n0, n1 = self.n0, self.n1
        box_ind = torch.arange(n0 * n1 * 4 * f).reshape(n0, n1, 4 * f)
        uni, c1, c2 = [box_ind[0, :, :f].flatten()], [], []
        for i in range(n0):
            if i > 0:
                l = box_ind[i, :, :f].flatten()
                r = box_ind[i - 1, :, f:2 * f].flatten()
                uni.append(l); c1.append(l); c2.append(r)
            uni.append(box_ind[i, 0, 2 * f:3 * f])
            d = box_ind[i, 1:, 2 * f:3 * f].flatten()
            u = box_ind[i, :n1 - 1, 3 * f:].flatten()
            uni.append(d); c1.append(d); c2.append(u)
            uni.append(box_ind[i, n1 - 1, 3 * f:])
        uni.append(box_ind[n0 - 1, :, f:2 * f].flatten())
        return torch.cat(uni), torch.cat(c1), torch.cat(c2)

    def iext_box(self):
        """Global face-DOF id of every leaf face slot (hps_multidomain_disc.py:63-81)."""
        f = self.p - 2
        n0, n1 = self.n0, self.n1
        i = torch.arange(n0).repeat_interleave(n1)
        j = torch.arange(n1).repeat(n0)
        off = (2 * n1 + 1) * i
        right = off + 2 * n1 + 1 + j
        if self.periodic:
            # periodic in x: the last column's right faces are the domain's
            # left boundary faces (slablu_solver.py:91-104), so the face
            # operator couples across the seam
            right = torch.where(i == n0 - 1, j, right)
        faces = torch.stack((off + j, right, off + n1 + j, off + n1 + j + 1), dim=1)
        return (faces[:, :, None] * f + torch.arange(f)[None, None, :]).reshape(self.nboxes, 4 * f)

    # -- device state -----------------------------------------------------------
    def _setup_device(self, device=None):
        dev = torch.device(device) if device is not None else (self.device or _default_device())
        if dev.type != "cuda":
            raise RuntimeError("the HPS leaf operators run on the GPU only (got device %s)" % dev)
        if self._dev != dev:
            self._dev = dev
            self.ops = LeafOps(self.H, dev)
            self.Iext_d = self.Iext.to(dev)
            self.xx_flat_d = self.grid_xx.reshape(-1, 2).to(dev)
        return dev

    def factor_leaves(self, device=None, dtype=torch.float64, chunk=None, store=True):
        """Evaluate the PDE coefficients at every node and condense every
        leaf.  chunk: leaves per batch (None: all in one batch; the
        reference sizes its chunks from free memory,
        hps_parallel_leaf_ops.py:102-141).  store=False keeps only the DtN
        maps: the leaf factors are recomputed, chunk by chunk, wherever they
        are needed again (leaf recovery, body loads, adjoint), as the
        reference's solve re-factors every leaf (hps_parallel_leaf_ops.py:
        73-92)."""
        dev = self._setup_device(device)
        self.leaf = None
        self.coefs = coef_arrays(self.pdo, self.xx_flat_d)
        nb, ni, nx = self.nboxes, self.ops.ni, self.ops.nx
        chunk = nb if chunk is None else max(1, min(int(chunk), nb))
        ev = self.leaf_event_pair
        if ev is not None:
            ev[0].record()
        if chunk >= nb and store:
            DtN, M, ipiv, info = self.ops.factor(self.coefs, nb, dtype)
        else:
            DtN = torch.empty(nb, nx, nx, dtype=dtype, device=dev)
            info = torch.zeros(nb, dtype=torch.int32, device=dev)
            M = torch.empty(nb, ni, ni + nx, dtype=dtype, device=dev) if store else None
            ipiv = torch.empty(nb, ni, dtype=torch.int32, device=dev) if store else None
            Mc = None if store else torch.empty(chunk, ni, ni + nx, dtype=dtype, device=dev)
            ipc = None if store else torch.empty(chunk, ni, dtype=torch.int32, device=dev)
            for s in range(0, nb, chunk):
                e = min(s + chunk, nb)
                out = (DtN[s:e], M[s:e] if store else Mc[:e - s], ipiv[s:e] if store else ipc[:e - s], info[s:e])
                self.ops.factor(self._coef_slice(s, e), e - s, dtype, out=out)
            del Mc, ipc
        if ev is not None:
            ev[1].record()
        self.leaf = dict(DtN=DtN, M=M, ipiv=ipiv, perm=hip.piv_to_perm(ipiv) if store else None, info=info,
                         dtype=dtype, chunk=chunk, stored=store)
        return DtN

    def _coef_slice(self, s, e):
        """Coefficient arrays of leaves s..e-1."""
        pp = self.p * self.p
        return {k: (None if v is None else v[s * pp:e * pp]) for k, v in self.coefs.items()}

    def leaf_chunks(self):
        """(s, e, M, ipiv, perm) over the leaves in the chunks of the
        factorization: slices of the kept factors, or the factors of leaves
        s..e-1 recomputed when only the DtN maps are kept (bit-identical:
        every leaf is factored independently)."""
        L = self.leaf
        nb, ni, nx = self.nboxes, self.ops.ni, self.ops.nx
        if L["stored"]:
            yield 0, nb, L["M"], L["ipiv"], L["perm"]
            return
        chunk = L["chunk"]
        dev = self._dev
        Mc = torch.empty(chunk, ni, ni + nx, dtype=L["dtype"], device=dev)
        Dc = torch.empty(chunk, nx, nx, dtype=L["dtype"], device=dev)
        ipc = torch.empty(chunk, ni, dtype=torch.int32, device=dev)
        infc = torch.empty(chunk, dtype=torch.int32, device=dev)
        for s in range(0, nb, chunk):
            e = min(s + chunk, nb)
            k = e - s
            self.ops.factor(self._coef_slice(s, e), k, L["dtype"], out=(Dc[:k], Mc[:k], ipc[:k], infc[:k]))
            yield s, e, Mc[:k], ipc[:k], hip.piv_to_perm(ipc[:k])

    def get_DtNs(self, device=None, mode="build", data=0, ff_body_func=None):
        """mode 'build': (nboxes, 4(p-2), 4(p-2)) leaf DtN maps (device tensor).
        mode 'solve': (nboxes, p^2, 2*nrhs) [leaf solution | interior residual]
        from per-leaf face data `data` (nboxes, 4(p-2), nrhs), as
        src_disc/hps_multidomain_disc.py:187-224 returns."""
        if mode == "build":
            return self.factor_leaves(device)
        if mode == "solve":
            if self.leaf is None:
                self.factor_leaves(device)
            uu, res = self._leaf_solve(data.to(self._dev), ff_body_func)
            return torch.cat((uu, res), dim=-1)
        raise ValueError("mode %r" % mode)

    def body_load(self, ff_body_func, nrhs=1, dtype=torch.float64):
        """f at the interior nodes of every leaf (nboxes, ni, nrhs)."""
        fb = ff_body_func(self.xx_flat_d)
        fb = torch.as_tensor(fb, device=self._dev).to(dtype).reshape(self.nboxes, self.p ** 2, -1)
        return fb[:, self.ops.Jc].contiguous()

    def _leaf_solve(self, u_bnd, ff_body_func=None):
        L = self.leaf
        u_bnd = u_bnd.to(L["dtype"]).contiguous()
        fb = None if ff_body_func is None else self.body_load(ff_body_func, u_bnd.shape[-1], L["dtype"])
        if L["stored"]:
            uc = self.ops.interior(L["M"], L["ipiv"], u_bnd, fb, perm=L.get("perm"))
        else:
            uc = torch.empty(self.nboxes, self.ops.ni, u_bnd.shape[-1], dtype=L["dtype"], device=u_bnd.device)
            for s, e, M, ipiv, perm in self.leaf_chunks():
                uc[s:e] = self.ops.interior(M, ipiv, u_bnd[s:e], None if fb is None else fb[s:e], perm=perm)
        uu = self.ops.assemble_nodes(u_bnd, uc)
        Au = self.ops.apply_aloc(uu, self.coefs)
        res = torch.zeros_like(uu)
        r = Au[:, self.ops.Jc]
        if fb is not None:
            r = r - fb
        res[:, self.ops.Jc] = r
        return uu, res

    def leaf_values(self, u_face):
        """Values at every leaf node (nboxes, p^2, nrhs) from the face values
        (leaf recovery without the residual; no host synchronization)."""
        L = self.leaf
        u_bnd = self.faces_to_boxes(u_face).to(L["dtype"]).contiguous()
        if L["stored"]:
            uc = self.ops.interior(L["M"], L["ipiv"], u_bnd, None, perm=L.get("perm"))
        else:
            uc = torch.empty(self.nboxes, self.ops.ni, u_bnd.shape[-1], dtype=L["dtype"], device=u_bnd.device)
            for s, e, M, ipiv, perm in self.leaf_chunks():
                uc[s:e] = self.ops.interior(M, ipiv, u_bnd[s:e], None, perm=perm)
        return self.ops.assemble_nodes(u_bnd, uc)

    def faces_to_boxes(self, x_face):
        """(nface, nrhs) face vector -> (nboxes, 4(p-2), nrhs) per-leaf data."""
        return x_face[self.Iext_d]

    def solve(self, device, uu_sol, ff_body_func=None):
        """Leaf recovery from the face solution on all face DOFs
        (hps_multidomain_disc.py:227-243).  Returns (solution at all leaf
        nodes (nboxes*p^2, nrhs), 2-norm of the leaf interior residuals)."""
        self._setup_device(device)
        u_face = uu_sol.to(self._dev)
        uu, res = self._leaf_solve(self.faces_to_boxes(u_face), ff_body_func)
        return uu.reshape(-1, uu.shape[-1]), torch.linalg.norm(res).item()

    # -- face operator ----------------------------------------------------------
    def apply_A(self, x_face, DtN=None, transpose=False):
        """y = A x (or A^T x) for a face vector (nface, nrhs), matrix-free
        from the DtNs: A = sum_b P_b^T DtN_b P_b, A^T = sum_b P_b^T DtN_b^T P_b."""
        DtN = self.leaf["DtN"] if DtN is None else DtN
        xb = x_face.to(DtN.dtype)[self.Iext_d].contiguous()
        zb = torch.empty_like(xb)
        hip.gemm(DtN.transpose(1, 2) if transpose else DtN, xb, zb)
        y = torch.zeros(self.nface, x_face.shape[-1], dtype=DtN.dtype, device=DtN.device)
        y.index_add_(0, self.Iext_d.reshape(-1), zb.reshape(-1, x_face.shape[-1]))
        return y

    def sparse_mat(self, device=None, verbose=False):
        """The assembled face operator as a scipy CSR matrix (host), with the
        reference's timing dictionary (hps_multidomain_disc.py:53-109)."""
        tic = time.time()
        if self.leaf is None:
            self.factor_leaves(device)
        torch.cuda.synchronize()
        toc_dtn = time.time() - tic
        tic = time.time()
        DtN = self.leaf["DtN"].double().cpu().numpy()
        I = self.Iext.numpy()
        nb, ne = I.shape
        rows = np.repeat(I, ne, axis=1).reshape(nb, ne, ne)
        cols = np.transpose(rows, (0, 2, 1))
        A = scipy.sparse.coo_matrix((DtN.ravel(), (rows.ravel(), cols.ravel())),
                                    shape=(self.nface, self.nface)).tocsr()
        toc_csr = time.time() - tic
        return A, dict(toc_DtN=toc_dtn, toc_forloop=0.0, toc_sparse=toc_csr)
