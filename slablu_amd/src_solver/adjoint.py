"""Adjoint of the FP64 SlabLU solve with respect to a per-node PDE
coefficient (the north star's adjoint form; the reference has the
transposed sweeps it builds on, block_tridiag_solver_dense.py:205-222 and
slablu_solver.py:900-943, but no gradient).

Problem: -Lap u - kh^2 b(x) u = 0 on the HPS grid, Dirichlet data g, with b
given at every leaf node (nboxes * p^2 values, the node order of
hps.xx_tot).  `HelmholtzAdjoint.solve(b)` returns u at every leaf node and is
differentiable: its backward is a hand-written torch.autograd.Function.

Gradient.  The un-condensed discrete system is
    F(u, b) = [A_cc u_c + A_cx u_x ;  sum_b P_b^T (N_c u_c + N_x u_x)] = 0
over the leaf interiors u_c and the face unknowns (I_Ctot); b enters only
the diagonal of A_cc (c = -kh^2 b at the interior rows), so
dF/db_i = -kh^2 u_i e_i and, for a loss L(u),
    dL/db_i = kh^2 lam_i u_i   at leaf-interior nodes, 0 elsewhere,
with F_u^T lam = dL/du.  Condensing F_u^T the same way the forward solve
condenses F_u:
    A_CC^T lam_C = g_C - (sum_b P_b^T S_b^T g_c)[I_Ctot],   S = A_cc^{-1} A_cx,
    lam_c = A_cc^{-T} (g_c - N_c^T lam_x)                    per leaf,
where g = dL/du reaches the face unknowns through the corner-inclusive
interpolation of the leaf's exterior nodes.  The backward reuses the stored
factors: S' (kept beside the leaf LU) for the face right-hand side, the
transposed two-level solve (slab sweeps with every block transposed, the
interface elimination transposed) for lam_C, and the leaf LU for the
transposed interior solves.  Under a leaf memory budget the leaf factors are
recomputed chunk by chunk instead of stored (twice in the backward: before
and after the face solve).
"""
import time

import torch

from .. import hip
from ..src_disc import pdo as pdo_mod
from .slablu_solver import Domain_Solver


class _NodeValues:
    """A PDO coefficient given by its values at the leaf nodes: the solver
    evaluates coefficients on hps.xx_tot (node order), so the callable
    returns the stored values."""

    def __init__(self):
        self.values = None

    def __call__(self, xx):
        if self.values is None or self.values.numel() != xx.shape[0]:
            raise ValueError("per-node coefficient of %s values for %d nodes"
                             % (None if self.values is None else self.values.numel(), xx.shape[0]))
        return self.values


class _CoefficientSolve(torch.autograd.Function):
    """u = solve(b); backward: dL/db from dL/du by the adjoint solve."""

    @staticmethod
    def forward(ctx, b, problem):
        u = problem._forward(b)
        ctx.problem = problem
        ctx.save_for_backward(b, u)
        return u

    @staticmethod
    def backward(ctx, grad_u):
        b, u = ctx.saved_tensors
        return ctx.problem._backward(b, u, grad_u), None


class HelmholtzAdjoint:
    """-Lap u - kh^2 b u = 0 with Dirichlet data uu_dir_func on box_geom
    (leaf half-width a, order p), b per leaf node; FP64 factor + solve by
    Domain_Solver and its adjoint.  leaf_memory_budget (bytes): leaf factors
    recomputed in chunks when they do not fit (Domain_Solver)."""

    def __init__(self, box_geom, kh, a, p, uu_dir_func, buf_constant=1.0, leaf_memory_budget=None, leaf_chunk=None,
                 device=None, periodic_bc=False):
        self._coef = _NodeValues()
        op = pdo_mod.PDO_2d(pdo_mod.ones, pdo_mod.ones, c=self._coef)
        self.dom = Domain_Solver(box_geom, op, kh, a, p=p, buf_constant=buf_constant, precision="fp64",
                                 device=device, periodic_bc=periodic_bc, leaf_memory_budget=leaf_memory_budget,
                                 leaf_chunk=leaf_chunk)
        self.kh = float(kh)
        self.p = p
        self.uu_dir_func = uu_dir_func
        hps = self.dom.hps
        self.nboxes = hps.nboxes
        self.nnodes = hps.nboxes * p * p
        self.xx = hps.xx_tot
        self._built_b = None
        self.timings = {}

    # -- helpers -----------------------------------------------------------------
    def interior_nodes(self):
        """Flat indices (into the nboxes*p^2 node vector) of every leaf's
        interior nodes, the nodes where b enters the equations."""
        Jc = torch.as_tensor(self.dom.hps.H.JJ.Jc, dtype=torch.long)
        return (torch.arange(self.nboxes)[:, None] * self.p * self.p + Jc[None, :]).reshape(-1)

    def observation_nodes(self, count=1024, seed=0):
        """`count` distinct seeded leaf-interior nodes (observation points of
        the loss)."""
        inter = self.interior_nodes()
        g = torch.Generator().manual_seed(seed)
        return inter[torch.randperm(inter.numel(), generator=g)[:count]].sort().values

    # -- forward / adjoint ---------------------------------------------------------
    def solve(self, b):
        """u (nboxes*p^2, nrhs) FP64 on the device at every leaf node for
        the coefficient b (nboxes*p^2,); differentiable in b."""
        return _CoefficientSolve.apply(b, self)

    def _forward(self, b):
        dom = self.dom
        bd = b.detach().to(dom.device, torch.float64).reshape(-1).contiguous()
        if bd.numel() != self.nnodes:
            raise ValueError("b must hold one value per leaf node (%d), got %d" % (self.nnodes, bd.numel()))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self._coef.values = (-(self.kh ** 2)) * bd
        dom.build(verbose=False)
        dom.solve(self.uu_dir_func)
        u = dom.sol_dev.reshape(self.nnodes, -1).to(torch.float64)
        torch.cuda.synchronize()
        self.timings["forward_s"] = time.perf_counter() - t0
        self._built_b = bd.clone()
        return u

    def _backward(self, b, u, grad_u):
        dom = self.dom
        hps, ops = dom.hps, dom.hps.ops
        bd = b.detach().to(dom.device, torch.float64).reshape(-1)
        if self._built_b is None or not torch.equal(self._built_b, bd):
            # the factors now held are those of another coefficient
            u = self._forward(bd)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        p, nb, ni, nx = self.p, self.nboxes, ops.ni, ops.nx
        nc = u.shape[-1]
        dev = dom.device
        g = grad_u.to(dev, torch.float64).reshape(nb, p * p, nc)
        g_c = g[:, ops.Jc].contiguous()
        # dL/du at the exterior nodes -> dL/d(face values) through Interp_mat
        g_ext = g[:, ops.Jxreo].contiguous()
        g_bnd = torch.empty(nb, nx, nc, dtype=torch.float64, device=dev)
        hip.gemm(ops.Intmap.transpose(0, 1), g_ext, g_bnd)
        # minus S'^T g_c, leaf by leaf (the factors kept or recomputed)
        for s, e, M, ipiv, perm in hps.leaf_chunks():
            hip.gemm(M[:, :, ni:].transpose(1, 2), g_c[s:e], g_bnd[s:e], alpha=-1.0, beta=1.0)
        Ic = dom._Ic()
        g_face = torch.zeros(hps.nface, nc, dtype=torch.float64, device=dev)
        g_face.index_add_(0, hps.Iext_d.reshape(-1), g_bnd.reshape(-1, nc))
        lam_C = dom.solve_reduced(g_face[Ic], transpose=True)
        lam_face = torch.zeros(hps.nface, nc, dtype=torch.float64, device=dev)
        lam_face[Ic] = lam_C
        lam_x = hps.faces_to_boxes(lam_face).contiguous()
        # lam_c = A_cc^{-T} (g_c - N_c^T lam_x)
        lam_c = g_c.clone()
        hip.gemm(ops.typed(torch.float64)["Nx_c"].transpose(0, 1), lam_x, lam_c, alpha=-1.0, beta=1.0)
        for s, e, M, ipiv, perm in hps.leaf_chunks():
            hip.getrs(M[:, :, :ni], ipiv, lam_c[s:e], trans=True)
        u_c = u.reshape(nb, p * p, nc)[:, ops.Jc]
        grad = torch.zeros(nb, p * p, dtype=torch.float64, device=dev)
        grad[:, ops.Jc] = (self.kh ** 2) * (lam_c * u_c).sum(-1)
        torch.cuda.synchronize()
        self.timings["backward_s"] = time.perf_counter() - t0
        return grad.reshape(b.shape).to(b.dtype)


def observed_loss(u, obs, d):
    """L = 1/2 ||u[obs] - d||^2 for the first column of u."""
    r = u[obs, 0] - d
    return 0.5 * torch.dot(r, r)
