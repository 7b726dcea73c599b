"""SlabLU two-level direct solver for HPS discretizations, on the GPU
(drop-in for src_solver/slablu_solver.py of the reference).

Domain_Solver(box_geom, pdo_op, kh, h, p, buf_constant, periodic_bc)
  .build(sparse_assembly, solver_type, verbose) -> info dict
  .solve(uu_dir_func, ff_body_func=None, known_sol=False)
        -> (sol_tot, rel_err, true_err, resloc_hps, toc_solve)

Factor (everything batched on the device; FP64 by default):
  1. leaf DtN maps of all leaves                   (hps_parallel_leaf_ops)
  2. Aloc: per leaf row of every slab, the block-tridiagonal system of the
     buf-1 interior vertical faces, gathered directly from the DtN blocks
     (the reference recovers the same blocks by probing A,
     slablu_solver.py:718-745) and LU-factored (Block_Tridiagonal).
  3. Ared: Schur complement on the interior horizontal faces of each slab,
     block tridiagonal over the n1-1 face rows (slablu_solver.py:768-882),
     formed exactly as A_rr - sum_j A_r,e_j A_e_j^{-1} A_e_j,r and factored.
  4. Slab Schur complements on [L; R] (the operator get_TXX_slabind samples,
     slablu_solver.py:1066-1211) formed densely with multi-RHS sweeps, and
     the interface block-tridiagonal system over the Npan-1 slab interfaces
     (slablu_solver.py:1215-1308) factored in FP64 (HBS_BlockTridiagonal).
Solve: right-hand side from the Dirichlet data, slab sweeps, interface
sweep, slab sweeps, leaf recovery (slablu_solver.py:489-666).

precision='mixed' runs steps 1-4a in FP32 (interface LU in FP64) and
refines the full discrete system (leaf collocation equations + flux
continuity on the faces) with FP64 residuals until the estimated error of
the face solution falls below `refine_tol` (relative, default 1e-8).  The
estimate is the change c_k of the last step, or, once two corrections after
the first step contract by rho = c_k / c_{k-1} < 0.1, the geometric tail
c_k rho / (1 - rho) that the iterate still lacks.  The default refinement is
flexible GMRES (FP64 Arnoldi on the un-condensed operator, right-
preconditioned by the FP32 two-level solve); refine_method='richardson' runs
plain iterative refinement instead (one FP32 solve per correction).
"""
import time
import warnings
from functools import reduce

import numpy as np
import torch

from .. import hip
from ..src_disc import hps_multidomain_disc
from .block_tridiag_separated import SeparatedBlockTridiagonal, separated_period
from .block_tridiag_solver_dense import Block_Tridiagonal
from .block_tridiag_solver_structured import HBS_BlockTridiagonal
from .hbs_reconstruction import HBS_matrix

L_, R_, D_, U_ = 0, 1, 2, 3


def get_nearest_div(n, x):
    """Divisor of n nearest to x, ties broken towards the larger divisor
    (slablu_solver.py:35-46)."""
    factors = list(reduce(list.__add__, ([i, n // i] for i in range(1, int(n ** 0.5) + 1) if n % i == 0)))
    best, dist = n, abs(n - x)
    for f in factors:
        d = abs(f - x)
        if d < dist or (d == dist and f > best):
            best, dist = f, d
    return int(best)


def _blockdiag(blocks):
    """(..., k, f, f) -> (..., k*f, k*f) block diagonal."""
    *lead, k, f, _ = blocks.shape
    out = torch.zeros(*lead, k, f, k, f, dtype=blocks.dtype, device=blocks.device)
    idx = torch.arange(k, device=blocks.device)
    nl = len(lead)
    perm = (nl,) + tuple(range(nl)) + (nl + 1, nl + 2)
    out[(slice(None),) * nl + (idx, slice(None), idx)] = blocks.permute(*perm)
    return out.reshape(*lead, k * f, k * f)


class Domain_Solver:
    def __init__(self, box_geom, pdo_op, kh, h, p=0, buf_constant=0.5, periodic_bc=False, precision="fp64",
                 device=None, refine_tol=1e-8, refine_maxit=60, refine_method="fgmres", refine_restart=40,
                 check_pivots="warn", on_unconverged="warn", leaf_memory_budget=None, leaf_chunk=None,
                 hbs_rank=None):
        if p == 0:
            raise NotImplementedError("finite-difference discretizations are not provided (HPS only, p > 0)")
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if check_pivots not in ("warn", "raise", "ignore") or on_unconverged not in ("warn", "raise", "ignore"):
            raise ValueError("check_pivots / on_unconverged must be 'warn', 'raise' or 'ignore'")
        self.kh = kh
        # slab Schur complements T_slab: formed exactly (None), or recovered as
        # HBS matrices of this rank from random samples ("auto": the
        # reference's rank rule, slablu_solver.py:1215-1229)
        if hbs_rank is not None and hbs_rank != "auto" and int(hbs_rank) < 1:
            raise ValueError("hbs_rank must be None, 'auto' or a positive rank")
        self.hbs_rank = hbs_rank
        self.periodic_bc = periodic_bc
        self.buf_constant = buf_constant
        self.precision = precision
        # elimination order of the interface system: twisted for FP64 (the
        # Schur complements of growing regions stay well conditioned); the
        # mixed form, which refines against the FP64 residual anyway, takes
        # the nested-dissection order with fewer dependent LU rounds (its
        # diagonal blocks, thin strips near resonance, can be far worse
        # conditioned: about 1e-13 backward error instead of 1e-16)
        self.interface_order = "auto" if precision == "mixed" else "twisted"
        if refine_method not in ("fgmres", "richardson"):
            raise ValueError("refine_method must be 'fgmres' or 'richardson'")
        self.refine_tol, self.refine_maxit = refine_tol, refine_maxit
        self.refine_method, self.refine_restart = refine_method, refine_restart
        # exact zero pivots of the leaf, slab and interface factorizations
        # (LAPACK's info) are read back at the end of build(); the reference
        # leaves them unchecked (torch lu_factor returns inf/nan silently)
        self.check_pivots = check_pivots
        # a refinement that stops at refine_maxit without meeting refine_tol
        self.on_unconverged = on_unconverged
        self.refine_converged = None
        self.refine_history, self.refine_residuals = [], []
        self.solve_info = {}
        # leaf stage under a memory budget (bytes for the kept leaf factors):
        # when all factors do not fit, only the DtN maps are kept and the
        # leaves are factored in chunks that fit, again wherever the factors
        # are needed (slablu_solver.py:781-790 sizes its chunks from free
        # memory); leaf_chunk forces a chunk size
        self.leaf_memory_budget, self.leaf_chunk = leaf_memory_budget, leaf_chunk
        # rem chains eliminated piecewise around separator blocks when long
        # enough (SeparatedBlockTridiagonal); False: one twisted chain each
        self.separated_rem = True
        # stage timings of build() from events on the stream (no host
        # synchronization between stages)
        self.stage_sync = True
        # from the second solve with one factorization on, FGMRES replays its
        # preconditioned operator step (FP32 correction + FP64 residual of
        # the full system) as a captured HIP graph; the first solve stays
        # eager, since capturing costs about two steps
        self.refine_graphs = True
        self._gsteps = {}
        self._solve_graphs = {}
        self.tuning = {}
        self._solves_since_build = 0
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.hps_disc(box_geom, h, p, pdo_op)
        self.hps_panel_split()
        self.disc = "hps"
        self.ntot = self.hps.xx_active.shape[0]
        I_L, I_C, Npan = self.I_L, self.I_C, self.Npan
        I_Ldir = self.inds_pans[0, I_L]
        I_Rdir = self.inds_pans[Npan - 1, self.I_R]
        I_Ddir = self.inds_pans[:, self.I_D].flatten()
        I_Udir = self.inds_pans[:, self.I_U].flatten()
        if periodic_bc:
            # x-periodic (slablu_solver.py:91-123): the last slab's right
            # faces are slab 0's left faces, every slab's left faces form an
            # interface, Dirichlet data only on the bottom and top
            if Npan < 2:
                raise ValueError("periodic boundary conditions need at least 2 slabs")
            self.inds_pans[Npan - 1, self.I_R] = self.inds_pans[0, I_L]
            self.I_slabC = self.inds_pans[:, I_C].flatten()
            self.I_slabX = self.inds_pans[:, I_L].flatten()
            self.I_Xtot = torch.hstack((I_Ddir, I_Udir))
        else:
            self.I_slabC = self.inds_pans[:, I_C].flatten()
            self.I_slabX = self.inds_pans[1:, I_L].flatten()
            self.I_Xtot = torch.hstack((I_Ldir, I_Rdir, I_Ddir, I_Udir))
        self.I_Ctot = torch.hstack((self.I_slabC, self.I_slabX))

    def _Ic(self):
        """I_Ctot on the device (cached: no host-to-device copy per call)."""
        if getattr(self, "_Ic_dev", None) is None:
            self._Ic_dev = self.I_Ctot.to(self.device)
        return self._Ic_dev

    # ------------------------------------------------------------------ setup
    def hps_disc(self, box_geom, a, p, pdo_op):
        """Leaf grid and slab width (slablu_solver.py:127-163)."""
        H = hps_multidomain_disc.HPS_Multidomain(pdo_op, box_geom, a, p, device=self.device,
                                                 periodic=self.periodic_bc)
        f = p - 2
        n0, n1 = H.n0, H.n1
        n_tmp = max(n0, n1) * f - 1
        buf_points = min(400, int(n_tmp ** (2 / 3) * self.buf_constant))
        buf = get_nearest_div(n0, max(int(buf_points / f) + 1, 2))
        if buf < 2:
            raise ValueError("each slab needs at least 2 leaf columns (n0=%d)" % n0)
        Npan = n0 // buf
        nfaces_pan = (2 * n1 + 1) * buf + n1
        self.inds_pans = torch.stack([torch.arange(nfaces_pan * f) + (2 * n1 + 1) * buf * j * f
                                      for j in range(Npan)])
        self.Npan, self.Npan_loc, self.buf_pans = Npan, n1, buf
        self.elim_nblocks, self.elim_bs = buf - 1, f
        self.rem_nblocks, self.rem_bs = n1 - 1, buf * f
        self.hps = H

    def hps_panel_split(self):
        """Slab-local index sets I_L, I_R, I_D, I_U, I_C = [I_elim, I_rem]
        (slablu_solver.py:165-202)."""
        f = self.hps.p - 2
        n1, buf = self.Npan_loc, self.buf_pans
        self.I_L = torch.arange(n1 * f)
        self.I_R = torch.arange(n1 * f) + (2 * n1 + 1) * buf * f
        bo = (2 * n1 + 1) * torch.arange(buf)
        t = torch.arange(f)
        self.I_D = ((bo + n1)[:, None] * f + t).flatten()
        self.I_U = ((bo + 2 * n1)[:, None] * f + t).flatten()
        jj = torch.arange(1, n1)
        I_rem = ((bo[None, :, None] + n1 + jj[:, None, None]) * f + t[None, None, :])
        I_elim = ((bo[None, 1:, None] + torch.arange(n1)[:, None, None]) * f + t[None, None, :])
        self.I_C = torch.hstack((I_elim.flatten(), I_rem.flatten()))

    # ------------------------------------------------------------------ build
    def build(self, sparse_assembly="reduced_gpu", solver_type="slabLU", verbose=True):
        if solver_type not in (None, "slabLU"):
            raise NotImplementedError("solver_type %r: only 'slabLU' is provided" % (solver_type,))
        self.sparse_assembly = sparse_assembly
        self.solver_type = "slabLU"
        self._gsteps = {}  # captured steps read the factors of the previous build
        self._solve_graphs = {}
        self._solves_since_build = 0
        self.wdt = torch.float32 if self.precision == "mixed" else torch.float64
        t = {}
        # stage boundaries are events on the stream; the host never waits in
        # between, so no stage starts behind a drained queue
        ev = {}

        def mark(name):
            if self.stage_sync:
                ev[name] = torch.cuda.Event(enable_timing=True)
                ev[name].record()

        self._mark = mark
        store, chunk = self._leaf_plan()
        torch.cuda.synchronize()
        tic0 = time.time()
        mark("start")
        self.hps.factor_leaves(self.device, self.wdt, chunk=chunk, store=store)
        mark("dtn")
        self.Dt7 = self._dt7(self.hps.leaf["DtN"])
        self.build_Aloc_solver()
        mark("aloc")
        self.construct_Ared()
        mark("ared")
        self.build_Tred_solver()
        mark("tred")
        torch.cuda.synchronize()
        toc_build = time.time() - tic0
        if self.stage_sync:
            el = lambda a, b: ev[a].elapsed_time(ev[b]) / 1e3
            t["toc_DtN"] = el("start", "dtn")
            t["toc_Aloc"] = el("dtn", "aloc")
            t["toc_Ared"] = el("aloc", "ared")
            t["toc_Tred"] = el("ared", "tred")
            t["toc_Tform"] = el("ared", "tform") if "tform" in ev else float("nan")
        else:
            t.update(toc_DtN=float("nan"), toc_Aloc=float("nan"), toc_Ared=float("nan"), toc_Tred=float("nan"),
                     toc_Tform=float("nan"))
        M = self.hps.leaf["M"]
        mem = ((M.numel() * M.element_size() if M is not None else 0) + self.hps.leaf["DtN"].numel() *
               self.hps.leaf["DtN"].element_size() + self.Aloc_solver.nbytes + self.Ared_solver.nbytes +
               (self.Tred_solver.nbytes if self.Tred_solver is not None else 0)) / 1e9
        info = dict(toc_build=toc_build, mem_build=mem, toc_assembly=t["toc_DtN"], **t)
        info["zero_pivots"] = self.zero_pivots()
        if verbose:
            print("SLABLU BUILD (%s): DtN %.4f s, Aloc %.4f s, Ared %.4f s, Tred %.4f s, total %.4f s, %.2f GB"
                  % (self.precision, t["toc_DtN"], t["toc_Aloc"], t["toc_Ared"], t["toc_Tred"], toc_build, mem))
        self.build_info = info
        return info

    def _leaf_plan(self):
        """(keep the leaf factors?, leaves per chunk) under leaf_memory_budget."""
        ni, nx = (self.hps.p - 2) ** 2, 4 * (self.hps.p - 2)
        per_leaf = ni * (ni + nx) * torch.tensor([], dtype=self.wdt).element_size()
        nb = self.hps.nboxes
        budget = self.leaf_memory_budget
        if budget is None or nb * per_leaf <= budget:
            return True, self.leaf_chunk
        if self.precision == "mixed":
            raise ValueError("the mixed-precision refinement reuses the leaf factors in every step: "
                             "leaf_memory_budget (%d bytes) must hold them (%d bytes)" % (budget, nb * per_leaf))
        chunk = max(1, int(budget // per_leaf))
        if self.leaf_chunk is not None:
            chunk = min(chunk, self.leaf_chunk)
        return False, chunk

    def zero_pivots(self):
        """Number of factorizations with an exactly zero pivot in each stage
        (leaf LUs, elim chains, rem chains, interface blocks), from the
        LAPACK-style info arrays the kernels fill; one device-to-host copy.
        A nonzero count warns or raises as check_pivots asks: the solve would
        carry inf/nan."""
        infos = dict(leaf=self.hps.leaf["info"], elim=self.Aloc_solver.info, rem=self.Ared_solver.info)
        if self.Tred_solver is not None:
            infos["interface"] = self.Tred_solver.info
        counts = torch.stack([(v != 0).sum() for v in infos.values()]).cpu().tolist()
        out = dict(zip(infos.keys(), (int(c) for c in counts)))
        bad = {k: v for k, v in out.items() if v}
        if bad and self.check_pivots != "ignore":
            msg = "zero pivots in the factorization (factorizations affected per stage: %s)" % bad
            if self.check_pivots == "raise":
                raise FloatingPointError(msg)
            warnings.warn(msg, RuntimeWarning, stacklevel=3)
        return out

    def _dt7(self, DtN):
        """DtN blocks indexed [slab, slab column, leaf row, a, :, b, :] with
        a, b in (l, r, d, u)."""
        f = self.hps.p - 2
        return DtN.view(self.Npan, self.buf_pans, self.Npan_loc, 4, f, 4, f)

    def _blk(self, c, a, b):
        """(Npan, |c|, n1, f, f) blocks DtN[a, b] of the leaves in slab column(s) c."""
        return self.Dt7[:, c, :, a, :, b, :]

    def elim_blocks(self):
        """Per leaf row, the chain of buf-1 vertical faces: diag_q =
        DtN_q[r,r] + DtN_{q+1}[l,l], sup_q = DtN_{q+1}[l,r], sub_q = DtN_{q+1}[r,l]
        -> (Npan*n1, m, f, f), (Npan*n1, m-1, f, f) x 2."""
        Np, n1, buf, f = self.Npan, self.Npan_loc, self.buf_pans, self.hps.p - 2
        perm = lambda x: x.permute(0, 2, 1, 3, 4).reshape(Np * n1, -1, f, f)
        diag = perm(self._blk(slice(0, buf - 1), R_, R_) + self._blk(slice(1, buf), L_, L_))
        sup = perm(self._blk(slice(1, buf - 1), L_, R_))
        sub = perm(self._blk(slice(1, buf - 1), R_, L_))
        return diag, sub, sup

    def rem_blocks(self):
        """A_rr of each slab: block tridiagonal over the n1-1 interior face
        rows, each block block-diagonal over the slab columns ->
        diag (Npan, n1-1, B, B), sup/sub (Npan, n1-2, B, B)."""
        n1 = self.Npan_loc
        mr = n1 - 1
        rows, nxt = slice(0, n1 - 1), slice(1, n1)
        bd = lambda x: _blockdiag(x.permute(0, 2, 1, 3, 4))
        d = bd(self.Dt7[:, :, rows, U_, :, U_, :] + self.Dt7[:, :, nxt, D_, :, D_, :])
        u = bd(self.Dt7[:, :, nxt, D_, :, U_, :])[:, :mr - 1]
        l = bd(self.Dt7[:, :, nxt, U_, :, D_, :])[:, :mr - 1]
        return d, u, l

    def interface_blocks(self):
        """A restricted to each slab interface (block diagonal over the leaf
        rows): (Npan-1, n1*f, n1*f); periodic: (Npan, ...) with the seam
        (interface 0, the left faces of slab 0) first."""
        Np, buf = self.Npan, self.buf_pans
        if self.periodic_bc:
            left = torch.roll(self.Dt7[:, buf - 1, :, R_, :, R_, :], 1, dims=0)
            return _blockdiag(left + self.Dt7[:, 0, :, L_, :, L_, :])
        return _blockdiag(self.Dt7[:Np - 1, buf - 1, :, R_, :, R_, :] + self.Dt7[1:, 0, :, L_, :, L_, :])

    def build_Aloc_solver(self):
        Np, n1, f = self.Npan, self.Npan_loc, self.hps.p - 2
        diag, sub, sup = self.elim_blocks()
        # FP32 (mixed precision): twisted elimination, half the dependent
        # block steps in the factorization and in every refinement sweep
        mixed = self.precision == "mixed"
        self.Aloc_solver = Block_Tridiagonal(Np * n1, self.buf_pans - 1, f, twisted=mixed)
        self.Aloc_solver.build_sweep(diag, sub, sup)

    def _couplings(self):
        """Dense per-leaf-row coupling blocks between the elim chain of leaf
        row j, the local rem faces [bottom faces; top faces] (2*buf*f) and the
        slab interfaces [L_j; R_j] (2f), for all (slab, row)."""
        Np, n1, buf, f = self.Npan, self.Npan_loc, self.buf_pans, self.hps.p - 2
        m, B = buf - 1, buf * f
        dt, dev = self.wdt, self.device
        q = torch.arange(m, device=dev)
        # R = A[elim, rem_local]: (Np, n1, m, f, 2, buf, f)
        Rm = torch.zeros(Np, n1, m, f, 2, buf, f, dtype=dt, device=dev)
        Cm = torch.zeros(Np, n1, 2, buf, f, m, f, dtype=dt, device=dev)
        for s_, side in ((0, D_), (1, U_)):
            # l face of leaf q+1 is elim face q; r face of leaf q is elim face q
            Rm[:, :, q, :, s_, q + 1, :] = self._blk(slice(1, buf), L_, side).permute(1, 0, 2, 3, 4)
            Rm[:, :, q, :, s_, q, :] = self._blk(slice(0, buf - 1), R_, side).permute(1, 0, 2, 3, 4)
            Cm[:, :, s_, q + 1, :, q, :] = self._blk(slice(1, buf), side, L_).permute(1, 0, 2, 3, 4)
            Cm[:, :, s_, q, :, q, :] = self._blk(slice(0, buf - 1), side, R_).permute(1, 0, 2, 3, 4)
        # E = A[elim, X_j], AXe = A[X_j, elim]
        Em = torch.zeros(Np, n1, m, f, 2, f, dtype=dt, device=dev)
        AXe = torch.zeros(Np, n1, 2, f, m, f, dtype=dt, device=dev)
        Em[:, :, 0, :, 0, :] = self._blk(0, R_, L_)
        Em[:, :, m - 1, :, 1, :] = self._blk(buf - 1, L_, R_)
        AXe[:, :, 0, :, 0, :] = self._blk(0, L_, R_)
        AXe[:, :, 1, :, m - 1, :] = self._blk(buf - 1, R_, L_)
        # Xr = A[rem_local, X_j], AXr = A[X_j, rem_local]
        Xr = torch.zeros(Np, n1, 2, buf, f, 2, f, dtype=dt, device=dev)
        AXr = torch.zeros(Np, n1, 2, f, 2, buf, f, dtype=dt, device=dev)
        for s_, side in ((0, D_), (1, U_)):
            Xr[:, :, s_, 0, :, 0, :] = self._blk(0, side, L_)
            Xr[:, :, s_, buf - 1, :, 1, :] = self._blk(buf - 1, side, R_)
            AXr[:, :, 0, :, s_, 0, :] = self._blk(0, L_, side)
            AXr[:, :, 1, :, s_, buf - 1, :] = self._blk(buf - 1, R_, side)
        nb = Np * n1
        self.cR = Rm.reshape(nb, m * f, 2 * B)
        self.cC = Cm.reshape(nb, 2 * B, m * f)
        self.cE = Em.reshape(nb, m * f, 2 * f)
        self.cAXe = AXe.reshape(nb, 2 * f, m * f)
        self.cXr = Xr.reshape(nb, 2 * B, 2 * f)
        self.cAXr = AXr.reshape(nb, 2 * f, 2 * B)

    def construct_Ared(self):
        """Schur complement of each slab onto its interior horizontal faces,
        block tridiagonal over the n1-1 face rows (blocks buf*f), factored."""
        diag, sub, sup = self.rem_schur_blocks()
        m, B = self.Npan_loc - 1, self.rem_bs
        # FP32 (mixed precision): triangular inverses of the diagonal blocks,
        # so the T_slab and refinement sweeps are block products; the FP64
        # path keeps the substitutions
        inv = self.precision == "mixed"
        d = separated_period(m, B, self.wdt) if self.separated_rem else None
        if d is not None:
            # pieces of d-1 blocks between separator blocks: far fewer
            # dependent block steps than one chain (block_tridiag_separated)
            self.Ared_solver = SeparatedBlockTridiagonal(self.Npan, m, B, d, inv_factors=inv)
        else:
            self.Ared_solver = Block_Tridiagonal(self.Npan, m, B, twisted=True, inv_factors=inv)
        self.Ared_solver.build_sweep(diag, sub, sup)

    def rem_schur_blocks(self):
        """(diag (Npan, n1-1, B, B), sub, sup (Npan, n1-2, B, B)) of the rem
        Schur complements A_rr - sum_j A_r,e_j A_e_j^{-1} A_e_j,r, before
        factorization (slablu_solver.py:768-882); also keeps the coupling
        solves W, Y and C Y the later stages use."""
        Np, n1, buf, f = self.Npan, self.Npan_loc, self.buf_pans, self.hps.p - 2
        m, B = buf - 1, buf * f
        self._couplings()
        nb = Np * n1
        # [W | Y] = A_ee^{-1} [R | E]
        RE = torch.cat((self.cR, self.cE), dim=2).contiguous()
        self.Aloc_solver.lusolve_sweep(RE)
        self.cW = RE[:, :, :2 * B]
        self.cY = RE[:, :, 2 * B:]
        CWY = torch.empty(nb, 2 * B, 2 * B + 2 * f, dtype=self.wdt, device=self.device)
        hip.gemm(self.cC, RE, CWY)
        CW = CWY[:, :, :2 * B].reshape(Np, n1, 2, B, 2, B)
        self.cCY = CWY[:, :, 2 * B:]
        mr = n1 - 1
        Arr_d, Arr_u, Arr_l = self.rem_blocks()
        diag = Arr_d - CW[:, :mr, 1, :, 1, :] - CW[:, 1:, 0, :, 0, :]
        sup = Arr_u - CW[:, 1:mr, 0, :, 1, :]
        sub = Arr_l - CW[:, 1:mr, 1, :, 0, :]
        return diag.contiguous(), sub.contiguous(), sup.contiguous()

    def _slab_u_local(self, upad, s):
        """(n1, 2B, ncols) view of the padded rem solution of slab s: rows of
        leaf row j are [bottom faces; top faces] = padded blocks j, j+1."""
        n1, B = self.Npan_loc, self.rem_bs
        nc = upad.shape[-1]
        return torch.as_strided(upad, (n1, 2 * B, nc), (B * nc, nc, 1),
                                upad.storage_offset() + s * upad.stride(0))

    def build_Tred_solver(self):
        """Dense slab Schur complements on [L; R] and the interface system."""
        Np, n1, buf, f = self.Npan, self.Npan_loc, self.buf_pans, self.hps.p - 2
        B, L = self.rem_bs, n1 * f
        nb = Np * n1
        dt, dev = self.wdt, self.device
        if Np < 2:
            self.Tred_solver = None
            return
        if self.hbs_rank is not None:
            Kf = self._Kf_hbs()
            if getattr(self, "_mark", None) is not None:
                self._mark("tform")
            self._interface_from_slabs(Kf)
            return
        # G = A[rem, X_j] - A[rem, e_j] Y_j, scattered into the padded rem rows
        G = self.cXr - self.cCY
        upad = torch.zeros(Np, (n1 + 1) * B, 2 * L, dtype=dt, device=dev)
        g6 = upad.view(Np, n1 + 1, B, 2, n1, f)
        G6 = G.reshape(Np, n1, 2, B, 2, f)
        ja = torch.arange(n1, device=dev)
        for sg in (0, 1):
            g6[:, ja + sg, :, :, ja, :] = G6[:, :, sg].permute(1, 0, 2, 3, 4)
        u = upad[:, B:n1 * B]
        self.Ared_solver.lusolve_sweep(u)
        # the padding blocks stand for the Dirichlet rows D and U: zero them
        upad[:, :B] = 0
        upad[:, n1 * B:] = 0
        # K_j = (A[X_j, rem] - A[X_j, e_j] W_j) u_local_j + A[X_j, e_j] Y_j
        H = self.cAXr.clone()
        hip.gemm(self.cAXe, self.cW, H, alpha=-1.0, beta=1.0)
        self.cH = H
        K = torch.empty(Np, n1, 2 * f, 2 * L, dtype=dt, device=dev)
        for s in range(Np):
            hip.gemm(H[s * n1:(s + 1) * n1], self._slab_u_local(upad, s), K[s])
        Q = torch.empty(nb, 2 * f, 2 * f, dtype=dt, device=dev)
        hip.gemm(self.cAXe, self.cY, Q)
        K6 = K.view(Np, n1, 2 * f, 2, n1, f)
        K6[:, ja, :, :, ja, :] += Q.view(Np, n1, 2 * f, 2, f).permute(1, 0, 2, 3, 4)
        Kf = K.view(Np, n1, 2, f, 2 * L).permute(0, 2, 1, 3, 4).reshape(Np, 2 * L, 2 * L)
        del upad, K
        if getattr(self, "_mark", None) is not None:
            self._mark("tform")
        self._interface_from_slabs(Kf)

    def _interface_from_slabs(self, Kf):
        """The interface block-tridiagonal system from the slab Schur
        complements Kf (Np, 2L, 2L) on [L faces; R faces], factored."""
        Np, L, dev = self.Npan, self.Npan_loc * (self.hps.p - 2), self.device
        Aii = self.interface_blocks()
        KR = Kf[:, L:, L:].to(torch.float64)
        KL = Kf[:, :L, :L].to(torch.float64)
        if self.periodic_bc:
            # interface j = left faces of slab j (j = 0 is the seam); slab j
            # couples interface j to interface j+1 (mod Npan)
            # (slablu_solver.py:1267-1305)
            diag = Aii.to(torch.float64) - torch.roll(KR, 1, dims=0) - KL
            sup = -Kf[:, :L, L:].to(torch.float64)
            sub = -Kf[:, L:, :L].to(torch.float64)
            self.Tred_solver = HBS_BlockTridiagonal(Np, L, cyclic=True, order=self.interface_order)
        else:
            # interface s between slab s and s+1
            diag = Aii.to(torch.float64) - KR[:Np - 1] - KL[1:]
            sup = -Kf[1:Np - 1, :L, L:].to(torch.float64)
            sub = -Kf[1:Np - 1, L:, :L].to(torch.float64)
            self.Tred_solver = HBS_BlockTridiagonal(Np - 1, L, order=self.interface_order)
        self.Tred_solver.build_lowmem(diag, None, sub, sup, device=dev, dtype=torch.float64)

    def hbs_rank_used(self):
        """Rank of the T_slab compression: 4 x the waves across a slab +
        1.5 p + 30 for "auto" (slablu_solver.py:1215-1229), else hbs_rank."""
        if self.hbs_rank == "auto":
            f = self.hps.p - 2
            buf, L = self.buf_pans * f, self.Npan_loc * f
            nwaves_thin = (self.kh / (2 * np.pi)) * (buf / L)
            return int(4 * nwaves_thin) + int(1.5 * self.hps.p) + 30
        return int(self.hbs_rank)

    def _slab_schur_apply(self, x, transpose=False):
        """K_s x (or K_s^T x) for every slab s, K_s = A[X_s, C_s] A_CC,s^{-1}
        A[C_s, X_s] the slab's Schur complement on its faces:
        x (Np, 2L, nc) [left faces; right faces] -> (Np, 2L, nc)."""
        Np, n1, f = self.Npan, self.Npan_loc, self.elim_bs
        nc = x.shape[-1]
        xs = x.view(Np, 2, n1, f, nc).permute(0, 2, 1, 3, 4).contiguous()
        v = self._ACX(xs.view(Np, n1, 2 * f, nc), transpose)
        _, ue, upad = self._slab_solve(v, transpose, want_sol=False)
        Z = self._AXC(ue, upad, transpose)
        return Z.permute(0, 2, 1, 3, 4).reshape(Np, 2 * n1 * f, nc)

    def _Kf_hbs(self):
        """Slab Schur complements recovered as HBS matrices from random
        samples (slablu_solver.py:1139-1213): K_s applied to
        blockdiag(Omega, Omega) and K_s^T to blockdiag(Psi, Psi) through the
        slab solves, the four L x L blocks of every slab reconstructed as one
        batch of 4 Npan HBS matrices (kept in self.Tslab_hbs), then densified
        for the interface factorization as the reference's build_lowmem
        does.  FP64 (Np, 2L, 2L)."""
        Np, L, dev, dt = self.Npan, self.Npan_loc * (self.hps.p - 2), self.device, self.wdt
        k = self.hbs_rank_used()
        hbs = HBS_matrix(4 * Np, L, k)
        ns = hbs.bs + k
        if ns > L or hbs.m < 2:
            raise ValueError("HBS rank %d: %d samples per side and %d leaf blocks for %d interface points "
                             "(needs at most L samples and two leaf blocks)" % (k, ns, hbs.m, L))
        g = torch.Generator(device=dev)
        g.manual_seed(20240613)
        Omega = torch.rand(L, ns, generator=g, dtype=torch.float64, device=dev)
        Psi = torch.rand(L, ns, generator=g, dtype=torch.float64, device=dev)
        X = torch.zeros(Np, 2 * L, 2 * ns, dtype=dt, device=dev)
        X[:, :L, :ns] = Omega.to(dt)
        X[:, L:, ns:] = Omega.to(dt)
        Yf = self._slab_schur_apply(X).double()
        X[:, :L, :ns] = Psi.to(dt)
        X[:, L:, ns:] = Psi.to(dt)
        Zf = self._slab_schur_apply(X, transpose=True).double()
        del X
        # blocks LL, LR, RL, RR of every slab: Y = K Omega, Z = K^T Psi
        Y = torch.stack((Yf[:, :L, :ns], Yf[:, :L, ns:], Yf[:, L:, :ns], Yf[:, L:, ns:]), 1)
        Z = torch.stack((Zf[:, :L, :ns], Zf[:, L:, :ns], Zf[:, :L, ns:], Zf[:, L:, ns:]), 1)
        hbs.reconstruct(Omega[None], Psi[None], Y.reshape(4 * Np, L, ns), Z.reshape(4 * Np, L, ns))
        self.Tslab_hbs = hbs
        Kd = hbs.todense().view(Np, 2, 2, L, L)
        return Kd.permute(0, 1, 3, 2, 4).reshape(Np, 2 * L, 2 * L)

    # ------------------------------------------------------------------ solve
    def _u_local_all(self, upad):
        Np, n1, B = upad.shape[0], self.Npan_loc, self.rem_bs
        nc = upad.shape[-1]
        return torch.as_strided(upad, (Np, n1, 2 * B, nc), (upad.stride(0), B * nc, nc, 1),
                                upad.storage_offset()).reshape(Np * n1, 2 * B, nc)

    def _scatter_rem(self, CT, out=None):
        """Sum per-leaf-row [bottom; top] rem contributions (Np*n1, 2B, nc)
        into the rem rows of each slab -> (Np, (n1-1)*B, nc): rem block j
        (between leaf rows j and j+1) gets the top part of leaf row j+1 and
        the bottom part of leaf row j; written to out if given."""
        n1, B = self.Npan_loc, self.rem_bs
        Np = CT.shape[0] // n1
        nc = CT.shape[-1]
        C5 = CT.view(Np, n1, 2, B, nc)
        if out is None:
            return torch.add(C5[:, 1:, 0], C5[:, :-1, 1]).view(Np, (n1 - 1) * B, nc)
        torch.add(C5[:, 1:, 0], C5[:, :-1, 1], out=out.view(Np, n1 - 1, B, nc))
        return out

    def _slab_solve(self, f_C, transpose=False, s0=0, s1=None, want_sol=True, out=None):
        """A_CC(slab)^{-1} f (transpose: A_CC(slab)^{-T} f) for the slabs
        s0..s1-1 through elim -> rem -> elim (slablu_solver.py:901-943);
        f_C (s1 - s0, |I_C|, nrhs) in the working precision.  Returns (sol,
        elim part (rows, m f, nrhs), padded rem part (slabs, (n1+1) B, nrhs));
        sol is None without want_sol, and written to out if given.
        Plain: t = A_ee^{-1} f_e, u = Ared^{-1}(f_r - A_re t),
        x_e = t - (A_ee^{-1} A_er) u.  Transposed: t = A_ee^{-T} f_e,
        u = Ared^{-T}(f_r - A_er^T t), x_e = A_ee^{-T}(f_e - A_re^T u)."""
        n1, m, f, B = self.Npan_loc, self.elim_nblocks, self.elim_bs, self.rem_bs
        s1 = self.Npan if s1 is None else s1
        Np = s1 - s0
        rows = slice(s0 * n1, s1 * n1)
        ne = n1 * m * f
        nc = f_C.shape[-1]
        fe = f_C[:, :ne].reshape(Np * n1, m * f, nc).contiguous()
        fr = f_C[:, ne:]
        # the elim sweep runs in place: on fe itself when fe is already a
        # private copy of f_C's rows (and the plain solve needs it no more)
        private = fe.untyped_storage().data_ptr() != f_C.untyped_storage().data_ptr()
        t = self.Aloc_solver.lusolve_sweep(fe if private and not transpose else fe.clone(), b_start=s0 * n1,
                                           b_end=s1 * n1, transpose=transpose)
        CT = torch.empty(Np * n1, 2 * B, nc, dtype=t.dtype, device=t.device)
        hip.gemm(self.cR[rows].transpose(1, 2) if transpose else self.cC[rows], t, CT)
        # padded rem right-hand side f_r - (scattered couplings), zero blocks
        # at both ends
        upad = torch.zeros(Np, (n1 + 1) * B, nc, dtype=f_C.dtype, device=f_C.device)
        g = self._scatter_rem(CT, out=upad[:, B:n1 * B])
        torch.sub(fr, g, out=g)
        self.Ared_solver.lusolve_sweep(upad[:, B:n1 * B], b_start=s0, b_end=s1, transpose=transpose)
        ul = self._u_local_all(upad)
        if transpose:
            ue = fe.clone()
            hip.gemm(self.cC[rows].transpose(1, 2), ul, ue, alpha=-1.0, beta=1.0)
            self.Aloc_solver.lusolve_sweep(ue, b_start=s0 * n1, b_end=s1 * n1, transpose=True)
        else:
            ue = t
            hip.gemm(self.cW[rows], ul, ue, alpha=-1.0, beta=1.0)
        sol = None
        if want_sol:
            sol = torch.empty_like(f_C) if out is None else out
            sol[:, :ne] = ue.view(Np, ne, nc)
            sol[:, ne:] = upad[:, B:n1 * B]
        return sol, ue, upad

    def solve_slab(self, Npan_start, Npan_end, f, transpose=False):
        """A_CC(slab)^{-1} f (or A_CC(slab)^{-T} f) for the slabs
        Npan_start..Npan_end-1; f (Npan_end - Npan_start, |I_C|, nrhs), in any
        dtype and device, returned in f's (slablu_solver.py:900-943)."""
        fd = torch.as_tensor(f).to(self.device, self.wdt)
        sol, _, _ = self._slab_solve(fd, transpose, Npan_start, Npan_end)
        return sol.to(f.device, f.dtype) if torch.is_tensor(f) else sol

    def _AXC(self, ue, upad, transpose=False):
        """Z_j = A[X_j, C] y for every (slab, leaf row), or (A[C, X_j])^T y
        with transpose: (Np, n1, 2, f, nc)."""
        n1, f = self.Npan_loc, self.elim_bs
        nc = ue.shape[-1]
        Np = ue.shape[0] // n1
        Ae, Ar = (self.cE.transpose(1, 2), self.cXr.transpose(1, 2)) if transpose else (self.cAXe, self.cAXr)
        Z = torch.empty(Np * n1, 2 * f, nc, dtype=ue.dtype, device=ue.device)
        hip.gemm(Ae, ue, Z)
        hip.gemm(Ar, self._u_local_all(upad), Z, beta=1.0)
        return Z.view(Np, n1, 2, f, nc)

    def _ACX(self, x, transpose=False):
        """A[C, X] x for every slab from x (Np, n1, 2f, nc), or (A[X, C])^T x
        with transpose -> (Np, |I_C|, nc)."""
        Np, n1, m, f = self.Npan, self.Npan_loc, self.elim_nblocks, self.elim_bs
        nc = x.shape[-1]
        Ae, Ar = (self.cAXe.transpose(1, 2), self.cAXr.transpose(1, 2)) if transpose else (self.cE, self.cXr)
        xe = x.reshape(Np * n1, 2 * f, nc).contiguous()
        e = torch.empty(Np * n1, m * f, nc, dtype=x.dtype, device=x.device)
        hip.gemm(Ae, xe, e)
        r = torch.empty(Np * n1, 2 * self.rem_bs, nc, dtype=x.dtype, device=x.device)
        hip.gemm(Ar, xe, r)
        out = torch.empty(Np, len(self.I_C), nc, dtype=x.dtype, device=x.device)
        out[:, :n1 * m * f] = e.view(Np, n1 * m * f, nc)
        self._scatter_rem(r, out=out[:, n1 * m * f:])
        return out

    def _interface_to_slabs(self, sx, dtype):
        """Interface values (interfaces, n1, f, nc) -> per-slab [L; R] face
        values (Np, n1, 2, f, nc) (zero on the Dirichlet sides)."""
        Np, n1, f = self.Npan, self.Npan_loc, self.elim_bs
        nc = sx.shape[-1]
        if self.periodic_bc:
            x = torch.empty(Np, n1, 2, f, nc, dtype=dtype, device=sx.device)
            x[:, :, 0] = sx.to(dtype)
            x[:, :, 1] = torch.roll(sx, -1, dims=0).to(dtype)
        else:
            x = torch.zeros(Np, n1, 2, f, nc, dtype=dtype, device=sx.device)
            x[1:, :, 0] = sx.to(dtype)
            x[:Np - 1, :, 1] = sx.to(dtype)
        return x

    def _slabs_to_interface(self, Z):
        """Per-slab [L; R] contributions (Np, n1, 2, f, nc) summed onto the
        interfaces -> (interfaces, n1, f, nc) in FP64 (interface j sees slab
        j-1 from the left and slab j from the right)."""
        Np = self.Npan
        if self.periodic_bc:
            return torch.roll(Z[:, :, 1], 1, dims=0).double() + Z[:, :, 0].double()
        return Z[:Np - 1, :, 1].double() + Z[1:, :, 0].double()

    def _interface_rhs(self, fX, Z):
        """fX - (per-slab contributions Z summed onto the interfaces), in
        FP64: fX (interfaces, n1, f, nc) in the working precision, Z as
        _slabs_to_interface takes it."""
        if self.periodic_bc:
            return fX.to(torch.float64) - self._slabs_to_interface(Z)
        Np = self.Npan
        acc = Z[:Np - 1, :, 1].to(torch.float64, copy=True)
        acc += Z[1:, :, 0]
        return torch.sub(fX, acc, out=acc)

    def solve_reduced(self, ff, transpose=False):
        """A_CC x = ff on the face unknowns I_Ctot (slablu_solver.py:543-590),
        or A_CC^T x = ff with transpose; ff (|I_Ctot|, nrhs) in the working
        precision (interface in FP64)."""
        Np, n1, f = self.Npan, self.Npan_loc, self.elim_bs
        nC = len(self.I_C)
        L = n1 * f
        nc = ff.shape[-1]
        ffC = ff[:Np * nC].view(Np, nC, nc)
        yC, ue, upad = self._slab_solve(ffC, transpose, want_sol=Np == 1)
        if Np == 1:
            return yC.reshape(Np * nC, nc)
        nI = Np if self.periodic_bc else Np - 1
        ffX = self._interface_rhs(ff[Np * nC:].view(nI, n1, f, nc), self._AXC(ue, upad, transpose))
        solX = self.Tred_solver.solve(ffX.reshape(nI * L, nc), transpose=transpose)
        x = self._interface_to_slabs(solX.view(nI, n1, f, nc), ff.dtype)
        out = torch.empty(Np * nC + nI * L, nc, dtype=ff.dtype, device=ff.device)
        acx = self._ACX(x.view(Np, n1, 2 * f, nc), transpose)
        self._slab_solve(torch.sub(ffC, acx, out=acx), transpose, out=out[:Np * nC].view(Np, nC, nc))
        out[Np * nC:] = solX
        return out

    def apply_Tred(self, rhs, transpose=False):
        """Action of the interface Schur complement T = A_XX - A_XC A_CC^{-1}
        A_CX (or T^T) on rhs (|I_slabX|, nrhs) (slablu_solver.py:963-1025),
        through the slab solves; FP64 result."""
        Np, n1, f = self.Npan, self.Npan_loc, self.elim_bs
        nI = Np if self.periodic_bc else Np - 1
        x64 = torch.as_tensor(rhs).to(self.device, torch.float64)
        nc = x64.shape[-1]
        sx = x64.view(nI, n1, f, nc)
        x = self._interface_to_slabs(sx, self.wdt)
        v = self._ACX(x.view(Np, n1, 2 * f, nc), transpose)
        _, ue, upad = self._slab_solve(v, transpose, want_sol=False)
        Z = self._slabs_to_interface(self._AXC(ue, upad, transpose))
        Aii = self.interface_blocks().to(torch.float64)
        if transpose:
            Aii = Aii.transpose(1, 2)
        y = torch.empty(nI, n1 * f, nc, dtype=torch.float64, device=self.device)
        hip.gemm(Aii, x64.view(nI, n1 * f, nc), y)
        return (y - Z.reshape(nI, n1 * f, nc)).reshape(nI * n1 * f, nc)

    def get_rhs(self, uu_dir_func, ff_body_func=None):
        """ff = -A[Ctot, Xtot] u_dir (slablu_solver.py:489-526) on the device
        (FP64), plus the Dirichlet values."""
        xx = self.hps.xx_active[self.I_Xtot]
        uu_dir = torch.as_tensor(uu_dir_func(xx)).to(self.device, torch.float64)
        if uu_dir.dim() == 1:
            uu_dir = uu_dir[:, None]
        x = torch.zeros(self.hps.nface, uu_dir.shape[1], dtype=torch.float64, device=self.device)
        x[self.I_Xtot.to(self.device)] = uu_dir
        return x, uu_dir

    def _face_residual(self, x_face, DtN):
        Ic = self._Ic()
        return self.hps.apply_A(x_face, DtN)[Ic]

    def solve_helper(self, uu_dir_func, ff_body_func=None):
        """Two-level solve; returns (face solution with Dirichlet values,
        relative residual of the reduced system, time)."""
        dev = self.device
        Ic = self._Ic()
        torch.cuda.synchronize()
        tic = time.time()
        x_face, uu_dir = self.get_rhs(uu_dir_func)
        DtN = self.hps.leaf["DtN"]
        ff = -self._face_residual(x_face, DtN)
        if ff_body_func is not None:
            fb = self.hps.body_load(ff_body_func, ff.shape[-1], DtN.dtype)
            ff = ff + self._body_reduce(fb)[Ic]
        sol = self.solve_reduced(ff.to(self.wdt))
        x_face[Ic] = sol.to(torch.float64)
        torch.cuda.synchronize()
        toc = time.time() - tic
        xc = torch.zeros_like(x_face)
        xc[Ic] = x_face[Ic]
        res = self._face_residual(xc, DtN).double() - ff.double()
        rel = (torch.linalg.norm(res) / torch.linalg.norm(ff.double())).item()
        return x_face, rel, toc

    def _body_reduce(self, fb, return_solve=False):
        """Face load -sum_b P_b N_c A_cc^{-1} f_b of interior loads fb
        (nboxes, ni, nrhs) (the reduction reduce_body intends,
        hps_parallel_leaf_ops.py:94-100); with return_solve, also the leaf
        solves A_cc^{-1} f_b."""
        ops, lf = self.hps.ops, self.hps.leaf
        ni = ops.ni
        w = fb.to(lf["dtype"], copy=True).contiguous()
        for s, e, M, ipiv, perm in self.hps.leaf_chunks():
            hip.getrs(M[:, :, :ni], ipiv, w[s:e], perm=perm)
        z = torch.empty(fb.shape[0], ops.nx, fb.shape[-1], dtype=w.dtype, device=w.device)
        hip.gemm(ops.typed(w.dtype)["Nx_c"], w, z, alpha=-1.0)
        y = torch.zeros(self.hps.nface, fb.shape[-1], dtype=w.dtype, device=w.device)
        y.index_add_(0, self.hps.Iext_d.reshape(-1), z.reshape(-1, fb.shape[-1]))
        return (y, w) if return_solve else y

    def _solve_face_device(self, x0):
        """The FP64 solve from the Dirichlet lift x0 (face vector, Dirichlet
        values at I_Xtot, zero elsewhere; nrhs columns) with no host
        synchronization, so that it can be captured: right-hand side -A x0,
        two-level solve, face solution, leaf recovery.  Returns (face
        solution, node values (nboxes, p^2, nrhs))."""
        Ic = self._Ic()
        ff = -self._face_residual(x0, self.hps.leaf["DtN"])
        sol = self.solve_reduced(ff.to(self.wdt))
        x = x0.clone()
        x[Ic] = sol.to(torch.float64)
        return x, self.hps.leaf_values(x)

    def solve_graph(self, uu_dir_func):
        """FP64 solve of all right-hand-side columns at once as one replay of
        a captured HIP graph (captured at the first call per column count of
        a factorization; the Dirichlet data are evaluated on the host
        before).  Returns the node values (nboxes*p^2, nrhs) on the device;
        the same values as solve() computes."""
        if self.precision != "fp64":
            raise ValueError("solve_graph runs the FP64 form (the mixed form refines with host decisions per step)")
        x0, _ = self.get_rhs(uu_dir_func)
        nc = x0.shape[1]
        g = self._solve_graphs.get(nc)
        if g is None:
            xin = torch.zeros_like(x0)
            cur = torch.cuda.current_stream(self.device)
            side = torch.cuda.Stream(self.device)
            side.wait_stream(cur)
            with torch.cuda.stream(side):                   # warm-up outside the capture
                for _ in range(2):
                    self._solve_face_device(xin)
            cur.wait_stream(side)
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                xo, uo = self._solve_face_device(xin)
            g = (graph, xin, xo, uo)
            self._solve_graphs[nc] = g
        graph, xin, xo, uo = g
        xin.copy_(x0)
        graph.replay()
        self.last_face = xo
        self.sol_dev = uo.reshape(-1, nc)
        return self.sol_dev

    def solve(self, uu_dir_func, ff_body_func=None, known_sol=False):
        """Returns (sol_tot at all leaf nodes (cpu), rel_err, true_err,
        resloc_hps, toc_solve) as slablu_solver.py:623-666 does."""
        if self.precision == "mixed":
            uu, rel, toc, resloc = self.solve_refined(uu_dir_func, ff_body_func)
            self.sol_dev = uu
        else:
            x_face, rel, toc = self.solve_helper(uu_dir_func, ff_body_func)
            torch.cuda.synchronize()
            tic = time.time()
            uu, resloc = self.hps.solve(self.device, x_face, ff_body_func)
            torch.cuda.synchronize()
            toc += time.time() - tic
            self.sol_dev = uu
        true_err = float("nan")
        if known_sol:
            uu_true = torch.as_tensor(uu_dir_func(self.hps.xx_tot.clone())).to(self.device, torch.float64)
            true_err = (torch.linalg.norm(uu.double() - uu_true) / torch.linalg.norm(uu_true)).item()
        return uu.double().cpu(), rel, true_err, resloc, toc

    # ------------------------------------------------------- mixed precision
    def full_residual(self, x_face, uc, uu=None, fb=None):
        """FP64 residual of the un-condensed system at the iterate given by
        the face vector x_face (Dirichlet values included) and the leaf
        interior values uc (nboxes, ni, nc): r_I = f - A_loc u at the leaf
        interiors (f = fb, the body load at the interior nodes, or zero),
        r_F = -(sum of outward fluxes) at the face DOFs I_Ctot.  One fused
        leaf launch; uu, if given, receives all node values."""
        hps = self.hps
        nc = x_face.shape[-1]
        rI, fl = hps.ops.residual(uc, hps.faces_to_boxes(x_face), hps.coefs, uu=uu)
        if fb is not None:
            rI += fb
        y = torch.zeros(hps.nface, nc, dtype=torch.float64, device=self.device)
        y.index_add_(0, hps.Iext_d.reshape(-1), fl.reshape(-1, nc))
        return rI, -y[self._Ic()]

    def _correction(self, rI, rF):
        """FP32 two-level solve of A d = r for a residual (rI, rF) of the full
        system: the face correction from the condensed system (interior
        residuals folded in through the leaf solves), then the leaf interiors.
        Returns (dI (nboxes, ni, nc), dC (|I_Ctot|, nc)) in FP64."""
        hps, ops, lf = self.hps, self.hps.ops, self.hps.leaf
        Ic = self._Ic()
        yb, w = self._body_reduce(rI, return_solve=True)
        ff = torch.add(rF, yb[Ic])
        dC = self.solve_reduced(ff.to(self.wdt)).double()
        dF = torch.zeros(hps.nface, dC.shape[-1], dtype=torch.float64, device=self.device)
        dF[Ic] = dC
        db = hps.faces_to_boxes(dF).to(self.wdt).contiguous()
        # A_cc^{-1}(r_I - A_cx d_x) = w - S' d_x with the stored S' = A_cc^{-1} A_cx
        hip.gemm(lf["M"][:, :, ops.ni:], db, w, alpha=-1.0, beta=1.0)
        return w.double(), dC

    def _apply_full(self, dI, dC):
        """A d of the un-condensed system (FP64) for a vector with zero
        Dirichlet data: interior values dI, face values dC on I_Ctot."""
        hps = self.hps
        Ic = self._Ic()
        nc = dC.shape[-1]
        x = torch.zeros(hps.nface, nc, dtype=torch.float64, device=self.device)
        x[Ic] = dC
        # full_residual with no load: r_I = -(A d)_I, and the summed outward
        # fluxes y give (A d)_F = y[I_Ctot]
        rI, fl = hps.ops.residual(dI, hps.faces_to_boxes(x), hps.coefs)
        y = torch.zeros(hps.nface, nc, dtype=torch.float64, device=self.device)
        y.index_add_(0, hps.Iext_d.reshape(-1), fl.reshape(-1, nc))
        return rI.neg_(), y[Ic]

    def _converged(self, c, prev):
        """Stop test on the relative change c of the last step (prev: the
        step before, or None): the error left after a geometrically
        converging step is about c rho / (1 - rho), rho = c / prev, used
        when rho < 0.1; otherwise c itself."""
        if prev is not None and prev > 0.0:
            rho = c / prev
            if rho < 0.1:
                c = c * rho / (1.0 - rho)
        return c <= self.refine_tol

    @staticmethod
    def _progressed(res, res_prev, res0):
        """Residual guard of the stop test: a step only counts as converged
        when it still reduced the FP64 residual (by at least half) or the
        residual is at the FP64 floor (1e-12 of the initial one).  A step
        that stagnates (a small update because the iteration cannot reduce
        the residual) makes a small change c_k and a small ratio rho, which
        the change-based test alone would accept."""
        if res0 <= 0.0 or res <= 1e-12 * res0:
            return True
        return res_prev is not None and res <= 0.5 * res_prev

    def _refine_richardson(self, x_face, uc, fb=None):
        """x <- x + M^{-1}(b - A x) until the face correction is small."""
        Ic = self._Ic()
        prev = None
        res0 = res_prev = None
        done = False
        for it in range(self.refine_maxit):
            rI, rF = self.full_residual(x_face, uc, fb=fb)
            dI, dC = self._correction(rI, rF)
            x_face[Ic] += dC
            uc += dI
            v = torch.stack((torch.linalg.norm(dC), torch.linalg.norm(x_face),
                             torch.sqrt(torch.sum(rI * rI) + torch.sum(rF * rF)))).cpu().tolist()
            rel_corr, res = v[0] / v[1], v[2]
            res0 = res if res0 is None else res0
            self.refine_history.append(rel_corr)
            self.refine_residuals.append(res / res0 if res0 > 0 else 0.0)
            # the residual is that of the iterate before this correction: a
            # step is judged on the progress of the step before it too
            if self._converged(rel_corr, prev) and self._progressed(res, res_prev, res0):
                done = True
                break
            prev = rel_corr if it >= 1 else None
            res_prev = res
        return done

    def _graph_step(self, nc):
        """The FGMRES operator step (rI, rF) -> (zI, zC, wI, wF) = (M^{-1} r,
        A M^{-1} r) captured once per column count as a HIP graph: one launch
        instead of a few hundred small ones per Krylov step.  Returns the
        graph and its static buffers."""
        g = self._gsteps.get(nc)
        if g is not None:
            return g
        hps, ops = self.hps, self.hps.ops
        rI = torch.zeros(hps.nboxes, ops.ni, nc, dtype=torch.float64, device=self.device)
        rF = torch.zeros(self._Ic().shape[0], nc, dtype=torch.float64, device=self.device)
        cur = torch.cuda.current_stream(self.device)
        side = torch.cuda.Stream(self.device)
        side.wait_stream(cur)
        with torch.cuda.stream(side):                   # warm-up outside the capture
            for _ in range(2):
                zI, zC = self._correction(rI, rF)
                self._apply_full(zI, zC)
        cur.wait_stream(side)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            zI, zC = self._correction(rI, rF)
            wI, wF = self._apply_full(zI, zC)
        g = (graph, rI, rF, zI, zC, wI, wF)
        self._gsteps[nc] = g
        return g

    def _refine_fgmres(self, x_face, uc, fb=None):
        """Flexible GMRES on the full system, right-preconditioned by the FP32
        two-level solve M: the Arnoldi relation A Z_j = V_{j+1} H_j holds in
        FP64 for the stored preconditioned directions Z_j = M^{-1} V_j, so the
        FP64 residual ||b - A x|| is minimised over x0 + span(Z_j) even though
        M is applied in FP32.  Stops when the change of the face solution
        between consecutive iterates is below refine_tol relative to the face
        solution (the quantity plain refinement tests, with the same error
        estimate), restarting every refine_restart steps."""
        hps, ops = self.hps, self.hps.ops
        dev = self.device
        Ic = self._Ic()
        nb, ni = hps.nboxes, ops.ni
        nI = nb * ni
        nc = x_face.shape[-1]
        ntot = nI + Ic.shape[0]
        # Krylov storage (V and Z: 2 * kmax + 1 vectors of ntot x nc values):
        # the restart length is capped by half the free device memory (free
        # on the device plus what the caching allocator holds unused)
        free = torch.cuda.mem_get_info(dev)[0] + torch.cuda.memory_reserved(dev) - torch.cuda.memory_allocated(dev)
        kmax = max(2, min(self.refine_restart, (free // 2) // max(1, 16 * ntot * nc)))
        self.refine_kmax = kmax

        def pack(aI, aC, out):
            out.view(nc, ntot)[:, :nI] = aI.reshape(nI, nc).T
            out.view(nc, ntot)[:, nI:] = aC.T

        def unpack(v):
            # fresh row-major tensors (a (..., 1) view of v.T is "contiguous"
            # with a non-unit last stride)
            aI = torch.empty(nb, ni, nc, dtype=torch.float64, device=dev)
            aC = torch.empty(ntot - nI, nc, dtype=torch.float64, device=dev)
            aI.view(nI, nc).copy_(v[:, :nI].T)
            aC.copy_(v[:, nI:].T)
            return aI, aC

        use_graph = self.refine_graphs and (self._solves_since_build > 0 or nc in self._gsteps)
        gstep = self._graph_step(nc) if use_graph else None
        steps = 0
        done = False
        prev = None
        res0 = None           # per column: norm of the first residual
        res_prev = None       # largest relative residual estimate of the step before
        while steps < self.refine_maxit and not done:
            rI, rF = self.full_residual(x_face, uc, fb=fb)
            V = torch.empty(nc, kmax + 1, ntot, dtype=torch.float64, device=dev)
            Z = torch.empty(nc, kmax, ntot, dtype=torch.float64, device=dev)
            pack(rI, rF, V[:, 0])
            beta = torch.linalg.norm(V[:, 0], dim=1)
            V[:, 0] /= beta.clamp_min(1e-300)[:, None]
            xc0 = x_face[Ic].T.contiguous()                       # (nc, |I_Ctot|)
            # one transfer: beta and |x_face|^2
            bx = torch.cat((beta, (torch.linalg.norm(x_face) ** 2).reshape(1))).cpu().numpy()
            beta, xf2 = bx[:nc], float(bx[nc])
            if res0 is None:
                res0 = np.where(beta > 0, beta, 1.0)
            H = np.zeros((nc, kmax + 1, kmax))
            # face parts of the directions: Gram matrix G and g = Z_face^T xc0,
            # so the face change and the face norm of every iterate follow on
            # the host from y (one device-to-host transfer per step):
            #   |sum dy_i Z_i|^2 = dy^T G dy,
            #   |x_face + sum y_i Z_i|^2 = |x_face|^2 + 2 y^T g + y^T G y
            G = np.zeros((nc, kmax, kmax))
            gx = np.zeros((nc, kmax))
            y_prev = np.zeros((nc, 0))
            for j in range(kmax):
                if gstep is not None:
                    graph, gI, gF, zI, zC, wI, wF = gstep
                    gI.view(nI, nc).copy_(V[:, j, :nI].T)
                    gF.copy_(V[:, j, nI:].T)
                    graph.replay()
                else:
                    zI, zC = self._correction(*unpack(V[:, j]))
                    wI, wF = self._apply_full(zI, zC)
                pack(zI, zC, Z[:, j])
                w = torch.empty(nc, ntot, dtype=torch.float64, device=dev)
                pack(wI, wF, w)
                h = torch.zeros(nc, j + 2, dtype=torch.float64, device=dev)
                for _ in range(2):                                  # classical Gram-Schmidt, twice
                    # projections of every column's w on its basis, one
                    # batched launch each (slab_mdot / slab_maxpy)
                    c = hip.mdot(V[:, :j + 1], w)
                    hip.maxpy(V[:, :j + 1], c, w, alpha=-1.0)
                    h[:, :j + 1] += c
                h[:, j + 1] = hip.mdot(w[:, None, :], w)[:, 0].sqrt()
                V[:, j + 1] = w / h[:, j + 1].clamp_min(1e-300)[:, None]
                gr = torch.empty(nc, j + 2, dtype=torch.float64, device=dev)
                zf = Z[:, j, nI:]
                hip.mdot(Z[:, :j + 1, nI:], zf, out=gr[:, :j + 1])
                gr[:, j + 1] = hip.mdot(xc0[:, None, :], zf)[:, 0]
                hg = torch.cat((h, gr), 1).cpu().numpy()
                H[:, :j + 2, j] = hg[:, :j + 2]
                G[:, :j + 1, j] = hg[:, j + 2:2 * j + 3]
                G[:, j, :j + 1] = hg[:, j + 2:2 * j + 3]
                gx[:, j] = hg[:, 2 * j + 3]
                # min ||beta e1 - H y|| for every column at once (batched QR of
                # the (j+2) x (j+1) Hessenberg matrices); the minimum is the
                # FP64 residual norm ||b - A x_j|| of the iterate (right
                # preconditioning), a free residual estimate
                Hj = H[:, :j + 2, :j + 1]
                Q, R = np.linalg.qr(Hj)
                qb = Q[:, 0, :] * beta[:, None]                       # Q^T (beta e1)
                y = np.linalg.solve(R, qb[:, :, None])[:, :, 0]
                rhs = np.zeros((nc, j + 2))
                rhs[:, 0] = beta
                rres = np.linalg.norm(rhs - np.einsum("cij,cj->ci", Hj, y), axis=1)
                res_rel = float(np.max(rres / res0))
                dy = y - np.pad(y_prev, ((0, 0), (0, 1)))
                Gj = G[:, :j + 1, :j + 1]
                num2 = float(np.einsum("ci,cij,cj->", dy, Gj, dy))
                xn2 = xf2 + float(np.einsum("ci,cij,cj->", y, Gj, y)) + 2.0 * float((y * gx[:, :j + 1]).sum())
                rel = float(np.sqrt(max(num2, 0.0)) / np.sqrt(max(xn2, 1e-300)))
                self.refine_history.append(rel)
                self.refine_residuals.append(res_rel)
                y_prev = y
                steps += 1
                done = self._converged(rel, prev) and self._progressed(res_rel, res_prev, 1.0)
                res_prev = res_rel
                # the first step moves from the starting guess, not by a
                # contraction: rates start from the second step
                prev = rel if steps >= 2 else None
                if done or steps >= self.refine_maxit:
                    break
            yd = torch.from_numpy(np.ascontiguousarray(y_prev)).to(dev)
            upd = torch.zeros(nc, ntot, dtype=torch.float64, device=dev)
            if y_prev.shape[1] > 0:
                hip.maxpy(Z[:, :y_prev.shape[1]], yd, upd)
            dI, dC = unpack(upd)
            x_face[Ic] += dC
            uc += dI
            del V, Z
        return done

    def solve_refined(self, uu_dir_func, ff_body_func=None):
        """Mixed precision: FP32 factors, FP64 residuals of the full system,
        flexible GMRES (or plain refinement) until the face solution changes
        by at most refine_tol (relative)."""
        hps, ops = self.hps, self.hps.ops
        dev = self.device
        lf = hps.leaf
        torch.cuda.synchronize()
        tic = time.time()
        x_face, uu_dir = self.get_rhs(uu_dir_func)
        x_lift = x_face.clone()
        nc = uu_dir.shape[1]
        uc = torch.zeros(hps.nboxes, ops.ni, nc, dtype=torch.float64, device=dev)
        # body load f at the leaf interior nodes (FP64), part of the
        # un-condensed residual r_I = f - A_loc u
        fb = None if ff_body_func is None else hps.body_load(ff_body_func, nc, torch.float64)
        self.refine_history = []
        self.refine_residuals = []
        if self.refine_method == "fgmres":
            done = self._refine_fgmres(x_face, uc, fb=fb)
            self._solves_since_build += 1
        else:
            done = self._refine_richardson(x_face, uc, fb=fb)
        self.refine_converged = bool(done)
        uu = torch.empty(hps.nboxes, ops.p * ops.p, nc, dtype=torch.float64, device=dev)
        rI, rF = self.full_residual(x_face, uc, uu=uu, fb=fb)
        torch.cuda.synchronize()
        toc = time.time() - tic
        resloc = torch.linalg.norm(rI).item()
        ff0 = -self._face_residual(x_lift, lf["DtN"]).double()
        if fb is not None:
            ff0 = ff0 + self._body_reduce(fb)[self._Ic()].double()
        rel = (torch.linalg.norm(rF) / torch.linalg.norm(ff0)).item()
        self.last_face = x_face
        self.solve_info = dict(refine_converged=self.refine_converged, refine_steps=len(self.refine_history),
                               refine_history=list(self.refine_history),
                               refine_residuals=list(self.refine_residuals), res_face=rel)
        if not done and self.on_unconverged != "ignore":
            msg = ("mixed-precision refinement stopped after %d steps without meeting refine_tol=%g "
                   "(last change %.3e, relative residual %.3e)"
                   % (len(self.refine_history), self.refine_tol,
                      self.refine_history[-1] if self.refine_history else float("nan"),
                      self.refine_residuals[-1] if self.refine_residuals else float("nan")))
            if self.on_unconverged == "raise":
                raise ArithmeticError(msg)
            warnings.warn(msg, RuntimeWarning, stacklevel=3)
        return uu.reshape(-1, nc), rel, toc, resloc

    # ------------------------------------------------------- self checks
    def test_slab_solve(self, nrhs=3, transpose=False):
        """||A_CC(slab) x - f|| / ||f|| (or with A_CC^T) for random f, all
        slabs at once (slablu_solver.py:945-959), using the matrix-free face
        operator."""
        Np, nC = self.Npan, len(self.I_C)
        f = torch.rand(Np, nC, nrhs, dtype=self.wdt, device=self.device)
        sol, _, _ = self._slab_solve(f, transpose)
        xf = torch.zeros(self.hps.nface, nrhs, dtype=torch.float64, device=self.device)
        IsC = self.I_slabC.to(self.device)
        xf[IsC] = sol.reshape(-1, nrhs).double()
        y = self.hps.apply_A(xf, self.hps.leaf["DtN"].double(), transpose)[IsC]
        return (torch.linalg.norm(y - f.reshape(-1, nrhs).double()) / torch.linalg.norm(f.double())).item()

    def test_Tred_solver(self, nrhs=3, transpose=False):
        """||T x - f|| / ||f|| for the interface solve of random f, T applied
        through the slab solves (apply_Tred, slablu_solver.py:689-695)."""
        nX = self.I_slabX.shape[0]
        f = torch.rand(nX, nrhs, dtype=torch.float64, device=self.device)
        x = self.Tred_solver.solve(f, transpose=transpose)
        y = self.apply_Tred(x, transpose=transpose)
        return (torch.linalg.norm(y - f) / torch.linalg.norm(f)).item()

    def test_solve_reduced(self, nrhs=3, transpose=False):
        """||A_CC x - f|| / ||f|| (or with A_CC^T) of the two-level solve for
        random f on all face unknowns I_Ctot."""
        Ic = self._Ic()
        f = torch.rand(Ic.shape[0], nrhs, dtype=torch.float64, device=self.device)
        x = self.solve_reduced(f.to(self.wdt), transpose=transpose).double()
        xf = torch.zeros(self.hps.nface, nrhs, dtype=torch.float64, device=self.device)
        xf[Ic] = x
        y = self.hps.apply_A(xf, self.hps.leaf["DtN"].double(), transpose)[Ic]
        return (torch.linalg.norm(y - f) / torch.linalg.norm(f)).item()

    def _apply_ACC(self, x, transpose=False):
        Ic = self._Ic()
        xf = torch.zeros(self.hps.nface, x.shape[1], dtype=torch.float64, device=self.device)
        xf[Ic] = x
        return self.hps.apply_A(xf, self.hps.leaf["DtN"].double(), transpose)[Ic]

    def norm_ACC(self, iters=30, seed=0):
        """||A_CC||_2 estimated by power iteration on A_CC^T A_CC (a lower
        bound that converges from below)."""
        g = torch.Generator(device=self.device).manual_seed(seed)
        v = torch.randn(self._Ic().shape[0], 1, dtype=torch.float64, device=self.device, generator=g)
        v /= torch.linalg.norm(v)
        s = 0.0
        for _ in range(iters):
            w = self._apply_ACC(self._apply_ACC(v), transpose=True)
            s = torch.linalg.norm(w).item() ** 0.5
            v = w / torch.linalg.norm(w)
        return s

    def backward_error(self, nrhs=3, transpose=False, normA=None):
        """Normwise backward error max_j ||A_CC x_j - f_j|| / (||A_CC|| ||x_j|| +
        ||f_j||) of the two-level solve for random f (the measure a stable
        direct solve bounds by a modest multiple of the unit round-off,
        whatever the conditioning), with the plain relative residuals and
        the amplification ||x_j|| / ||f_j||."""
        Ic = self._Ic()
        nA = self.norm_ACC() if normA is None else normA
        f = torch.rand(Ic.shape[0], nrhs, dtype=torch.float64, device=self.device)
        x = self.solve_reduced(f.to(self.wdt), transpose=transpose).double()
        r = torch.linalg.norm(self._apply_ACC(x, transpose) - f, dim=0)
        nx, nf = torch.linalg.norm(x, dim=0), torch.linalg.norm(f, dim=0)
        eta = (r / (nA * nx + nf)).max().item()
        return dict(backward_error=eta, residual=(r / nf).max().item(), amplification=(nx / nf).max().item(),
                    norm_A=nA)
