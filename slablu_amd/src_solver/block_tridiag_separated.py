"""Block-tridiagonal chains eliminated piecewise (nested dissection of the
1-D chain) on the GPU: the rem-chain factorization of the slab stage with
fewer dependent steps (the factorization of block_tridiag_solver_dense.py:
152-177 and its sweeps, :179-222, in another elimination order).

A chain of m blocks is padded with identity blocks to m' = P d and cut into
P periods of d blocks: blocks 0..d-2 of each period form a piece, block d-1
is a separator.  Ordering all piece blocks before the separators,

    [ A_pp  E ] [x_p]   [f_p]      A_pp = blockdiag(pieces) (tridiagonal each)
    [ C   A_ss] [x_s] = [f_s],     E, C: the couplings of every separator
                                   to the last block of the piece before it
                                   and the first block of the piece after it,

the pieces of every chain are factored at once (twisted block LU, d-1
blocks), the spikes V = A_pp^{-1} E come from one multi-column sweep, and
the separators' Schur complement T = A_ss - C V is again block tridiagonal
(P blocks: each piece couples its two separators) and is factored the same
way.  Dependent block steps: about (d-1)/2 + P/2 instead of m/2.

Solve:      y = A_pp^{-1} f_p,  x_s = T^{-1}(f_s - C y),  x_p = y - V x_s.
Transposed: y = A_pp^{-T} f_p,  x_s = T^{-T}(f_s - E^T y),  x_p = A_pp^{-T}(f_p - C^T x_s).
"""
import math

import torch

from .. import hip


def separated_period(m, bs, dtype):
    """Period d (piece of d-1 blocks + a separator) for a chain of m blocks,
    or None when the chain is too short to gain from the split."""
    if m < 12:
        return None
    # one-launch chains run twisted (both ends towards the middle); wider
    # blocks run a few batched launches per block step, top-down
    fused = hip.btri_fused(dtype, bs)
    best = None
    for d in range(3, m // 2 + 1):
        P = -(-m // d)
        steps = ((d - 1) // 2 + 1 + P // 2 + 1) if fused else (d - 1) + P
        waste = P * d - m
        key = (steps + 0.25 * waste, waste)
        if best is None or key < best[0]:
            best = (key, d)
    return best[1]


class SeparatedBlockTridiagonal:
    def __init__(self, b, m, bs, period, inv_factors=False):
        self.b, self.m, self.bs, self.d = b, m, bs, period
        self.P = -(-m // period)
        self.mp = self.P * period
        self.inv_factors = inv_factors

    # -- factorization --------------------------------------------------------
    def build_sweep(self, diag, sub, sup):
        """diag (b, m, bs, bs), sub/sup (b, m-1, bs, bs): sub[i] is block
        (i+1, i), sup[i] block (i, i+1)."""
        b, m, bs, d, P, mp = self.b, self.m, self.bs, self.d, self.P, self.mp
        dev, dt = diag.device, diag.dtype
        l = d - 1
        self.dtype, self.device = dt, dev
        # pieces and separators factor twisted where their blocks fit the
        # one-launch chain kernels; inverse factors need the twisted form
        self.twisted = hip.btri_fused(dt, bs)
        if not self.twisted:
            self.inv_factors = False
        # padded chain: identity diagonal blocks, zero couplings (the last
        # coupling slot of every chain is always padding)
        if mp > m or not diag.is_contiguous():
            Dp = torch.empty(b, mp, bs, bs, dtype=dt, device=dev)
            Dp[:, :m] = diag
            if mp > m:
                Dp[:, m:] = torch.eye(bs, dtype=dt, device=dev)
        else:
            Dp = diag                                                # only read below
        Lp = torch.nn.functional.pad(sub, (0, 0, 0, 0, 0, mp - m + 1))   # Lp[i] = block (i+1, i)
        Up = torch.nn.functional.pad(sup, (0, 0, 0, 0, 0, mp - m + 1))   # Up[i] = block (i, i+1)
        nb = b * P
        Dv, Lv, Uv = (t.view(nb, d, bs, bs) for t in (Dp, Lp, Up))
        # pieces: blocks 0..l-1 of every period
        self.pd = Dv[:, :l].contiguous()
        self.pl = Lv[:, :l - 1].contiguous()
        self.pu = Uv[:, :l - 1].contiguous()
        self.pipiv = torch.zeros(nb, l, bs, dtype=torch.int32, device=dev)
        self.pinfo = torch.zeros(nb, dtype=torch.int32, device=dev)
        hip.btri_factor(self.pd, self.pl, self.pu, self.pipiv, info=self.pinfo, twisted=self.twisted)
        self.plinv = self.slinv = None
        if self.inv_factors:
            # inverse factors first: the spike sweep below then runs as block
            # products too
            self.plinv = torch.empty_like(self.pd)
            hip.btri_invert_diag(self.pd, self.pipiv, self.plinv)
        # couplings of the separator of period g (block l) to its pieces:
        # strided views of the padded couplings (one block per period)
        self.C_last = Lv[:, l - 1]                # (sep g, last block of piece g)
        self.E_last = Uv[:, l - 1]                # (last block of piece g, sep g)
        self.C_next = Uv[:, l]                    # (sep g, first block of piece g+1)
        self.E_next = Lv[:, l]                    # (first block of piece g+1, sep g)
        # spikes: V[g] = piece_g^{-1} [E_prev | E_last] (E_prev: the coupling of
        # the piece's first block to the separator before it).  Separators of
        # a chain never couple across chains: E_next of a chain's last period
        # is a padding zero, so the first piece of the next chain gets none
        V = torch.zeros(nb, l, bs, 2, bs, dtype=dt, device=dev)
        V[1:, 0, :, 0] = self.E_next[:-1]
        V[:, l - 1, :, 1] = self.E_last
        self.V = V.view(nb, l * bs, 2 * bs)
        self._piece_solve(self.V, False, 0, nb)
        V5 = self.V.view(nb, l, bs, 2, bs)
        # separator Schur complement, block tridiagonal over the P separators.
        # The products run flat over all b P periods: the one pairing a
        # chain's last period with the next chain's first is zero (that
        # C_next is padding, and the next chain's first piece has no E_prev
        # spike), and its slot is dropped
        T = Dv[:, l].clone()                                             # (nb, bs, bs)
        hip.gemm(self.C_last, V5[:, l - 1, :, 1], T, alpha=-1.0, beta=1.0)
        if P > 1:
            # diagonal: minus (sep g -> first block of piece g+1) V^prev of piece g+1
            hip.gemm(self.C_next[:nb - 1], V5[1:, 0, :, 0], T[:nb - 1], alpha=-1.0, beta=1.0)
            tsup = torch.empty(b, P, bs, bs, dtype=dt, device=dev)
            tsub = torch.empty(b, P, bs, bs, dtype=dt, device=dev)
            # T(g, g+1) = -C_next[g] V^last_{g+1}[first block]
            hip.gemm(self.C_next[:nb - 1], V5[1:, 0, :, 1], tsup.view(nb, bs, bs)[:nb - 1], alpha=-1.0)
            # T(g+1, g) = -C_last[g+1] V^prev_{g+1}[last block]
            hip.gemm(self.C_last[1:], V5[1:, l - 1, :, 0], tsub.view(nb, bs, bs)[:nb - 1], alpha=-1.0)
            tsup, tsub = tsup[:, :P - 1], tsub[:, :P - 1]
        else:
            tsup = tsub = torch.zeros(b, 1, bs, bs, dtype=dt, device=dev)
        self.sd, self.sl, self.su = T.view(b, P, bs, bs), tsub, tsup
        self.sipiv = torch.zeros(b, P, bs, dtype=torch.int32, device=dev)
        self.sinfo = torch.zeros(b, dtype=torch.int32, device=dev)
        hip.btri_factor(self.sd, self.sl, self.su, self.sipiv, info=self.sinfo, twisted=self.twisted)
        if self.inv_factors:
            self.slinv = torch.empty_like(self.sd)
            hip.btri_invert_diag(self.sd, self.sipiv, self.slinv)
        self.info = torch.maximum(self.pinfo.view(b, P).amax(1), self.sinfo)
        self.nbytes = sum(t.numel() * t.element_size() for t in
                          (self.pd, self.pl, self.pu, self.sd, self.sl, self.su, self.V, self.C_last, self.E_last,
                           self.C_next, self.E_next)) + \
            (2 * (self.pd.numel() + self.sd.numel()) * self.pd.element_size() if self.inv_factors else 0)

    # -- sweeps ---------------------------------------------------------------
    def _piece_solve(self, x, transpose, g0, g1):
        """In place on x (periods g0..g1-1, l bs rows, nrhs; any batch stride)."""
        s = slice(g0, g1)
        if self.plinv is not None:
            hip.btri_solve(self.pd[s], self.pl[s], self.pu[s], None, x, transpose=transpose, twisted=self.twisted,
                           linv=self.plinv[s])
        else:
            hip.btri_solve(self.pd[s], self.pl[s], self.pu[s], self.pipiv[s], x, transpose=transpose,
                           twisted=self.twisted)

    def _sep_solve(self, x, transpose, b0, b1):
        s = slice(b0, b1)
        if self.slinv is not None:
            hip.btri_solve(self.sd[s], self.sl[s], self.su[s], None, x, transpose=transpose, twisted=self.twisted,
                           linv=self.slinv[s])
        else:
            hip.btri_solve(self.sd[s], self.sl[s], self.su[s], self.sipiv[s], x, transpose=transpose,
                           twisted=self.twisted)

    def lusolve_sweep(self, f, b_start=None, b_end=None, transpose=False):
        """Solve in place on f (chains b_start..b_end-1, m*bs rows, nrhs)."""
        b0 = 0 if b_start is None else b_start
        b1 = self.b if b_end is None else b_end
        m, bs, d, P, mp = self.m, self.bs, self.d, self.P, self.mp
        b = b1 - b0
        g0, g1 = b0 * P, b1 * P
        l = d - 1
        fv = f.reshape(b, m * bs, -1)
        nc = fv.shape[-1]
        nb = b * P
        tr = lambda X: X.transpose(-1, -2)
        W = torch.empty(b, mp * bs, nc, dtype=fv.dtype, device=fv.device)
        W[:, :m * bs] = fv
        if mp > m:
            W[:, m * bs:] = 0
        W4 = W.view(nb, d, bs, nc)
        xp = W.view(nb, d * bs, nc)[:, :l * bs]           # piece rows (strided batch)
        C_last, C_next = self.C_last[g0:g1], self.C_next[g0:g1]
        E_last, E_next = self.E_last[g0:g1], self.E_next[g0:g1]
        Cl, Cn, El, En = (tr(E_last), tr(E_next), C_last, C_next) if transpose else (C_last, C_next, E_last, E_next)
        # y = A_pp^{-1} f_p (or A_pp^{-T} f_p)
        self._piece_solve(xp, transpose, g0, g1)
        # g = f_s - (C or E^T) y: separator g sees the last block of piece g
        # and the first block of piece g+1
        g = W4[:, l].contiguous()                               # (nb, bs, nc)
        hip.gemm(Cl, W4[:, l - 1], g, alpha=-1.0, beta=1.0)
        if nb > 1:
            hip.gemm(Cn[:nb - 1], W4[1:, 0], g[:nb - 1], alpha=-1.0, beta=1.0)
        # chains do not couple: the last separator of a chain has no next piece
        # (its C_next block is a padding zero), so nothing crosses chains
        xs = g.view(b, P * bs, nc)
        self._sep_solve(xs, transpose, b0, b1)
        xs4 = xs.view(nb, bs, nc)
        if not transpose:
            # x_p = y - V^prev x_s(g-1) - V^last x_s(g); V^prev of a chain's
            # first piece is zero (no separator before it), so the product
            # with the previous chain's last separator adds nothing
            Vg = self.V[g0:g1]
            hip.gemm(Vg[:, :, bs:], xs4, xp, alpha=-1.0, beta=1.0)
            if nb > 1:
                hip.gemm(Vg[1:, :, :bs], xs4[:-1], xp[1:], alpha=-1.0, beta=1.0)
        else:
            # x_p = A_pp^{-T} (f_p - C^T x_s): piece g's last block gets
            # C_last[g]^T x_s(g), its first block C_next[g-1]^T x_s(g-1) (zero
            # for the first piece of a chain: the C_next of a chain's last
            # separator is a padding block)
            W[:, :m * bs] = fv
            if mp > m:
                W[:, m * bs:] = 0
            hip.gemm(tr(El), xs4, W4[:, l - 1], alpha=-1.0, beta=1.0)
            if nb > 1:
                hip.gemm(tr(En[:nb - 1]), xs4[:-1], W4[1:, 0], alpha=-1.0, beta=1.0)
            self._piece_solve(xp, True, g0, g1)
        W4[:, l] = xs4
        fv.copy_(W[:, :m * bs])
        return fv
