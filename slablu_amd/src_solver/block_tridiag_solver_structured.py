"""Interface block-tridiagonal solver (drop-in for
src_solver/block_tridiag_solver_structured.py of the reference).

The reference keeps the interface blocks HBS-compressed and densifies each
one before its LU (block_tridiag_solver_structured.py:34-76), running the
Schur recursion S_0 = D_0, S_j = D_j - sub_{j-1} S_{j-1}^{-1} sup_{j-1} from
the first block to the last.  Here the blocks are formed exactly and densely
on the GPU and eliminated from both ends at once (a twisted block
factorization): the top chain runs the same recursion over blocks 0..k-1,
the bottom chain runs its mirror image over blocks m-1..k+1,

    T_{m-1} = D_{m-1},  T_i = D_i - sup_i T_{i+1}^{-1} sub_i,

and the two meet at the middle block k = m // 2,

    M = D_k - sub_{k-1} S_{k-1}^{-1} sup_{k-1} - sup_k T_{k+1}^{-1} sub_k.

Each step factors one block of each chain as a batch of two, so the number
of dependent dense LU factorizations drops from m to m // 2 + 1 for the same
block products.  Each block is kept in inverse form (S^{-1} and
G = S^{-1} * coupling towards the middle, FP64), so a solve is a sequence of
matrix products: forward sweeps down both chains, the middle block, then
back-substitution out to both ends.

In the chain order [top chain, bottom chain, middle] the factorization is
T = LL UU with LL lower block-bidiagonal (diagonal S_t, sub-diagonal the
couplings Cl_t, middle row [sub_{k-1}, sup_k, M]) and UU unit upper
(super-diagonal G_t).  The transposed solve (the interface half of the
transposed two-level solve, slablu_solver.py:963-1025 for the operator)
runs T^T = UU^T LL^T: a forward sweep with the G_t^T, then a backward
sweep with the S_t^{-T} and Cl_t^T.

The other order offered is nested dissection: the even blocks are mutually
decoupled and inverted in one batched round, and the odd blocks' Schur
complement is block tridiagonal again (half as many blocks), solved by this
class recursively.  The dense LUs are latency-bound (every panel column
waits on the one before; a batch of 2 or of 5 matrices takes the same
time), so what counts is the number of dependent rounds: 4 instead of 5 for
9 blocks, 4 instead of 6 for 11.  It inverts the diagonal blocks of the
original system, which can be far worse conditioned than the Schur
complements the twisted order inverts, so it serves the mixed-precision
preconditioner, not the FP64 solve.
"""
import torch

from .. import hip


def _dense(x, device):
    if x is None:
        return None
    if hasattr(x, "todense") and not torch.is_tensor(x):
        x = x.todense(device)
    return torch.as_tensor(x).to(device)


def _inverse(S):
    """S (b, n, n) -> S^{-1} (b, n, n) = U^{-1} L^{-1} P from getrf (S is
    overwritten by its factors); getri inverts both triangles by recursive
    doubling, so no step runs down the diagonal block by block."""
    b, n, _ = S.shape
    ipiv = torch.empty(b, n, dtype=torch.int32, device=S.device)
    info = torch.zeros(b, dtype=torch.int32, device=S.device)
    hip.getrf(S, ipiv, info=info)
    return hip.getri(S, ipiv), info


class HBS_BlockTridiagonal:
    def __init__(self, b, n, cyclic=False, order="twisted"):
        """order: "twisted" (both ends towards the middle), "separated" (the
        even blocks first, then the odd blocks' Schur complement) or "auto"
        (separated from 6 blocks on, where it has fewer dependent steps).
        The separated order inverts the even diagonal blocks themselves,
        which can be much worse conditioned than the twisted order's Schur
        complements (on a Helmholtz interface at 10 points per wavelength:
        condition numbers up to 4e6 against 7e4, backward error about 1e-13
        against 1e-16), so it is meant for a preconditioner."""
        assert b < n
        if cyclic and b < 2:
            raise ValueError("a cyclic block-tridiagonal system needs at least 2 blocks")
        if order not in ("auto", "twisted", "separated"):
            raise ValueError("order must be 'auto', 'twisted' or 'separated'")
        self.b, self.n, self.cyclic, self.order = b, n, cyclic, order

    @staticmethod
    def dependent_steps(m, order="auto"):
        """Dependent rounds of dense LU + inverse for m blocks (acyclic):
        twisted m // 2 + 1; separated 1 + the steps of the odd blocks'
        system of m // 2 blocks."""
        if m <= 0:
            return 0
        if HBS_BlockTridiagonal._separated(m, order):
            return 1 + HBS_BlockTridiagonal.dependent_steps(m // 2, "auto" if order == "auto" else "twisted")
        return m // 2 + 1

    @staticmethod
    def _separated(m, order):
        return (order == "separated" and m >= 2) or (order == "auto" and m >= 6)

    def build_lowmem(self, hbs_diag0, hbs_diag1, hbs_sub, hbs_sup, device=None, dtype=torch.float64):
        """diag = diag0 + diag1 (b, n, n); sub, sup (b-1, n, n), with sub[j]
        the block (j+1, j) and sup[j] the block (j, j+1); dense tensors or
        objects with todense().  Cyclic: sub, sup (b, n, n), sub[b-1] the
        corner block (0, b-1) and sup[b-1] the corner block (b-1, 0)."""
        if self.cyclic:
            return self._build_cyclic(hbs_diag0, hbs_diag1, hbs_sub, hbs_sup, device, dtype)
        device = torch.device("cuda") if device is None else torch.device(device)
        m, n = self.b, self.n
        diag = _dense(hbs_diag0, device).to(dtype)
        if hbs_diag1 is not None:
            diag = diag + _dense(hbs_diag1, device).to(dtype)
        diag = diag.reshape(m, n, n)
        if m > 1:
            sub = _dense(hbs_sub, device).to(dtype).reshape(m - 1, n, n)
            sup = _dense(hbs_sup, device).to(dtype).reshape(m - 1, n, n)
        self.dtype, self.device = dtype, device
        self.separated = self._separated(m, self.order)
        if self.separated:
            return self._build_separated(diag, sub, sup)
        k = m // 2
        L = k                                    # chain length (bottom padded)
        self.k, self.L = k, L
        self.info = torch.zeros(2, dtype=torch.int32, device=device)
        if L > 0:
            pad = L - (m - 1 - k)                # identity blocks in front of the bottom chain
            Dc = torch.zeros(2, L, n, n, dtype=dtype, device=device)
            Cl = torch.zeros(2, L, n, n, dtype=dtype, device=device)   # coupling to the previous chain block
            Cr = torch.zeros(2, L, n, n, dtype=dtype, device=device)   # coupling to the next block (last: middle)
            for t in range(L):
                Dc[0, t] = diag[t]
                if t > 0:
                    Cl[0, t] = sub[t - 1]
                Cr[0, t] = sup[t]
                if t < pad:
                    Dc[1, t] = torch.eye(n, dtype=dtype, device=device)
                else:
                    i = m - 1 - (t - pad)
                    Dc[1, t] = diag[i]
                    if t > pad:
                        Cl[1, t] = sup[i]
                    Cr[1, t] = sub[i - 1]
            Sinv = torch.empty_like(Dc)
            G = torch.empty_like(Dc)
            for t in range(L):
                S = Dc[:, t].clone()
                if t > 0:
                    hip.gemm(Cl[:, t], G[:, t - 1], S, alpha=-1.0, beta=1.0)
                Sinv[:, t], info = _inverse(S)
                self.info = torch.maximum(self.info, info)
                hip.gemm(Sinv[:, t], Cr[:, t], G[:, t])
            M = diag[k].clone().unsqueeze(0)
            hip.gemm(sub[k - 1], G[0, L - 1].unsqueeze(0), M, alpha=-1.0, beta=1.0)
            self.has_bot = m - 1 - k > 0
            if self.has_bot:
                hip.gemm(sup[k], G[1, L - 1].unsqueeze(0), M, alpha=-1.0, beta=1.0)
            self.Sinv, self.G, self.Cl = Sinv, G, Cl
            self.sub_mid = sub[k - 1].contiguous()
            self.sup_mid = sup[k].contiguous() if self.has_bot else None
            self.pad = pad
            del Dc, Cr
        else:
            M = diag[k].clone().unsqueeze(0)
        self.Minv, info = _inverse(M)
        self.info[0] = torch.maximum(self.info[0], info[0])
        self.nbytes = sum(t.numel() * t.element_size()
                          for t in ([self.Minv] + ([self.Sinv, self.G, self.Cl] if L > 0 else [])))
        self.dtype, self.device = dtype, device
        self.build_mode = "lowmem"

    build = build_lowmem

    def _build_separated(self, diag, sub, sup):
        """Nested-dissection order: the even blocks E = {0, 2, ...} are
        decoupled from each other, so they are inverted in one batch; the odd
        blocks O = {1, 3, ...} are left with the Schur complement

            T = A_OO - A_OE A_EE^{-1} A_EO,

        again block tridiagonal over the m // 2 odd blocks (odd block q is
        block 2q + 1; its even neighbours are q and q + 1):

            T_qq      = D_{2q+1} - sub_{2q} GR_q - sup_{2q+1} GL_q,
            T_{q,q+1} = -sup_{2q+1} GR_{q+1},   T_{q+1,q} = -sub_{2q+2} GL_q,

        with GR_q = E_q^{-1} sup_{2q} (even q towards its right odd
        neighbour) and GL_q = E_{q+1}^{-1} sub_{2q+1} (even q + 1 towards its
        left one), and T is factored by this class again.  A path of 9
        blocks takes 4 dependent rounds of LU + inverse this way against 5
        twisted (the fewest any elimination order allows for it); the price
        is the products that form T."""
        m, n = self.b, self.n
        ne, no = (m + 1) // 2, m // 2
        self.ne, self.no = ne, no
        Einv, info_e = _inverse(diag[0::2].clone())
        GL = torch.empty(ne - 1, n, n, dtype=self.dtype, device=self.device)
        GR = torch.empty(no, n, n, dtype=self.dtype, device=self.device)
        if ne > 1:
            hip.gemm(Einv[1:], sub[1::2], GL)
        hip.gemm(Einv[:no], sup[0::2], GR)
        T = diag[1::2].clone()
        hip.gemm(sub[0::2], GR, T, alpha=-1.0, beta=1.0)
        if ne > 1:
            hip.gemm(sup[1::2], GL, T[:ne - 1], alpha=-1.0, beta=1.0)
        Tsub = Tsup = None
        if no > 1:
            Tsup = torch.empty(no - 1, n, n, dtype=self.dtype, device=self.device)
            Tsub = torch.empty(no - 1, n, n, dtype=self.dtype, device=self.device)
            hip.gemm(sup[1::2][:no - 1], GR[1:], Tsup, alpha=-1.0)
            hip.gemm(sub[2::2][:no - 1], GL[:no - 1], Tsub, alpha=-1.0)
        self.Tsolver = HBS_BlockTridiagonal(no, n, order="auto" if self.order == "auto" else "twisted")
        self.Tsolver.build_lowmem(T, None, Tsub, Tsup, device=self.device, dtype=self.dtype)
        del T, Tsub, Tsup
        self.Einv, self.GL, self.GR = Einv, GL, GR
        self.sub, self.sup = sub.contiguous(), sup.contiguous()
        self.info = torch.cat((info_e, self.Tsolver.info))
        self.nbytes = self.Tsolver.nbytes + sum(t.numel() * t.element_size()
                                                for t in (self.Einv, self.GL, self.GR, self.sub, self.sup))
        self.build_mode = "lowmem"

    def _solve_separated(self, f, transpose=False):
        """x_O = T^{-1}(f_O - A_OE A_EE^{-1} f_E), x_E = A_EE^{-1}(f_E -
        A_EO x_O), as products with E^{-1}, GL, GR and the couplings.
        Transposed: y = A_EE^{-T} f_E, x_O = T^{-T}(f_O - A_EO^T y),
        x_E = A_EE^{-T}(f_E - A_OE^T x_O)."""
        m, n, ne, no = self.b, self.n, self.ne, self.no
        nrhs = f.shape[-1]
        tr = lambda X: X.transpose(-1, -2)
        fb = f.to(self.device, self.dtype).reshape(m, n, nrhs)
        sub, sup = self.sub, self.sup
        x = torch.empty(m, n, nrhs, dtype=self.dtype, device=self.device)
        xe, xo = x[0::2], x[1::2]          # the even blocks hold y = E^{-1} f_E (E^{-T}) first
        g = fb[1::2].clone()
        if transpose:
            hip.gemm(tr(self.Einv), fb[0::2], xe)
            hip.gemm(tr(sup[0::2]), xe[:no], g, alpha=-1.0, beta=1.0)
            if ne > 1:
                hip.gemm(tr(sub[1::2]), xe[1:], g[:ne - 1], alpha=-1.0, beta=1.0)
            xo.copy_(self.Tsolver.solve(g.reshape(no * n, nrhs), transpose=True).reshape(no, n, nrhs))
            h = fb[0::2].clone()
            if ne > 1:
                hip.gemm(tr(sup[1::2]), xo[:ne - 1], h[1:], alpha=-1.0, beta=1.0)
            hip.gemm(tr(sub[0::2]), xo, h[:no], alpha=-1.0, beta=1.0)
            hip.gemm(tr(self.Einv), h, xe)
        else:
            hip.gemm(self.Einv, fb[0::2], xe)
            hip.gemm(sub[0::2], xe[:no], g, alpha=-1.0, beta=1.0)
            if ne > 1:
                hip.gemm(sup[1::2], xe[1:], g[:ne - 1], alpha=-1.0, beta=1.0)
            xo.copy_(self.Tsolver.solve(g.reshape(no * n, nrhs)).reshape(no, n, nrhs))
            if ne > 1:
                hip.gemm(self.GL, xo[:ne - 1], xe[1:], alpha=-1.0, beta=1.0)
            hip.gemm(self.GR, xo, xe[:no], alpha=-1.0, beta=1.0)
        return x.reshape(m * n, nrhs)

    def _build_cyclic(self, hbs_diag0, hbs_diag1, hbs_sub, hbs_sup, device, dtype):
        """Cyclic system (the periodic interface ring, block_tridiag_solver_
        structured.py:118-155) as an acyclic chain over blocks 0..b-2 with a
        border at block b-1:
            [T  B] [x ]   [f ]      B: blocks (0, b-1) and (b-2, b-1)
            [C  D] [xb] = [fb],     C: blocks (b-1, 0) and (b-1, b-2)
        W = T^{-1} B by the chain solver, then the Schur complement
        S = D - C W (FP64, explicit inverse)."""
        device = torch.device("cuda") if device is None else torch.device(device)
        m, n = self.b, self.n
        diag = _dense(hbs_diag0, device).to(dtype)
        if hbs_diag1 is not None:
            diag = diag + _dense(hbs_diag1, device).to(dtype)
        diag = diag.reshape(m, n, n)
        sub = _dense(hbs_sub, device).to(dtype).reshape(m, n, n)
        sup = _dense(hbs_sup, device).to(dtype).reshape(m, n, n)
        self.inner = HBS_BlockTridiagonal(m - 1, n, order=self.order)
        self.inner.build_lowmem(diag[:m - 1], None, sub[:m - 2] if m > 2 else None,
                                sup[:m - 2] if m > 2 else None, device=device, dtype=dtype)
        # border column B and row C (blocks may coincide when b = 2)
        Bcol = torch.zeros(m - 1, n, n, dtype=dtype, device=device)
        Bcol[0] += sub[m - 1]
        Bcol[m - 2] += sup[m - 2]
        self.C0 = sup[m - 1].contiguous()        # block (b-1, 0)
        self.Cl = sub[m - 2].contiguous()        # block (b-1, b-2)
        # the nonzero blocks of B (for the transposed solve)
        self.B0, self.Bl = Bcol[0].clone(), Bcol[m - 2].clone()
        self.Wt = None                           # T^{-T} C^T, formed at the first transposed solve
        W = self.inner.solve(Bcol.reshape((m - 1) * n, n)).reshape(m - 1, n, n)
        S = diag[m - 1].clone().unsqueeze(0)
        hip.gemm(self.C0, W[0].unsqueeze(0), S, alpha=-1.0, beta=1.0)
        hip.gemm(self.Cl, W[m - 2].unsqueeze(0), S, alpha=-1.0, beta=1.0)
        self.Sinv, info = _inverse(S)
        self.W = W.reshape((m - 1) * n, n).contiguous()
        self.info = torch.maximum(self.inner.info, torch.zeros_like(self.inner.info))
        self.info[0] = torch.maximum(self.info[0], info[0])
        self.nbytes = self.inner.nbytes + sum(t.numel() * t.element_size()
                                              for t in (self.Sinv, self.W, self.C0, self.Cl))
        self.dtype, self.device = dtype, device
        self.build_mode = "lowmem"

    def _solve_cyclic(self, f, transpose=False):
        """[T B; C D] [x; xb] = [f; fb]: y = T^{-1} f, xb = S^{-1}(fb - C y),
        x = y - W xb.  Transposed: y = T^{-T} f, xb = S^{-T}(fb - B^T y),
        x = y - T^{-T} C^T xb."""
        m, n = self.b, self.n
        nrhs = f.shape[-1]
        tr = lambda X: X.transpose(-1, -2)
        fb = f.to(self.device, self.dtype).reshape(m * n, nrhs)
        y = self.inner.solve(fb[:(m - 1) * n], transpose=transpose).reshape(m - 1, n, nrhs)
        r = fb[(m - 1) * n:].clone().unsqueeze(0)
        if transpose:
            hip.gemm(tr(self.B0), y[0].unsqueeze(0), r, alpha=-1.0, beta=1.0)
            if m > 2:
                hip.gemm(tr(self.Bl), y[m - 2].unsqueeze(0), r, alpha=-1.0, beta=1.0)
            if self.Wt is None:
                Ct = torch.zeros(m - 1, n, n, dtype=self.dtype, device=self.device)
                Ct[0] += tr(self.C0)
                Ct[m - 2] += tr(self.Cl)
                self.Wt = self.inner.solve(Ct.reshape((m - 1) * n, n), transpose=True).contiguous()
            Sm, Wm = tr(self.Sinv), self.Wt
        else:
            hip.gemm(self.C0, y[0].unsqueeze(0), r, alpha=-1.0, beta=1.0)
            hip.gemm(self.Cl, y[m - 2].unsqueeze(0), r, alpha=-1.0, beta=1.0)
            Sm, Wm = self.Sinv, self.W
        xb = torch.empty(1, n, nrhs, dtype=self.dtype, device=self.device)
        hip.gemm(Sm, r, xb)
        x = torch.empty(m * n, nrhs, dtype=self.dtype, device=self.device)
        xs = y.reshape((m - 1) * n, nrhs).clone().unsqueeze(0)
        hip.gemm(Wm.unsqueeze(0), xb, xs, alpha=-1.0, beta=1.0)
        x[:(m - 1) * n] = xs[0]
        x[(m - 1) * n:] = xb[0]
        return x

    def _solve_transposed(self, f):
        """T^T x = f with the factors of the twisted elimination: forward
        z_t = f_t - G_{t-1}^T z_{t-1} down both chains, z_k = f_k - G^T z of
        both chain ends, x_k = M^{-T} z_k, then backward x_t = S_t^{-T}(z_t -
        Cl_{t+1}^T x_{t+1}) out to both ends (the middle couplings sub_{k-1},
        sup_k in place of Cl_L)."""
        m, n, k, L = self.b, self.n, self.k, self.L
        nrhs = f.shape[-1]
        tr = lambda X: X.transpose(-1, -2)
        fb = f.to(self.device, self.dtype).reshape(m, n, nrhs)
        x = torch.empty(m, n, nrhs, dtype=self.dtype, device=self.device)
        z_mid = fb[k].clone().unsqueeze(0)
        if L > 0:
            pad = self.pad
            bot = [m - 1 - (t - pad) if t >= pad else None for t in range(L)]
            Z = torch.zeros(2, L, n, nrhs, dtype=self.dtype, device=self.device)
            for t in range(L):
                Z[0, t] = fb[t]
                if bot[t] is not None:
                    Z[1, t] = fb[bot[t]]
                if t > 0:
                    hip.gemm(tr(self.G[:, t - 1]), Z[:, t - 1], Z[:, t], alpha=-1.0, beta=1.0)
            hip.gemm(tr(self.G[0, L - 1]).unsqueeze(0), Z[0, L - 1].unsqueeze(0), z_mid, alpha=-1.0, beta=1.0)
            if self.has_bot:
                hip.gemm(tr(self.G[1, L - 1]).unsqueeze(0), Z[1, L - 1].unsqueeze(0), z_mid, alpha=-1.0, beta=1.0)
        xk = torch.empty(1, n, nrhs, dtype=self.dtype, device=self.device)
        hip.gemm(tr(self.Minv), z_mid, xk)
        x[k] = xk[0]
        if L > 0:
            X = torch.empty(2, L, n, nrhs, dtype=self.dtype, device=self.device)
            W = torch.empty(2, n, nrhs, dtype=self.dtype, device=self.device)
            for t in range(L - 1, -1, -1):
                W.copy_(Z[:, t])
                if t == L - 1:
                    hip.gemm(tr(self.sub_mid).unsqueeze(0), xk, W[0:1], alpha=-1.0, beta=1.0)
                    if self.has_bot:
                        hip.gemm(tr(self.sup_mid).unsqueeze(0), xk, W[1:2], alpha=-1.0, beta=1.0)
                else:
                    hip.gemm(tr(self.Cl[:, t + 1]), X[:, t + 1], W, alpha=-1.0, beta=1.0)
                hip.gemm(tr(self.Sinv[:, t]), W, X[:, t])
            x[:L] = X[0]
            for t in range(pad, L):
                x[m - 1 - (t - pad)] = X[1, t]
        return x.reshape(m * n, nrhs)

    def solve(self, f, transpose=False):
        """f (b*n, nrhs) -> solution (b*n, nrhs) of T x = f, or T^T x = f
        with transpose; f is not modified."""
        if self.cyclic:
            return self._solve_cyclic(f, transpose)
        if self.separated:
            return self._solve_separated(f, transpose)
        if transpose:
            return self._solve_transposed(f)
        m, n, k, L = self.b, self.n, self.k, self.L
        nrhs = f.shape[-1]
        fb = f.to(self.device, self.dtype).reshape(m, n, nrhs)
        x = torch.empty(m, n, nrhs, dtype=self.dtype, device=self.device)
        rhs_mid = fb[k].clone().unsqueeze(0)
        if L > 0:
            pad = self.pad
            Y = torch.zeros(2, L, n, nrhs, dtype=self.dtype, device=self.device)
            top = list(range(L))
            bot = [m - 1 - (t - pad) if t >= pad else None for t in range(L)]
            W = torch.empty(2, n, nrhs, dtype=self.dtype, device=self.device)
            for t in range(L):
                W[0] = fb[top[t]]
                if bot[t] is None:
                    W[1] = 0
                else:
                    W[1] = fb[bot[t]]
                if t > 0:
                    hip.gemm(self.Cl[:, t], Y[:, t - 1], W, alpha=-1.0, beta=1.0)
                hip.gemm(self.Sinv[:, t], W, Y[:, t])
            hip.gemm(self.sub_mid, Y[0, L - 1].unsqueeze(0), rhs_mid, alpha=-1.0, beta=1.0)
            if self.has_bot:
                hip.gemm(self.sup_mid, Y[1, L - 1].unsqueeze(0), rhs_mid, alpha=-1.0, beta=1.0)
        xk = torch.empty(1, n, nrhs, dtype=self.dtype, device=self.device)
        hip.gemm(self.Minv, rhs_mid, xk)
        x[k] = xk[0]
        if L > 0:
            X = torch.empty(2, L, n, nrhs, dtype=self.dtype, device=self.device)
            nxt = xk.expand(2, n, nrhs)
            for t in range(L - 1, -1, -1):
                X[:, t] = Y[:, t]
                hip.gemm(self.G[:, t], nxt, X[:, t], alpha=-1.0, beta=1.0)
                nxt = X[:, t]
            x[:L] = X[0]
            for t in range(pad, L):
                x[m - 1 - (t - pad)] = X[1, t]
        return x.reshape(m * n, nrhs)

    solve_lowmem = solve
