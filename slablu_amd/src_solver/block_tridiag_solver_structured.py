"""Interface block-tridiagonal solver (drop-in for
src_solver/block_tridiag_solver_structured.py of the reference).

The reference keeps the interface blocks HBS-compressed and densifies each
one before its LU (block_tridiag_solver_structured.py:34-76).  Here the
blocks are formed exactly and densely on the GPU, and the Schur recursion
S_0 = D_0, S_j = D_j - sub_{j-1} S_{j-1}^{-1} sup_{j-1} is run as one
(b = 1) block-tridiagonal chain in block-Thomas form with S_j^{-1} and
G_j = S_j^{-1} sup_j kept (slab_btri_inv_factor, FP64), so that the solve
is a sequence of matrix products.
"""
import torch

from .. import hip


def _dense(x, device):
    if x is None:
        return None
    if hasattr(x, "todense") and not torch.is_tensor(x):
        x = x.todense(device)
    return torch.as_tensor(x).to(device)


class HBS_BlockTridiagonal:
    def __init__(self, b, n, cyclic=False):
        assert b < n
        if cyclic:
            raise NotImplementedError("periodic (cyclic) interface systems are not provided")
        self.b, self.n, self.cyclic = b, n, cyclic

    def build_lowmem(self, hbs_diag0, hbs_diag1, hbs_sub, hbs_sup, device=None, dtype=torch.float64):
        """diag = diag0 + diag1 (b, n, n); sub, sup (b-1, n, n); dense
        tensors or objects with todense()."""
        device = torch.device("cuda") if device is None else torch.device(device)
        b, n = self.b, self.n
        diag = _dense(hbs_diag0, device).to(dtype)
        if hbs_diag1 is not None:
            diag = diag + _dense(hbs_diag1, device).to(dtype)
        self.diag = diag.reshape(1, b, n, n).contiguous()
        if b > 1:
            self.sub = _dense(hbs_sub, device).to(dtype).reshape(1, b - 1, n, n).contiguous()
            self.sup = _dense(hbs_sup, device).to(dtype).reshape(1, b - 1, n, n).contiguous()
        else:
            self.sub = torch.zeros(1, 1, n, n, device=device, dtype=dtype)
            self.sup = torch.zeros(1, 1, n, n, device=device, dtype=dtype)
        self.info = torch.zeros(1, dtype=torch.int32, device=device)
        hip.btri_inv_factor(self.diag, self.sub, self.sup, info=self.info)
        self.nbytes = (self.diag.numel() + self.sub.numel() + self.sup.numel()) * self.diag.element_size()
        self.build_mode = "lowmem"

    build = build_lowmem

    def solve(self, f):
        """f (b*n, nrhs) -> solution (b*n, nrhs); f is not modified."""
        nrhs = f.shape[-1]
        x = f.to(self.diag.device, self.diag.dtype).reshape(1, self.b * self.n, nrhs).contiguous().clone()
        hip.btri_inv_solve(self.diag, self.sub, self.sup, x)
        return x.reshape(self.b * self.n, nrhs)

    solve_lowmem = solve
