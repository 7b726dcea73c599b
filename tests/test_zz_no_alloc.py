"""slab_getri_noalloc (the inverse with the library's static per-device
scratch instead of a workspace) through the C ABI against slab_getri (the
same kernels with a caller workspace): bit-identical results on two calls
in a row, the second smaller, and a problem larger than the scratch
rejected before any launch."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _factored(b, n, dtype, seed):
    from slablu_amd import hip
    g = torch.Generator().manual_seed(seed)
    A = (torch.randn(b, n, n, generator=g, dtype=torch.float64) + 2 * torch.eye(n, dtype=torch.float64))
    A = A.to("cuda", dtype).contiguous()
    ipiv = torch.empty(b, n, dtype=torch.int32, device="cuda")
    hip.getrf(A, ipiv)
    return A, ipiv


def test_getri_noalloc_matches_getri():
    from slablu_amd import hip
    print("case getri_noalloc: FP64 1 x 1000, then FP32 3 x 300, then FP64 4 x 100", flush=True)
    for b, n, dtype, seed in ((1, 1000, torch.float64, 1), (3, 300, torch.float32, 2), (4, 100, torch.float64, 3)):
        LU, ipiv = _factored(b, n, dtype, seed)
        ref = hip.getri(LU, ipiv)
        X = torch.full_like(LU, float("nan"))
        torch.cuda.synchronize()
        before = torch.cuda.memory_allocated()
        hip.getri_noalloc(LU, ipiv, X)
        after = torch.cuda.memory_allocated()
        torch.cuda.synchronize()
        assert after == before
        assert torch.equal(X, ref), (b, n, dtype)
        # and it is the inverse
        I = torch.eye(n, dtype=torch.float64, device="cuda")
        res = (X.double() @ _reassemble(LU, ipiv).double() - I).abs().max().item()
        assert res < (1e-9 if dtype == torch.float64 else 1e-2), res


def _reassemble(LU, ipiv):
    """P^T L U from packed getrf factors (row interchanges undone in reverse)."""
    L = torch.tril(LU, -1) + torch.eye(LU.shape[1], dtype=LU.dtype, device=LU.device)
    U = torch.triu(LU)
    A = L @ U
    piv = ipiv.cpu().numpy() - 1
    for bb in range(A.shape[0]):
        for i in range(A.shape[1] - 1, -1, -1):
            p = int(piv[bb, i])
            if p != i:
                A[bb, [i, p]] = A[bb, [p, i]]
    return A


def test_getri_noalloc_rejects_oversize():
    from slablu_amd import hip
    LU, ipiv = _factored(2, 1000, torch.float64, 4)
    X = torch.zeros_like(LU)
    with pytest.raises(ValueError, match="static scratch"):
        hip.getri_noalloc(LU, ipiv, X)
    torch.cuda.synchronize()
    assert torch.count_nonzero(X) == 0
