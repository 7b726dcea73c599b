"""Parity of the batched HIP kernels (C ABI through slablu_amd.hip) against
FP64 CPU references (torch / scipy on the host)."""
import numpy as np
import pytest
import scipy.linalg
import torch

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-12, torch.float32: 2e-5}


def _hip():
    from slablu_amd import hip
    return hip


def _rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return (torch.linalg.norm(a - b) / max(torch.linalg.norm(b), 1e-300)).item()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 33, 5), (64, 64, 16), (130, 70, 45), (200, 3, 97)])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_gemm(dtype, m, n, k, ta, tb):
    hip = _hip()
    g = torch.Generator().manual_seed(m * 1000 + n * 10 + k)
    batch = 3
    A0 = torch.randn(batch, k, m, generator=g, dtype=torch.float64) if ta else torch.randn(batch, m, k, generator=g, dtype=torch.float64)
    B0 = torch.randn(batch, n, k, generator=g, dtype=torch.float64) if tb else torch.randn(batch, k, n, generator=g, dtype=torch.float64)
    C0 = torch.randn(batch, m, n, generator=g, dtype=torch.float64)
    A = A0.to("cuda", dtype)
    B = B0.to("cuda", dtype)
    C = C0.to("cuda", dtype)
    Av = A.transpose(1, 2) if ta else A
    Bv = B.transpose(1, 2) if tb else B
    hip.gemm(Av, Bv, C, alpha=-0.5, beta=2.0)
    ref = -0.5 * (A0.transpose(1, 2) if ta else A0) @ (B0.transpose(1, 2) if tb else B0) + 2.0 * C0
    torch.cuda.synchronize()
    assert _rel(C, ref) < TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("k0", [0, 32, 96, 352])
def test_gemm_leaf_update_views(dtype, k0):
    """The leaf LU's trailing update on strided sub-views of a batch of
    400 x 480 matrices (16-byte vector loads whenever rows, offsets, k and n
    allow them; partial 64-row tiles at the bottom edge)."""
    hip = _hip()
    g = torch.Generator().manual_seed(k0 + 7)
    M0 = torch.randn(5, 400, 480, generator=g, dtype=torch.float64)
    M = M0.to("cuda", dtype)
    A, B, C = M[:, k0 + 32:, k0:k0 + 32], M[:, k0:k0 + 32, k0 + 32:], M[:, k0 + 32:, k0 + 32:]
    ref = M0[:, k0 + 32:, k0 + 32:] - 0.5 * M0[:, k0 + 32:, k0:k0 + 32] @ M0[:, k0:k0 + 32, k0 + 32:]
    hip.gemm(A, B, C, alpha=-0.5, beta=1.0)
    torch.cuda.synchronize()
    assert _rel(C, ref) < TOL[dtype]
    # the rest of the matrix is untouched
    assert torch.equal(M[:, :k0 + 32].cpu(), M0[:, :k0 + 32].to(dtype))
    assert torch.equal(M[:, k0 + 32:, :k0 + 32].cpu(), M0[:, k0 + 32:, :k0 + 32].to(dtype))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("m,n,k", [(100, 36, 32), (64, 4, 8), (7, 128, 4), (130, 68, 44)])
def test_gemm_vector_shapes(dtype, m, n, k):
    """Shapes whose k and n are multiples of the vector width, with partial
    tiles in every direction, against the FP64 product."""
    hip = _hip()
    g = torch.Generator().manual_seed(m + n + k)
    A0 = torch.randn(3, m, k, generator=g, dtype=torch.float64)
    B0 = torch.randn(3, k, n, generator=g, dtype=torch.float64)
    C0 = torch.randn(3, m, n, generator=g, dtype=torch.float64)
    C = C0.to("cuda", dtype)
    hip.gemm(A0.to("cuda", dtype), B0.to("cuda", dtype), C, alpha=1.5, beta=-1.0)
    torch.cuda.synchronize()
    assert _rel(C, 1.5 * A0 @ B0 - C0) < TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("m,k", [(5, 32), (400, 80), (1001, 200), (37, 256), (64, 257), (130, 1000)])
@pytest.mark.parametrize("n", [1, 3, 8])
def test_gemv_paths(dtype, m, k, n):
    """Products with at most 8 right-hand columns (the 16-lanes-per-row
    kernel for k <= 256, 64 lanes per row above), with A a column slice of
    a wider matrix as in the leaf recovery, against the FP64 product."""
    hip = _hip()
    g = torch.Generator().manual_seed(m * 7 + k + n)
    W0 = torch.randn(3, m, k + 40, generator=g, dtype=torch.float64)
    B0 = torch.randn(3, k, n, generator=g, dtype=torch.float64)
    C0 = torch.randn(3, m, n, generator=g, dtype=torch.float64)
    W = W0.to("cuda", dtype)
    C = C0.to("cuda", dtype)
    hip.gemm(W[:, :, 40:], B0.to("cuda", dtype), C, alpha=-1.0, beta=0.5)
    torch.cuda.synchronize()
    assert _rel(C, -W0[:, :, 40:] @ B0 + 0.5 * C0) < TOL[dtype]


def test_gemm_broadcast_and_zero_beta():
    hip = _hip()
    g = torch.Generator().manual_seed(1)
    A0 = torch.randn(40, 50, generator=g, dtype=torch.float64)
    B0 = torch.randn(5, 50, 30, generator=g, dtype=torch.float64)
    C = torch.full((5, 40, 30), float("nan"), dtype=torch.float64, device="cuda")
    hip.gemm(A0.cuda(), B0.cuda(), C, alpha=1.0, beta=0.0)
    torch.cuda.synchronize()
    assert _rel(C, A0 @ B0) < 1e-13


@pytest.mark.parametrize("m,n,k,beta", [(80, 80, 400, 1.0), (56, 56, 196, 1.0), (49, 80, 64, 0.0), (80, 64, 100, 0.5)])
def test_gemm_shared_operator_small_outputs(m, n, k, beta):
    """FP32 products of one shared operator with 70 matrices into small
    outputs (the leaf DtN product N_x S'), which run as one workgroup per
    matrix: against the FP64 product; C is not read when beta is 0."""
    hip = _hip()
    g = torch.Generator().manual_seed(m + n + k)
    A0 = torch.randn(m, k, generator=g, dtype=torch.float64)
    B0 = torch.randn(70, k, n, generator=g, dtype=torch.float64)
    C0 = torch.randn(70, m, n, generator=g, dtype=torch.float64)
    C = C0.to("cuda", torch.float32) if beta != 0.0 else torch.full((70, m, n), float("nan"), device="cuda")
    hip.gemm(A0.to("cuda", torch.float32), B0.to("cuda", torch.float32), C, alpha=-1.0, beta=beta)
    ref = -(A0.float().double() @ B0.float().double()) + beta * C0.float().double()
    torch.cuda.synchronize()
    assert torch.isfinite(C).all()
    assert _rel(C, ref) < TOL[torch.float32]


def _tri(b, m, lower, g, dtype):
    A = torch.randn(b, m, m, generator=g, dtype=torch.float64) * 0.1
    A = torch.tril(A) if lower else torch.triu(A)
    A = A + torch.diag_embed(1.0 + torch.rand(b, m, generator=g, dtype=torch.float64))
    return A


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("m,n", [(1, 1), (20, 7), (45, 130), (100, 33)])
@pytest.mark.parametrize("lower", [True, False])
@pytest.mark.parametrize("unit", [True, False])
@pytest.mark.parametrize("transA", [False, True])
def test_trsm(dtype, m, n, lower, unit, transA):
    hip = _hip()
    g = torch.Generator().manual_seed(m + 7 * n)
    b = 2
    # view lower/upper as given after the optional transpose
    A0 = _tri(b, m, lower != transA, g, dtype)
    Aview0 = A0.transpose(1, 2) if transA else A0
    B0 = torch.randn(b, m, n, generator=g, dtype=torch.float64)
    A = A0.to("cuda", dtype)
    Av = A.transpose(1, 2) if transA else A
    X = B0.to("cuda", dtype).clone()
    hip.trsm(Av, X, lower=lower, unit=unit)
    Aref = Aview0.clone()
    if unit:
        Aref = Aref - torch.diag_embed(torch.diagonal(Aref, dim1=1, dim2=2)) + torch.eye(m, dtype=torch.float64)
    ref = torch.linalg.solve_triangular(Aref, B0, upper=not lower)
    torch.cuda.synchronize()
    assert _rel(X, ref) < 50 * TOL[dtype]


@pytest.mark.parametrize("m,n", [(400, 80), (196, 56), (100, 96), (36, 52), (64, 80), (48, 49)])
def test_trsm_upper_many_matrices(m, n):
    """FP32 upper solves of 70 matrices (the leaf stage's S' = U^{-1} Y
    shape: U and the right-hand sides side by side in one row-major
    matrix), which run as one workgroup per matrix with the right-hand
    sides in LDS: against a float64 triangular solve, for m not a multiple
    of the 32-row block and n not a multiple of 16 (n = 49, not a multiple
    of 4, takes the blocked path)."""
    hip = _hip()
    g = torch.Generator().manual_seed(m + n)
    b = 70
    U0 = _tri(b, m, False, g, torch.float32)
    Y0 = torch.randn(b, m, n, generator=g, dtype=torch.float64)
    M = torch.cat((U0, Y0), dim=2).to("cuda", torch.float32).contiguous()
    hip.trsm(M[:, :, :m], M[:, :, m:], lower=False, unit=False)
    ref = torch.linalg.solve_triangular(M[:, :, :m].double().cpu(), Y0.float().double(), upper=True)
    torch.cuda.synchronize()
    assert _rel(M[:, :, m:].double().cpu(), ref) < 50 * TOL[torch.float32]
    # the triangle itself is untouched
    assert torch.equal(M[:, :, :m].cpu(), U0.float())


def _lu_check(Af, LUf, piv, n):
    """||P L U - A|| / ||A|| for one square matrix (numpy)."""
    L = np.tril(LUf, -1) + np.eye(n)
    U = np.triu(LUf)
    PA = Af.copy()
    for i in range(n):
        p = piv[i] - 1
        if p != i:
            PA[[i, p]] = PA[[p, i]]
    return np.linalg.norm(L @ U - PA) / np.linalg.norm(Af)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("n", [1, 5, 20, 33, 100, 257, 700, 1000])
def test_getrf_square_matches_lapack(dtype, n):
    hip = _hip()
    g = torch.Generator().manual_seed(n)
    b = 4
    A0 = torch.randn(b, n, n, generator=g, dtype=torch.float64)
    A = A0.to("cuda", dtype).contiguous()
    ipiv = torch.zeros(b, n, dtype=torch.int32, device="cuda")
    info = torch.zeros(b, dtype=torch.int32, device="cuda")
    hip.getrf(A, ipiv, info=info)
    torch.cuda.synchronize()
    LU = A.double().cpu().numpy()
    pv = ipiv.cpu().numpy()
    assert (info.cpu() == 0).all()
    for k in range(b):
        lu_ref, piv_ref = scipy.linalg.lu_factor(A0[k].numpy())
        if dtype == torch.float64:
            assert np.array_equal(pv[k] - 1, piv_ref)
            assert np.abs(LU[k] - lu_ref).max() < 1e-10 * max(1, np.abs(lu_ref).max())
        assert _lu_check(A0[k].to(dtype).double().numpy(), LU[k], pv[k], n) < 10 * TOL[dtype] * max(1, n / 10)


@pytest.mark.parametrize("n", [1100, 1500])
def test_getrf_tall_panel_phases(n):
    """1025-2048 rows start with 8-column panels of four rows per thread and
    switch to wider panels as the rows below the diagonal shrink (16, then
    32 columns); pivots as LAPACK's, factors to rounding."""
    hip = _hip()
    g = torch.Generator().manual_seed(n + 1)
    A0 = torch.randn(1, n, n, generator=g, dtype=torch.float64)
    A = A0.to("cuda").contiguous()
    ipiv = torch.zeros(1, n, dtype=torch.int32, device="cuda")
    hip.getrf(A, ipiv)
    torch.cuda.synchronize()
    lu_ref, piv_ref = scipy.linalg.lu_factor(A0[0].numpy())
    assert np.array_equal(ipiv.cpu().numpy()[0] - 1, piv_ref)
    assert np.abs(A.cpu().numpy()[0] - lu_ref).max() < 1e-9 * max(1, np.abs(lu_ref).max())


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_getrf_rectangular_and_bordered(dtype):
    """[A11 A12; A21 A22] with pivots restricted to the first ni rows:
    A12 <- L^{-1} P A12 and A22 <- A22 - A21 A11^{-1} A12."""
    hip = _hip()
    g = torch.Generator().manual_seed(3)
    b, ni, nx = 3, 100, 40
    M0 = torch.randn(b, ni + nx, ni + nx, generator=g, dtype=torch.float64)
    M0[:, :ni, :ni] += 5 * torch.eye(ni, dtype=torch.float64)
    for bordered in (False, True):
        nrows = ni + nx if bordered else ni
        M = M0[:, :nrows].to("cuda", dtype).contiguous()
        ipiv = torch.zeros(b, ni, dtype=torch.int32, device="cuda")
        hip.getrf(M, ipiv, nelim=ni, npiv=ni)
        torch.cuda.synchronize()
        Mc = M.double().cpu()
        A11, A12 = M0[:, :ni, :ni], M0[:, :ni, ni:]
        # U^{-1} (L^{-1} P A12) = A11^{-1} A12
        X = torch.linalg.solve_triangular(torch.triu(Mc[:, :ni, :ni]), Mc[:, :ni, ni:], upper=True)
        assert _rel(X, torch.linalg.solve(A11, A12)) < 100 * TOL[dtype]
        if bordered:
            S = M0[:, ni:, ni:] - M0[:, ni:, :ni] @ torch.linalg.solve(A11, A12)
            assert _rel(Mc[:, ni:, ni:], S) < 100 * TOL[dtype]


def _getrf_residual(A0, M, piv, nelim):
    """||P A - L U|| / ||A|| for getrf's packed output of an nrows x ncols
    matrix with nelim eliminated columns: L = the unit-lower multipliers in
    the first nelim columns (identity below for the rest), U = the rest of
    M, whose trailing rows are the Schur complement of the bordered rows."""
    A0 = np.asarray(A0, dtype=np.float64)
    M = np.asarray(M, dtype=np.float64)
    nrows, ncols = M.shape
    PA = A0.copy()
    for i, p in enumerate(piv[:nelim]):
        if p != i:
            PA[[i, p]] = PA[[p, i]]
    L = np.eye(nrows)
    L[:, :nelim] += np.tril(M[:, :nelim], -1)
    U = M.copy()
    U[:, :nelim] = np.triu(M[:, :nelim])
    return np.linalg.norm(L @ U - PA) / np.linalg.norm(A0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("b,nrows,ncols,nelim,npiv", [(2, 1000, 1000, 1000, 1000), (1, 300, 300, 300, 300),
                                                       (3, 700, 780, 700, 700), (2, 600, 650, 450, 500),
                                                       (1, 1500, 1500, 1500, 1500), (80, 400, 480, 400, 400),
                                                       (80, 300, 320, 260, 280)])
def test_getrf_panel_phases_reconstruct(dtype, b, nrows, ncols, nelim, npiv):
    """getrf's register panels change shape with the rows left below the
    diagonal (8/16/32 columns; few heavy wavefronts for FP32 and small FP64
    batches, more lighter ones for large FP64 batches) and move rows as
    16-byte vectors where aligned.  On square, rectangular (extra columns)
    and bordered (rows past npiv never pivot) shapes at both batch sizes:
    P A = L U with the bordered Schur complement, to the precision's
    rounding; no pivot row at or past npiv; FP64 square pivots as LAPACK's."""
    hip = _hip()
    g = torch.Generator().manual_seed(nrows + 7 * ncols + b)
    M0 = torch.randn(b, nrows, ncols, generator=g, dtype=torch.float64)
    M = M0.to("cuda", dtype).contiguous()
    ipiv = torch.zeros(b, nelim, dtype=torch.int32, device="cuda")
    info = torch.zeros(b, dtype=torch.int32, device="cuda")
    hip.getrf(M, ipiv, nelim=nelim, npiv=npiv, info=info)
    torch.cuda.synchronize()
    Mc, pc = M.double().cpu(), ipiv.cpu()
    assert (info.cpu() == 0).all()
    tol = 1e-13 if dtype == torch.float64 else 2e-5
    for k in range(0, b, max(1, b // 4)):
        assert (pc[k].numpy() - 1 < npiv).all()
        res = _getrf_residual(M0[k].to(dtype).double().numpy(), Mc[k].numpy(), pc[k].numpy() - 1, nelim)
        assert res < tol * max(1, nrows / 100)
        if dtype == torch.float64 and nrows == ncols == nelim == npiv:
            lu_ref, piv_ref = scipy.linalg.lu_factor(M0[k].numpy())
            assert np.array_equal(pc[k].numpy() - 1, piv_ref)
            assert np.abs(Mc[k].numpy() - lu_ref).max() < 1e-9 * max(1, np.abs(lu_ref).max())


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("trans", [False, True])
def test_getrs(dtype, trans):
    hip = _hip()
    g = torch.Generator().manual_seed(11)
    b, n, nrhs = 3, 70, 9
    A0 = torch.randn(b, n, n, generator=g, dtype=torch.float64) + 4 * torch.eye(n, dtype=torch.float64)
    B0 = torch.randn(b, n, nrhs, generator=g, dtype=torch.float64)
    A = A0.to("cuda", dtype).contiguous()
    ipiv = torch.zeros(b, n, dtype=torch.int32, device="cuda")
    hip.getrf(A, ipiv)
    X = B0.to("cuda", dtype).contiguous()
    hip.getrs(A, ipiv, X, trans=trans)
    ref = torch.linalg.solve(A0.transpose(1, 2) if trans else A0, B0)
    torch.cuda.synchronize()
    assert _rel(X, ref) < 100 * TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("n", [1, 5, 31, 32, 33, 100, 257, 1000])
def test_getri_inverse(dtype, n):
    """getrf + getri against LAPACK's inverse; the residual A X - I relative
    to |A| |X| is independent of the conditioning of the random matrices."""
    hip = _hip()
    g = torch.Generator().manual_seed(n)
    b = 2 if n >= 257 else 3
    A0 = torch.randn(b, n, n, generator=g, dtype=torch.float64)
    A = A0.to("cuda", dtype).contiguous()
    ipiv = torch.zeros(b, n, dtype=torch.int32, device="cuda")
    hip.getrf(A, ipiv)
    X = hip.getri(A, ipiv).double().cpu()
    eye = torch.eye(n, dtype=torch.float64)
    eps = torch.finfo(dtype).eps
    for i in range(b):
        res = torch.linalg.norm(A0[i] @ X[i] - eye) / (torch.linalg.norm(A0[i]) * torch.linalg.norm(X[i]))
        assert res < 10 * n ** 0.5 * eps, (i, res.item())
        ref = torch.linalg.inv(A0[i])
        cond = torch.linalg.cond(A0[i]).item()
        assert _rel(X[i], ref) < 20 * cond * eps, (i, _rel(X[i], ref), cond)


@pytest.mark.parametrize("n", [7, 64, 1000])
def test_getri_permutation_is_exact(n):
    """A permutation matrix factors as P^T I I; its inverse is its transpose,
    bit for bit, so any slip in the column interchanges shows."""
    hip = _hip()
    g = torch.Generator().manual_seed(3 + n)
    P = torch.stack([torch.eye(n, dtype=torch.float64)[torch.randperm(n, generator=g)] for _ in range(2)])
    A = P.to("cuda").contiguous()
    ipiv = torch.zeros(2, n, dtype=torch.int32, device="cuda")
    hip.getrf(A, ipiv)
    X = hip.getri(A, ipiv).cpu()
    assert torch.equal(X, P.transpose(1, 2))


def _btri_dense(diag, sub, sup):
    b, m, bs, _ = diag.shape
    M = torch.zeros(b, m * bs, m * bs, dtype=torch.float64)
    for i in range(m):
        M[:, i * bs:(i + 1) * bs, i * bs:(i + 1) * bs] = diag[:, i]
        if i < m - 1:
            M[:, (i + 1) * bs:(i + 2) * bs, i * bs:(i + 1) * bs] = sub[:, i]
            M[:, i * bs:(i + 1) * bs, (i + 1) * bs:(i + 2) * bs] = sup[:, i]
    return M


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("b,m,bs", [(3, 5, 20), (2, 1, 16), (1, 4, 100), (4, 2, 37), (2, 3, 24), (2, 11, 100),
                                    (2, 6, 154), (1, 3, 200)])
@pytest.mark.parametrize("transpose", [False, True])
@pytest.mark.parametrize("twisted", [False, True])
def test_btri_factor_solve(dtype, b, m, bs, transpose, twisted):
    """Block-tridiagonal LU factor + sweeps (sequential, or twisted: both
    chain ends towards block m//2), plain and transposed, against a dense
    solve.  Blocks wider than the one-launch chain kernels hold (FP64 above
    112, FP32 above 160) run the per-step batched launches, top-down only."""
    hip = _hip()
    if twisted and not hip.btri_fused(dtype, bs):
        pytest.skip("twisted factors need blocks the one-launch chain kernels hold")
    g = torch.Generator().manual_seed(b * 100 + m * 10 + bs)
    diag0 = torch.randn(b, m, bs, bs, generator=g, dtype=torch.float64) + 3 * torch.eye(bs, dtype=torch.float64)
    sub0 = torch.randn(b, max(m - 1, 1), bs, bs, generator=g, dtype=torch.float64) * 0.3
    sup0 = torch.randn(b, max(m - 1, 1), bs, bs, generator=g, dtype=torch.float64) * 0.3
    nrhs = 6
    f0 = torch.randn(b, m * bs, nrhs, generator=g, dtype=torch.float64)
    diag, sub, sup = (t.to("cuda", dtype).contiguous() for t in (diag0, sub0, sup0))
    ipiv = torch.zeros(b, m, bs, dtype=torch.int32, device="cuda")
    hip.btri_factor(diag, sub, sup, ipiv, twisted=twisted)
    f = f0.to("cuda", dtype).contiguous()
    hip.btri_solve(diag, sub, sup, ipiv, f, transpose=transpose, twisted=twisted)
    M = _btri_dense(diag0, sub0[:, :m - 1], sup0[:, :m - 1])
    ref = torch.linalg.solve(M.transpose(1, 2) if transpose else M, f0)
    torch.cuda.synchronize()
    assert _rel(f, ref) < 200 * TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("b,m,bs", [(3, 5, 20), (2, 1, 16), (2, 2, 37), (2, 11, 100), (3, 4, 100)])
@pytest.mark.parametrize("nrhs", [1, 3, 32, 70, 150])
@pytest.mark.parametrize("transpose", [False, True])
def test_btri_twisted_inverse_factors(dtype, b, m, bs, nrhs, transpose):
    """Twisted block LU whose diagonal blocks are applied through their
    triangular inverses (L^{-1} P, U^{-1}): the same solution as the
    substitution sweeps and as a dense solve, plain and transposed."""
    hip = _hip()
    g = torch.Generator().manual_seed(b * 1000 + m * 100 + bs + nrhs)
    diag0 = torch.randn(b, m, bs, bs, generator=g, dtype=torch.float64) + 3 * torch.eye(bs, dtype=torch.float64)
    sub0 = torch.randn(b, max(m - 1, 1), bs, bs, generator=g, dtype=torch.float64) * 0.3
    sup0 = torch.randn(b, max(m - 1, 1), bs, bs, generator=g, dtype=torch.float64) * 0.3
    f0 = torch.randn(b, m * bs, nrhs, generator=g, dtype=torch.float64)
    diag, sub, sup = (t.to("cuda", dtype).contiguous() for t in (diag0, sub0, sup0))
    ipiv = torch.zeros(b, m, bs, dtype=torch.int32, device="cuda")
    hip.btri_factor(diag, sub, sup, ipiv, twisted=True)
    f_lu = f0.to("cuda", dtype).contiguous()
    hip.btri_solve(diag, sub, sup, ipiv, f_lu, twisted=True, transpose=transpose)
    uinv = diag.clone()
    linv = torch.empty_like(uinv)
    hip.btri_invert_diag(uinv, ipiv, linv)
    f = f0.to("cuda", dtype).contiguous()
    hip.btri_solve(uinv, sub, sup, None, f, twisted=True, linv=linv, transpose=transpose)
    M = _btri_dense(diag0, sub0[:, :m - 1], sup0[:, :m - 1])
    ref = torch.linalg.solve(M.transpose(1, 2) if transpose else M, f0)
    torch.cuda.synchronize()
    assert _rel(f, ref) < 200 * TOL[dtype]
    assert _rel(f, f_lu.double().cpu()) < 200 * TOL[dtype]
    # U^{-1} is upper triangular and L^{-1} P has the unit lower triangle
    # of L^{-1} in its permuted columns: L^{-1} P P^T L = I on every block
    Ui = uinv.double().cpu()
    assert torch.count_nonzero(torch.tril(Ui, -1)) == 0


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("twisted", [False, True])
@pytest.mark.parametrize("nrhs", [1, 70])
@pytest.mark.parametrize("transpose", [False, True])
def test_btri_solve_block_sparse_rhs(dtype, twisted, nrhs, transpose):
    """Right-hand sides whose leading blocks are zero (a different first
    nonzero block per column, one all-zero 32-column tile): the forward
    sweeps start at the first nonzero block and must give the dense
    solution."""
    hip = _hip()
    b, m, bs = 2, 11, 24
    g = torch.Generator().manual_seed(7 + nrhs)
    diag0 = torch.randn(b, m, bs, bs, generator=g, dtype=torch.float64) + 3 * torch.eye(bs, dtype=torch.float64)
    sub0 = torch.randn(b, m - 1, bs, bs, generator=g, dtype=torch.float64) * 0.3
    sup0 = torch.randn(b, m - 1, bs, bs, generator=g, dtype=torch.float64) * 0.3
    f0 = torch.zeros(b, m, bs, nrhs, dtype=torch.float64)
    for c in range(nrhs):
        if 32 <= c < 64:
            continue                                   # an all-zero column tile
        first = (c * 5 + 3) % m
        f0[:, first:min(m, first + 2), :, c] = torch.randn(b, min(m, first + 2) - first, bs, generator=g,
                                                          dtype=torch.float64)
    f0 = f0.reshape(b, m * bs, nrhs)
    diag, sub, sup = (t.to("cuda", dtype).contiguous() for t in (diag0, sub0, sup0))
    ipiv = torch.zeros(b, m, bs, dtype=torch.int32, device="cuda")
    hip.btri_factor(diag, sub, sup, ipiv, twisted=twisted)
    f = f0.to("cuda", dtype).contiguous()
    hip.btri_solve(diag, sub, sup, ipiv, f, twisted=twisted, transpose=transpose)
    M = _btri_dense(diag0, sub0, sup0)
    ref = torch.linalg.solve(M.transpose(1, 2) if transpose else M, f0)
    torch.cuda.synchronize()
    assert _rel(f, ref) < 200 * TOL[dtype]
    if nrhs > 64:
        assert torch.count_nonzero(f[:, :, 32:64]) == 0


@pytest.mark.parametrize("p", [10, 22])
@pytest.mark.parametrize("general", [False, True])
@pytest.mark.parametrize("nrhs", [1, 3])
def test_leaf_residual_matches_dense_operator(p, general, nrhs):
    """Fused FP64 leaf residual (node assembly through Interp_mat, -(A_loc u)
    at the interiors, N_x u at the faces) against the oracle's dense p^2 x p^2
    leaf operator and N_x."""
    from oracle import hps_oracle as O
    from slablu_amd.src_disc.hps_parallel_leaf_ops import LeafOps
    from slablu_amd.src_disc.hps_subdomain_disc import HPS_Disc
    hip = _hip()
    a = 0.013
    ones = lambda xx: np.ones(xx.shape[0])
    if general:
        pdo = O.PDO(lambda xx: 1 + 0.3 * np.sin(7 * xx[:, 0]), lambda xx: 2 + xx[:, 1] ** 2,
                    c12=lambda xx: 0.1 * xx[:, 0] * xx[:, 1], c1=lambda xx: np.cos(xx[:, 1]),
                    c2=lambda xx: -xx[:, 0], c=lambda xx: -400 * (1 + 0.2 * np.cos(9 * xx[:, 0])))
    else:
        pdo = O.PDO(ones, ones, c=lambda xx: -300.0 * np.ones(xx.shape[0]))
    hps = O.HPSOracle(pdo, np.array([[0.0, 6 * a], [0.0, 4 * a]]), a, p)
    nb = hps.nboxes
    A = hps.aloc(np.arange(nb))
    Jc, Jxr = hps.JJ["Jc"], hps.JJ["Jxreorder"]
    xx = hps.grid_xx.reshape(-1, 2)
    coefs = {}
    for name in ("c11", "c22", "c12", "c1", "c2", "c"):
        fn = getattr(pdo, name)
        if fn is not None:
            coefs[name] = torch.tensor(np.asarray(fn(xx), dtype=np.float64).reshape(-1), device="cuda")
    ops = LeafOps(HPS_Disc(a, p), "cuda")
    f = p - 2
    ni, nx = f * f, 4 * f
    rng = np.random.default_rng(p + 10 * nrhs)
    uc = rng.standard_normal((nb, ni, nrhs))
    ub = rng.standard_normal((nb, nx, nrhs))
    u = np.zeros((nb, p * p, nrhs))
    u[:, Jc] = uc
    u[:, Jxr] = hps.Interp_mat[None] @ ub
    rI_ref = -(A @ u)[:, Jc]
    fl_ref = hps.Nx[None] @ u
    uu = torch.empty(nb, p * p, nrhs, dtype=torch.float64, device="cuda")
    rI, fl = ops.residual(torch.tensor(uc, device="cuda"), torch.tensor(ub, device="cuda"), coefs, uu=uu)
    torch.cuda.synchronize()
    assert np.linalg.norm(uu.cpu().numpy() - u) / np.linalg.norm(u) < 1e-14
    assert np.linalg.norm(rI.cpu().numpy() - rI_ref) / np.linalg.norm(rI_ref) < 1e-13
    assert np.linalg.norm(fl.cpu().numpy() - fl_ref) / np.linalg.norm(fl_ref) < 1e-13


@pytest.mark.parametrize("p", [10, 16, 22])
@pytest.mark.parametrize("general", [False, True])
def test_leaf_assemble_bit_exact(p, general):
    from oracle import hps_oracle as O
    hip = _hip()
    rng = np.random.default_rng(p)
    a = 0.013
    ones = lambda xx: np.ones(xx.shape[0])
    if general:
        pdo = O.PDO(lambda xx: 1 + 0.3 * np.sin(7 * xx[:, 0]), lambda xx: 2 + xx[:, 1] ** 2,
                    c12=lambda xx: 0.1 * xx[:, 0] * xx[:, 1], c1=lambda xx: np.cos(xx[:, 1]),
                    c2=lambda xx: -xx[:, 0], c=lambda xx: -400 * (1 + 0.2 * np.cos(9 * xx[:, 0])))
    else:
        pdo = O.PDO(ones, ones, c=lambda xx: -300.0 * np.ones(xx.shape[0]))
    hps = O.HPSOracle(pdo, np.array([[0.0, 6 * a], [0.0, 4 * a]]), a, p)
    nb = hps.nboxes
    A = hps.aloc(np.arange(nb))
    Jc, Jx = hps.JJ["Jc"], hps.JJ["Jx"]
    Jtot = np.hstack((Jc, Jx))
    xx = hps.grid_xx.reshape(-1, 2)
    coefs = {}
    for name in ("c11", "c22", "c12", "c1", "c2", "c"):
        fn = getattr(pdo, name)
        if fn is not None:
            coefs[name] = torch.tensor(np.asarray(fn(xx), dtype=np.float64).reshape(-1), device="cuda")
    D = torch.tensor(hps.D, device="cuda")
    Dsq = torch.tensor(hps.Dsq, device="cuda")
    Nxtot = torch.tensor(np.ascontiguousarray(hps.Nx[:, Jtot]), device="cuda")
    ni, nx = len(Jc), len(Jx)
    M = torch.zeros(nb, ni + nx, ni + nx, dtype=torch.float64, device="cuda")
    hip.leaf_assemble(M, p, 1, D, Dsq, coefs, Nxtot)
    torch.cuda.synchronize()
    Mc = M.cpu().numpy()
    ref = A[:, Jc][:, :, Jtot]
    assert np.array_equal(Mc[:, :ni], ref)
    assert np.array_equal(Mc[:, ni:], np.broadcast_to(hps.Nx[:, Jtot], (nb, nx, ni + nx)))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("b,m,bs", [(5, 4, 20), (3, 7, 100), (2, 3, 154), (1, 3, 200), (4, 1, 33), (2, 2, 16)])
@pytest.mark.parametrize("nrhs", [1, 7, 40])
def test_btri_inv_factor_solve(dtype, b, m, bs, nrhs):
    """Block-Thomas factorization with explicit inverses (fused in LDS for
    small blocks, batched getrf/getrs/gemm otherwise) vs a dense solve."""
    hip = _hip()
    g = torch.Generator().manual_seed(b * 1000 + m * 10 + bs + nrhs)
    # diagonal shift ~ 2 sqrt(bs) keeps the blocks well conditioned at every size
    shift = 3 + 2 * bs ** 0.5
    diag0 = torch.randn(b, m, bs, bs, generator=g, dtype=torch.float64) + shift * torch.eye(bs, dtype=torch.float64)
    sub0 = torch.randn(b, max(m - 1, 1), bs, bs, generator=g, dtype=torch.float64) * 0.3
    sup0 = torch.randn(b, max(m - 1, 1), bs, bs, generator=g, dtype=torch.float64) * 0.3
    f0 = torch.randn(b, m * bs, nrhs, generator=g, dtype=torch.float64)
    diag, sub, sup = (t.to("cuda", dtype).contiguous() for t in (diag0, sub0, sup0))
    info = torch.zeros(b, dtype=torch.int32, device="cuda")
    hip.btri_inv_factor(diag, sub, sup, info=info)
    f = f0.to("cuda", dtype).contiguous()
    hip.btri_inv_solve(diag, sub, sup, f)
    M = _btri_dense(diag0, sub0[:, :m - 1], sup0[:, :m - 1])
    ref = torch.linalg.solve(M, f0)
    torch.cuda.synchronize()
    assert int(info.abs().max()) == 0
    assert _rel(f, ref) < 200 * TOL[dtype]
    # the diagonal slot holds S_0^{-1} = D_0^{-1}
    assert _rel(diag[:, 0], torch.linalg.inv(diag0[:, 0])) < 200 * TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("n", [33, 100, 400])
@pytest.mark.parametrize("nrhs", [1, 3, 8])
def test_getrs_small_rhs(dtype, n, nrhs):
    hip = _hip()
    g = torch.Generator().manual_seed(n + nrhs)
    b = 5
    A0 = torch.randn(b, n, n, generator=g, dtype=torch.float64) + 2 * torch.eye(n, dtype=torch.float64)
    B0 = torch.randn(b, n, nrhs, generator=g, dtype=torch.float64)
    A = A0.to("cuda", dtype).contiguous()
    ipiv = torch.zeros(b, n, dtype=torch.int32, device="cuda")
    hip.getrf(A, ipiv)
    X = B0.to("cuda", dtype).contiguous()
    hip.getrs(A, ipiv, X)
    torch.cuda.synchronize()
    assert _rel(X, torch.linalg.solve(A0, B0)) < 500 * TOL[dtype]


@pytest.mark.parametrize("form", ["inv", "lu"])
@pytest.mark.parametrize("twisted", [False, True])
def test_block_tridiagonal_api(form, twisted):
    """Block_Tridiagonal(b, m, bs) drop-in: build_sweep / lusolve_sweep on
    slices of the batch, as slablu_solver.py uses it."""
    from slablu_amd.src_solver.block_tridiag_solver_dense import Block_Tridiagonal
    g = torch.Generator().manual_seed(7)
    b, m, bs, nrhs = 6, 5, 20, 3
    diag0 = torch.randn(b, m, bs, bs, generator=g, dtype=torch.float64) + 4 * torch.eye(bs, dtype=torch.float64)
    sub0 = torch.randn(b, m - 1, bs, bs, generator=g, dtype=torch.float64) * 0.3
    sup0 = torch.randn(b, m - 1, bs, bs, generator=g, dtype=torch.float64) * 0.3
    T = Block_Tridiagonal(b, m, bs, form=form, twisted=twisted)
    T.build_sweep(diag0.cuda(), sub0.cuda(), sup0.cuda())
    assert T.Tdata.shape == (b, T.storage_blocks, bs, bs)
    f0 = torch.randn(b, m * bs, nrhs, generator=g, dtype=torch.float64)
    M = _btri_dense(diag0, sub0, sup0)
    x = T.lusolve_sweep(f0[2:5].cuda().clone(), b_start=2, b_end=5)
    torch.cuda.synchronize()
    assert _rel(x, torch.linalg.solve(M[2:5], f0[2:5])) < 1e-11
    if form == "lu":
        xt = T.lusolve_sweep(f0.cuda().clone(), transpose=True)
        torch.cuda.synchronize()
        assert _rel(xt, torch.linalg.solve(M.transpose(1, 2), f0)) < 1e-11


@pytest.mark.parametrize("order", ["twisted", "separated", "auto"])
@pytest.mark.parametrize("nb", [1, 2, 3, 4, 5, 6, 8, 9, 13])
def test_hbs_block_tridiagonal_api(nb, order):
    """Twisted (two-ended) and separated (even blocks, then the odd blocks'
    Schur complement, recursively) block elimination against a dense solve,
    plain and transposed, for odd, even and degenerate block counts; the
    separated order's dependent steps are counted from its levels."""
    from slablu_amd.src_solver.block_tridiag_solver_structured import HBS_BlockTridiagonal
    g = torch.Generator().manual_seed(9 + nb)
    n = 300 if nb < 13 else 120
    # well-conditioned blocks (random parts of spectral norm about 2 and 1
    # around 5 I), so that the tolerance measures the algebra of each order
    # and not the conditioning of the blocks it inverts
    s = 1.0 / n ** 0.5
    d0 = torch.randn(nb, n, n, generator=g, dtype=torch.float64) * s + 5 * torch.eye(n, dtype=torch.float64)
    d1 = torch.randn(nb, n, n, generator=g, dtype=torch.float64) * 0.1 * s
    sub = torch.randn(max(nb - 1, 0), n, n, generator=g, dtype=torch.float64) * 0.5 * s
    sup = torch.randn(max(nb - 1, 0), n, n, generator=g, dtype=torch.float64) * 0.5 * s
    H = HBS_BlockTridiagonal(nb, n, order=order)
    H.build_lowmem(d0, d1, sub, sup, device="cuda")
    f = torch.randn(nb * n, 2, generator=g, dtype=torch.float64)
    x = H.solve(f.cuda())
    xt = H.solve(f.cuda(), transpose=True)
    M = _btri_dense((d0 + d1)[None], sub[None], sup[None])[0]
    torch.cuda.synchronize()
    assert _rel(x, torch.linalg.solve(M, f)) < 1e-11
    assert _rel(xt, torch.linalg.solve(M.T, f)) < 1e-11
    assert int((H.info != 0).sum()) == 0
    # the levels this build ran: one batched round per separated level
    levels, h = 0, H
    while getattr(h, "separated", False):
        levels, h = levels + 1, h.Tsolver
    assert levels + (h.b // 2 + 1) == HBS_BlockTridiagonal.dependent_steps(nb, order)
    if order == "separated" and nb >= 2 or order == "auto" and nb >= 6:
        assert H.separated


@pytest.mark.parametrize("nb", [2, 3, 4, 7, 10])
def test_hbs_block_tridiagonal_cyclic(nb):
    """Cyclic (periodic) interface ring: chain over blocks 0..nb-2 bordered by
    block nb-1, against a dense solve of the ring matrix (for nb = 2 the two
    corner couplings add onto the same blocks)."""
    from slablu_amd.src_solver.block_tridiag_solver_structured import HBS_BlockTridiagonal
    g = torch.Generator().manual_seed(19 + nb)
    n = 200
    diag = torch.randn(nb, n, n, generator=g, dtype=torch.float64) + 6 * torch.eye(n, dtype=torch.float64)
    sub = torch.randn(nb, n, n, generator=g, dtype=torch.float64) * 0.2
    sup = torch.randn(nb, n, n, generator=g, dtype=torch.float64) * 0.2
    M = torch.zeros(nb * n, nb * n, dtype=torch.float64)
    for j in range(nb):
        k = (j + 1) % nb
        M[j * n:(j + 1) * n, j * n:(j + 1) * n] += diag[j]
        M[j * n:(j + 1) * n, k * n:(k + 1) * n] += sup[j]      # block (j, j+1)
        M[k * n:(k + 1) * n, j * n:(j + 1) * n] += sub[j]      # block (j+1, j)
    for order in ("twisted", "separated"):
        H = HBS_BlockTridiagonal(nb, n, cyclic=True, order=order)
        H.build_lowmem(diag, None, sub, sup, device="cuda")
        f = torch.randn(nb * n, 3, generator=g, dtype=torch.float64)
        x = H.solve(f.cuda())
        xt = H.solve(f.cuda(), transpose=True)
        torch.cuda.synchronize()
        assert _rel(x, torch.linalg.solve(M, f)) < 1e-11
        assert _rel(xt, torch.linalg.solve(M.T, f)) < 1e-11


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("nrhs", [1, 2, 16, 40])
def test_btri_twisted_rem_shape(dtype, nrhs):
    """Twisted factor + sweeps at the slab-chain shape (bs = 100, odd and
    even chain lengths) for single, few and many right-hand sides."""
    hip = _hip()
    for b, m in ((2, 9), (1, 8)):
        g = torch.Generator().manual_seed(31 * m + nrhs)
        bs = 100
        diag0 = torch.randn(b, m, bs, bs, generator=g, dtype=torch.float64) + 3 * torch.eye(bs, dtype=torch.float64)
        sub0 = torch.randn(b, m - 1, bs, bs, generator=g, dtype=torch.float64) * 0.3
        sup0 = torch.randn(b, m - 1, bs, bs, generator=g, dtype=torch.float64) * 0.3
        f0 = torch.randn(b, m * bs, nrhs, generator=g, dtype=torch.float64)
        diag, sub, sup = (t.to("cuda", dtype).contiguous() for t in (diag0, sub0, sup0))
        ipiv = torch.zeros(b, m, bs, dtype=torch.int32, device="cuda")
        hip.btri_factor(diag, sub, sup, ipiv, twisted=True)
        f = f0.to("cuda", dtype).contiguous()
        hip.btri_solve(diag, sub, sup, ipiv, f, twisted=True)
        ref = torch.linalg.solve(_btri_dense(diag0, sub0, sup0), f0)
        torch.cuda.synchronize()
        assert _rel(f, ref) < 200 * TOL[dtype], (b, m)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("n,ld", [(400, 480), (100, 100), (37, 41), (64, 64), (130, 131)])
@pytest.mark.parametrize("nrhs", [1, 3, 8])
def test_getrs_perm_matches_getrs(dtype, n, ld, nrhs):
    """One-wavefront-per-matrix solve with the composed pivot permutation
    against the reference getrs (aligned and unaligned row strides, partial
    last diagonal blocks)."""
    hip = _hip()
    g = torch.Generator().manual_seed(n + nrhs)
    b = 9
    A0 = torch.randn(b, n, ld, generator=g, dtype=torch.float64)
    A0[:, :, :n] += 4 * torch.eye(n, dtype=torch.float64)
    M = A0.to("cuda", dtype).contiguous()
    piv = torch.zeros(b, n, dtype=torch.int32, device="cuda")
    hip.getrf(M, piv, nelim=n, npiv=n)
    perm = hip.piv_to_perm(piv)
    f0 = torch.randn(b, n, nrhs, generator=g, dtype=torch.float64).to("cuda", dtype)
    x1 = f0.clone()
    hip.getrs(M[:, :, :n], piv, x1)
    x2 = f0.clone()
    hip.getrs(M[:, :, :n], piv, x2, perm=perm)
    ref = torch.linalg.solve(A0[:, :, :n], f0.double().cpu())
    torch.cuda.synchronize()
    assert _rel(x2, ref) < 200 * TOL[dtype]
    assert _rel(x2, x1.double().cpu()) < 50 * TOL[dtype]


def test_launch_probe_counts_gemm_launches():
    """The launch probe times exactly the launches of the chosen GEMM kernel:
    FP32 products with 16-byte rows (k_gemm<float, int, true, true>) are
    counted with 2 m n k flops each, FP64 and unaligned products are not;
    launches past the capacity are reported as untimed."""
    hip = _hip()
    A = torch.randn(3, 64, 96, device="cuda")
    B = torch.randn(3, 96, 128, device="cuda")
    C = torch.empty(3, 64, 128, device="cuda")
    hip.probe_start(hip.PROBE_GEMM_F32, 4)
    for _ in range(5):
        hip.gemm(A, B, C)
    hip.gemm(A.double(), B.double(), C.double())          # another kernel
    hip.gemm(A[:, :, 1:], B[:, 1:, :], C)                  # unaligned rows: another variant
    r = hip.probe_stop()
    assert r["launches"] == 4 and r["dropped"] == 1
    assert r["flops"] == 4 * 2 * 3 * 64 * 128 * 96
    assert r["bytes"] == 4 * 3 * 4 * (64 * 96 + 96 * 128 + 64 * 128)
    assert r["total_ms"] > 0
    torch.cuda.synchronize()
    assert torch.allclose(C, torch.bmm(A[:, :, 1:], B[:, 1:, :]), rtol=1e-4, atol=1e-3)
    with pytest.raises(RuntimeError):
        hip.probe_stop()                                     # not running
