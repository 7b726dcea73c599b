"""ctypes binding of libslabhip.so (include/slab_hip.h) for torch tensors.

Every function takes torch tensors that live on the GPU and passes their
data pointers and strides to the C ABI; all launches are asynchronous on the
current torch stream.  The library is mandatory: it is loaded on first use
and a missing library raises there; there is no CPU fallback anywhere in
the product path.
"""
import ctypes
import os

import torch

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libslabhip.so")

F64, F32 = 0, 1
_i64 = ctypes.c_int64
_int = ctypes.c_int
_vp = ctypes.c_void_p
_dbl = ctypes.c_double


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libslabhip.so is not built (%s); build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python slablu_amd/_build.py`" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "slab_version": ([], _int),
        "slab_gemm": ([_int, _int, _int, _int, _int, _dbl, _vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64, _dbl, _vp,
                       _i64, _i64, _i64, _vp], _int),
        "slab_trsm": ([_int, _int, _int, _int, _int, _int, _vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64, _vp], _int),
        "slab_laswp": ([_int, _int, _int, _int, _int, _int, _vp, _i64, _i64, _i64, _vp, _i64, _vp], _int),
        "slab_getrf": ([_int, _int, _int, _int, _int, _int, _vp, _i64, _i64, _vp, _i64, _vp, _vp], _int),
        "slab_getrs": ([_int, _int, _int, _int, _int, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _i64, _vp], _int),
        "slab_piv_to_perm": ([_int, _int, _vp, _i64, _vp, _i64, _vp], _int),
        "slab_getri_ws": ([_int, _int, _int], ctypes.c_size_t),
        "slab_orth": ([_int, _int, _int, _int, _vp, _i64, _i64, _vp, _i64, _i64, _vp], _int),
        "slab_getri": ([_int, _int, _int, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _i64, _vp, _vp], _int),
        "slab_getri_noalloc": ([_int, _int, _int, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _i64, _vp], _int),
        "slab_getrs_perm": ([_int, _int, _int, _int, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _i64, _vp], _int),
        "slab_btri_factor": ([_int, _int, _int, _int, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64,
                              _i64, _vp, _vp], _int),
        "slab_btri_solve": ([_int, _int, _int, _int, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64,
                             _i64, _vp, _i64, _i64, _int, _int, _vp], _int),
        "slab_btri_inv_fused": ([_int, _int], _int),
        "slab_btri_invert_diag": ([_int, _int, _int, _int, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _vp, _vp], _int),
        "slab_btri_solve_twisted_inv": ([_int, _int, _int, _int, _vp, _i64, _i64, _vp, _vp, _i64, _i64, _vp, _i64,
                                         _i64, _vp, _i64, _i64, _int, _int, _vp], _int),
        "slab_btri_fused": ([_int, _int], _int),
        "slab_btri_factor_twisted": ([_int, _int, _int, _int, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp,
                                      _i64, _i64, _vp, _vp], _int),
        "slab_btri_solve_twisted": ([_int, _int, _int, _int, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp,
                                     _i64, _i64, _vp, _i64, _i64, _int, _int, _vp], _int),
        "slab_btri_inv_factor": ([_int, _int, _int, _int, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp,
                                  _vp, _vp, _vp], _int),
        "slab_btri_inv_solve": ([_int, _int, _int, _int, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp,
                                 _i64, _i64, _int, _vp, _vp], _int),
        "slab_leaf_assemble": ([_int, _int, _int, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64,
                                _vp], _int),
        "slab_leaf_assemble_nodes": ([_int, _int, _int, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                      _i64, _i64, _vp], _int),
        "slab_set_sweep_tile": ([_int, _int, _int, _int], _int),
        "slab_mdot_ws": ([_int, _int, _i64], ctypes.c_size_t),
        "slab_mdot": ([_int, _int, _i64, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _vp], _int),
        "slab_maxpy": ([_int, _int, _i64, _dbl, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp], _int),
        "slab_probe_start": ([_int, _int], _int),
        "slab_probe_stop": ([ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.POINTER(_dbl), ctypes.POINTER(_dbl),
                             ctypes.POINTER(_dbl)], _int),
        "slab_leaf_residual": ([_int, _int, _int, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp],
                               _int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


class _Lazy:
    _lib = None

    def __getattr__(self, name):
        if _Lazy._lib is None:
            _Lazy._lib = _load()
        return getattr(_Lazy._lib, name)


LIB = _Lazy()


def _check(rc, name):
    if rc != 0:
        raise RuntimeError("%s failed with status %d" % (name, rc))


def _dt(t):
    if t.dtype == torch.float64:
        return F64
    if t.dtype == torch.float32:
        return F32
    raise TypeError("unsupported dtype %s" % t.dtype)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _mat(t):
    """(batch, rows, cols) strides of a 2-D or 3-D tensor view."""
    if t.dim() == 2:
        return 1, t.stride(0), t.stride(1), 0
    if t.dim() == 3:
        return t.shape[0], t.stride(1), t.stride(2), t.stride(0)
    raise ValueError("expected a 2-D or 3-D tensor, got shape %s" % (tuple(t.shape),))


def _require(cond, msg):
    if not cond:
        raise ValueError(msg)


def version():
    return LIB.slab_version()


def gemm(A, B, C, alpha=1.0, beta=0.0):
    """C = alpha A@B + beta C for (batched) strided views; 2-D operands are
    broadcast over the batch of C."""
    bc, crs, ccs, cbs = _mat(C)
    ba, ars, acs, abs_ = _mat(A)
    bb, brs, bcs, bbs = _mat(B)
    m, n = C.shape[-2], C.shape[-1]
    k = A.shape[-1]
    _require(A.shape[-2] == m and B.shape[-2] == k and B.shape[-1] == n, "gemm shape mismatch")
    _require(ba in (1, bc) and bb in (1, bc), "gemm batch mismatch")
    _require(A.dtype == B.dtype == C.dtype and A.is_cuda and B.is_cuda and C.is_cuda, "gemm dtype/device mismatch")
    if ba == 1 and A.dim() == 3:
        abs_ = 0
    if bb == 1 and B.dim() == 3:
        bbs = 0
    _check(LIB.slab_gemm(_dt(C), bc, m, n, k, float(alpha), _p(A), ars, acs, abs_, _p(B), brs, bcs, bbs,
                         float(beta), _p(C), crs, ccs, cbs, _stream()), "slab_gemm")
    return C


def trsm(A, B, lower, unit):
    """B <- A^{-1} B in place (A triangular as viewed)."""
    bb, brs, bcs, bbs = _mat(B)
    ba, ars, acs, abs_ = _mat(A)
    m, n = B.shape[-2], B.shape[-1]
    _require(A.shape[-1] == m and A.shape[-2] == m, "trsm shape mismatch")
    _require(ba in (1, bb), "trsm batch mismatch")
    if ba == 1 and A.dim() == 3:
        abs_ = 0
    _check(LIB.slab_trsm(_dt(B), bb, m, n, int(lower), int(unit), _p(A), ars, acs, abs_, _p(B), brs, bcs, bbs,
                         _stream()), "slab_trsm")
    return B


def laswp(A, ipiv, k1, k2, forward=True):
    bb, rs, cs, bs = _mat(A)
    pbs = ipiv.stride(0) if ipiv.dim() == 2 else 0
    _check(LIB.slab_laswp(_dt(A), bb, A.shape[-1], k1, k2, int(forward), _p(A), rs, cs, bs, _p(ipiv), pbs,
                          _stream()), "slab_laswp")
    return A


def getrf(A, ipiv, nelim=None, npiv=None, info=None):
    """In-place batched LU (row-major, unit column stride) of A (batch, nrows,
    ncols) eliminating the first nelim columns; ipiv (batch, >= nelim) int32."""
    _require(A.dim() == 3 and A.stride(2) == 1, "getrf expects a (batch, rows, cols) row-major tensor")
    b, nr, nc = A.shape
    nelim = min(nr, nc) if nelim is None else nelim
    npiv = nr if npiv is None else npiv
    _require(ipiv.dtype == torch.int32 and ipiv.shape[0] == b and ipiv.shape[1] >= nelim, "ipiv shape/dtype")
    _check(LIB.slab_getrf(_dt(A), b, nr, nc, nelim, npiv, _p(A), A.stride(1), A.stride(0), _p(ipiv), ipiv.stride(0),
                          _p(info), _stream()), "slab_getrf")
    return A


def piv_to_perm(ipiv, n=None):
    """(batch, >= n) 1-based getrf pivots -> (batch, n) int32 row permutation:
    row r of P B is row perm[r] of B."""
    b = ipiv.shape[0]
    n = ipiv.shape[1] if n is None else n
    _require(ipiv.dtype == torch.int32 and ipiv.stride(-1) == 1, "ipiv layout")
    perm = torch.empty(b, n, dtype=torch.int32, device=ipiv.device)
    _check(LIB.slab_piv_to_perm(b, n, _p(ipiv), ipiv.stride(0), _p(perm), perm.stride(0), _stream()),
           "slab_piv_to_perm")
    return perm


def orth(A, Q=None):
    """Q (batch, m, k) with orthonormal columns spanning the columns of each
    A (batch, m, k), m >= k, k <= 256, by Householder QR (A is destroyed)."""
    _require(A.dim() == 3 and A.stride(2) == 1 and A.shape[1] >= A.shape[2] and A.shape[2] <= 256,
             "orth expects (batch, m, k) row-major blocks with m >= k, k <= 256")
    b, m, k = A.shape
    if Q is None:
        Q = torch.empty(b, m, k, dtype=A.dtype, device=A.device)
    _require(Q.shape == A.shape and Q.dtype == A.dtype and Q.stride(2) == 1, "orth output layout")
    _check(LIB.slab_orth(_dt(A), b, m, k, _p(A), A.stride(1), A.stride(0), _p(Q), Q.stride(1), Q.stride(0),
                         _stream()), "slab_orth")
    return Q


def getri(LU, ipiv, X=None):
    """Batched inverse from getrf factors: X (batch, n, n) = A^{-1}, out of
    place (a new tensor unless X is given)."""
    _require(LU.dim() == 3 and LU.stride(2) == 1 and LU.shape[1] == LU.shape[2], "getri expects square row-major LU")
    b, n, _ = LU.shape
    _require(ipiv.dtype == torch.int32 and ipiv.shape[0] == b and ipiv.shape[1] >= n and ipiv.stride(-1) == 1,
             "ipiv shape/dtype")
    if X is None:
        X = torch.empty(b, n, n, dtype=LU.dtype, device=LU.device)
    _require(X.shape == LU.shape and X.dtype == LU.dtype and X.stride(2) == 1, "getri output layout")
    ws = torch.empty(int(LIB.slab_getri_ws(_dt(LU), b, n)), dtype=torch.uint8, device=LU.device)
    _check(LIB.slab_getri(_dt(LU), b, n, _p(LU), LU.stride(1), LU.stride(0), _p(ipiv), ipiv.stride(0), _p(X),
                          X.stride(1), X.stride(0), _p(ws), _stream()), "slab_getri")
    return X


GETRI_NOALLOC_MAX_BYTES = 40 * 1024 * 1024


def getri_noalloc(LU, ipiv, X):
    """getri into a caller-provided X with no device allocation: the
    workspace is the library's static per-device scratch array (at most
    GETRI_NOALLOC_MAX_BYTES; larger problems raise ValueError).  Calls that
    overlap on one device race on that array."""
    _require(LU.dim() == 3 and LU.stride(2) == 1 and LU.shape[1] == LU.shape[2], "getri expects square row-major LU")
    b, n, _ = LU.shape
    _require(ipiv.dtype == torch.int32 and ipiv.shape[0] == b and ipiv.shape[1] >= n and ipiv.stride(-1) == 1,
             "ipiv shape/dtype")
    _require(X.shape == LU.shape and X.dtype == LU.dtype and X.stride(2) == 1, "getri output layout")
    rc = LIB.slab_getri_noalloc(_dt(LU), b, n, _p(LU), LU.stride(1), LU.stride(0), _p(ipiv), ipiv.stride(0), _p(X),
                                X.stride(1), X.stride(0), _stream())
    if rc == -2:
        raise ValueError("slab_getri_noalloc: %d x %d x %d needs %d bytes of workspace, more than the %d of its "
                         "static scratch" % (b, n, n, int(LIB.slab_getri_ws(_dt(LU), b, n)), GETRI_NOALLOC_MAX_BYTES))
    _check(rc, "slab_getri_noalloc")
    return X


def getrs(LU, ipiv, B, trans=False, perm=None):
    """In-place batched solve op(A) X = B with getrf factors; with perm (from
    piv_to_perm) and at most 8 columns the one-wavefront-per-matrix kernel
    runs."""
    _require(LU.dim() == 3 and B.dim() == 3 and LU.stride(2) == 1 and B.stride(2) == 1, "getrs layouts")
    b, n, _ = LU.shape
    if perm is not None and not trans and 1 <= B.shape[2] <= 8:
        _check(LIB.slab_getrs_perm(_dt(B), b, n, B.shape[2], _p(LU), LU.stride(1), LU.stride(0), _p(perm),
                                   perm.stride(0), _p(B), B.stride(1), B.stride(0), _stream()), "slab_getrs_perm")
        return B
    _check(LIB.slab_getrs(_dt(B), b, n, B.shape[2], int(trans), _p(LU), LU.stride(1), LU.stride(0), _p(ipiv),
                          ipiv.stride(0), _p(B), B.stride(1), B.stride(0), _stream()), "slab_getrs")
    return B


def btri_fused(dtype, bs):
    """True when bs x bs blocks run in the fused LU-form chain kernels (the
    twisted factorization requires it)."""
    return LIB.slab_btri_fused(0 if dtype == torch.float64 else 1, int(bs)) == 1


def btri_factor(diag, sub, sup, ipiv, info=None, twisted=False):
    """Block-tridiagonal LU in place: diag (b, m, bs, bs), sub/sup (b, m-1, bs,
    bs), ipiv (b, m, bs) int32; blocks must be contiguous row-major.
    twisted: eliminate from both chain ends towards block m//2 (factors then
    serve btri_solve(..., twisted=True) only)."""
    b, m, bs, _ = diag.shape
    for t in (diag, sub, sup):
        _require(t.stride(-1) == 1 and t.stride(-2) == bs, "btri blocks must be contiguous row-major")
    _require(ipiv.stride(-1) == 1, "ipiv layout")
    fn = LIB.slab_btri_factor_twisted if twisted else LIB.slab_btri_factor
    _check(fn(_dt(diag), b, m, bs, _p(diag), diag.stride(0), diag.stride(1), _p(sub),
              sub.stride(0), sub.stride(1) if m > 1 else 0, _p(sup), sup.stride(0),
              sup.stride(1) if m > 1 else 0, _p(ipiv), ipiv.stride(0), ipiv.stride(1), _p(info),
              _stream()), "slab_btri_factor_twisted" if twisted else "slab_btri_factor")


def btri_invert_diag(diag, ipiv, linv):
    """After btri_factor: diag blocks <- U^{-1}, linv (same shape) <- L^{-1} P;
    btri_solve(..., linv=linv) then sweeps with block products."""
    b, m, bs, _ = diag.shape
    _require(diag.is_contiguous() and linv.is_contiguous() and linv.shape == diag.shape and
             linv.dtype == diag.dtype and ipiv.is_contiguous() and ipiv.shape == (b, m, bs),
             "btri_invert_diag layouts")
    ws = torch.empty(int(LIB.slab_getri_ws(_dt(diag), b * m, bs)), dtype=torch.uint8, device=diag.device)
    _check(LIB.slab_btri_invert_diag(_dt(diag), b, m, bs, _p(diag), diag.stride(0), diag.stride(1), _p(ipiv),
                                     ipiv.stride(0), ipiv.stride(1), _p(linv), _p(ws), _stream()),
           "slab_btri_invert_diag")
    return linv


def btri_solve(diag, sub, sup, ipiv, f, transpose=False, twisted=False, linv=None):
    """In-place sweeps on f (b, m*bs, nrhs) (row-major rows); linv: the
    inverse factors of btri_invert_diag (twisted factors only)."""
    b, m, bs, _ = diag.shape
    _require(f.dim() == 3 and f.shape[0] == b and f.shape[1] == m * bs and f.stride(2) == 1, "btri rhs layout")
    if linv is not None:
        _require(twisted, "inverse-factor sweeps need twisted factors")
        _check(LIB.slab_btri_solve_twisted_inv(_dt(diag), b, m, bs, _p(diag), diag.stride(0), diag.stride(1),
                                               _p(linv), _p(sub), sub.stride(0), sub.stride(1) if m > 1 else 0,
                                               _p(sup), sup.stride(0), sup.stride(1) if m > 1 else 0, _p(f),
                                               f.stride(0), f.stride(1), f.shape[2], int(transpose), _stream()),
               "slab_btri_solve_twisted_inv")
        return f
    args = (_dt(diag), b, m, bs, _p(diag), diag.stride(0), diag.stride(1), _p(sub), sub.stride(0),
            sub.stride(1) if m > 1 else 0, _p(sup), sup.stride(0), sup.stride(1) if m > 1 else 0, _p(ipiv),
            ipiv.stride(0), ipiv.stride(1), _p(f), f.stride(0), f.stride(1), f.shape[2])
    if twisted:
        _check(LIB.slab_btri_solve_twisted(*args, int(transpose), _stream()), "slab_btri_solve_twisted")
    else:
        _check(LIB.slab_btri_solve(*args, int(transpose), _stream()), "slab_btri_solve")
    return f


def btri_inv_fused(dtype, bs):
    return LIB.slab_btri_inv_fused(F64 if dtype == torch.float64 else F32, bs) == 1


def _blk_strides(t, m):
    return t.stride(0), (t.stride(1) if m > 1 else 0)


def btri_inv_factor(diag, sub, sup, info=None):
    """Block-Thomas factorization in place (diag_i <- S_i^{-1}, sup_i <- G_i);
    diag (b, m, bs, bs), sub/sup (b, >= m-1, bs, bs), contiguous blocks."""
    b, m, bs, _ = diag.shape
    for t in (diag, sub, sup):
        _require(t.stride(-1) == 1 and t.stride(-2) == bs and t.dtype == diag.dtype, "btri block layout")
    ws = iws = None
    if not btri_inv_fused(diag.dtype, bs):
        ws = torch.empty(2 * b * bs * bs, dtype=diag.dtype, device=diag.device)
        iws = torch.empty(b * bs, dtype=torch.int32, device=diag.device)
    ds, di = diag.stride(0), diag.stride(1)
    ls, li = _blk_strides(sub, m)
    us, ui = _blk_strides(sup, m)
    _check(LIB.slab_btri_inv_factor(_dt(diag), b, m, bs, _p(diag), ds, di, _p(sub), ls, li, _p(sup), us, ui, _p(info),
                                    _p(ws), _p(iws), _stream()), "slab_btri_inv_factor")


def btri_inv_solve(diag, sub, sup, f):
    """In-place sweeps on f (b, m*bs, nrhs) with the factors of btri_inv_factor."""
    b, m, bs, _ = diag.shape
    _require(f.dim() == 3 and f.shape[0] == b and f.shape[1] == m * bs and f.stride(2) == 1, "btri rhs layout")
    _require(f.dtype == diag.dtype, "btri rhs dtype")
    nrhs = f.shape[2]
    ws = None
    if not btri_inv_fused(diag.dtype, bs):
        ws = torch.empty(b * bs * max(nrhs, 1), dtype=diag.dtype, device=diag.device)
    ds, di = diag.stride(0), diag.stride(1)
    ls, li = _blk_strides(sub, m)
    us, ui = _blk_strides(sup, m)
    _check(LIB.slab_btri_inv_solve(_dt(diag), b, m, bs, _p(diag), ds, di, _p(sub), ls, li, _p(sup), us, ui, _p(f),
                                   f.stride(0), f.stride(1), nrhs, _p(ws), _stream()), "slab_btri_inv_solve")
    return f


def set_sweep_tile(dtype, bs, inv_factors, nr):
    """Right-hand-side tile width of the chain sweeps for blocks of size bs
    (0: built-in rule); returns the previous width."""
    rc = LIB.slab_set_sweep_tile(F64 if dtype == torch.float64 else F32, int(bs), int(bool(inv_factors)), int(nr))
    if rc < 0:
        raise ValueError("slab_set_sweep_tile(%s, %d, %d, %d) rejected" % (dtype, bs, inv_factors, nr))
    return rc


# most vectors one slab_mdot / slab_maxpy launch takes per column (their
# per-thread partial sums are register arrays of this length); more run as
# several launches
KRYLOV_MAX = 32


def mdot(V, w, out=None):
    """out[c, i] = V[c, i, :] . w[c, :] for all columns c at once (FP64):
    V (nc, k, n) and w (nc, n) with unit stride along n; out (nc, k)."""
    nc, k, n = V.shape
    _require(V.dtype == torch.float64 and w.dtype == torch.float64 and V.stride(2) == 1 and w.stride(-1) == 1 and
             tuple(w.shape) == (nc, n) and k >= 1, "mdot layouts")
    if out is None:
        out = torch.empty(nc, k, dtype=torch.float64, device=V.device)
    _require(out.stride(-1) == 1 and tuple(out.shape) == (nc, k), "mdot output layout")
    if k > KRYLOV_MAX:
        for c0 in range(0, k, KRYLOV_MAX):
            mdot(V[:, c0:c0 + KRYLOV_MAX], w, out=out[:, c0:c0 + KRYLOV_MAX])
        return out
    ws = torch.empty(max(1, int(LIB.slab_mdot_ws(nc, k, n)) // 8), dtype=torch.float64, device=V.device)
    _check(LIB.slab_mdot(nc, k, n, _p(V), V.stride(0), V.stride(1), _p(w), w.stride(0), _p(out), out.stride(0),
                         _p(ws), _stream()), "slab_mdot")
    return out


def maxpy(V, coef, w, alpha=1.0):
    """w[c, :] += alpha sum_i coef[c, i] V[c, i, :] for all columns c (FP64)."""
    nc, k, n = V.shape
    _require(V.dtype == torch.float64 and w.dtype == torch.float64 and coef.dtype == torch.float64 and
             V.stride(2) == 1 and w.stride(-1) == 1 and coef.stride(-1) == 1 and tuple(w.shape) == (nc, n) and
             tuple(coef.shape) == (nc, k) and k >= 1, "maxpy layouts")
    if k > KRYLOV_MAX:
        for c0 in range(0, k, KRYLOV_MAX):
            maxpy(V[:, c0:c0 + KRYLOV_MAX], coef[:, c0:c0 + KRYLOV_MAX], w, alpha)
        return w
    _check(LIB.slab_maxpy(nc, k, n, float(alpha), _p(V), V.stride(0), V.stride(1), _p(coef), coef.stride(0), _p(w),
                          w.stride(0), _stream()), "slab_maxpy")
    return w


def leaf_assemble(M, p, border, D, Dsq, coefs, Nxtot):
    """Fill M (nboxes, nrows, >= n_i + n_x) with the condensation rows of every
    leaf; coefs: dict of FP64 (nboxes*p*p) tensors for c11, c22 and optionally
    c12, c1, c2, c."""
    _require(M.dim() == 3 and M.stride(2) == 1, "leaf M layout")
    for k in ("c11", "c22"):
        _require(coefs.get(k) is not None, "coefficient %s required" % k)
    args = [coefs.get(k) for k in ("c11", "c22", "c12", "c1", "c2", "c")]
    for a in args:
        _require(a is None or (a.dtype == torch.float64 and a.is_contiguous() and a.numel() == M.shape[0] * p * p),
                 "coefficient arrays must be contiguous FP64 of nboxes*p*p")
    _check(LIB.slab_leaf_assemble(_dt(M), M.shape[0], p, int(border), _p(D), _p(Dsq), *[_p(a) for a in args],
                                  _p(Nxtot), _p(M), M.stride(1), M.stride(0), _stream()), "slab_leaf_assemble")
    return M


def leaf_assemble_nodes(M, p, xnodes, D, Dsq, coefs):
    """Fill M (nboxes, (p-2)^2, >= (p-2)^2 + len(xnodes)) with [A_cc | A_c,X]
    of every leaf, the X columns at the leaf nodes xnodes (int32 device
    tensor of node ids i*p + j)."""
    _require(M.dim() == 3 and M.stride(2) == 1, "leaf M layout")
    _require(xnodes.dtype == torch.int32 and xnodes.is_contiguous(), "xnodes must be contiguous int32")
    for k in ("c11", "c22"):
        _require(coefs.get(k) is not None, "coefficient %s required" % k)
    args = [coefs.get(k) for k in ("c11", "c22", "c12", "c1", "c2", "c")]
    for a in args:
        _require(a is None or (a.dtype == torch.float64 and a.is_contiguous() and a.numel() == M.shape[0] * p * p),
                 "coefficient arrays must be contiguous FP64 of nboxes*p*p")
    _check(LIB.slab_leaf_assemble_nodes(_dt(M), M.shape[0], p, xnodes.numel(), _p(xnodes), _p(D), _p(Dsq),
                                        *[_p(a) for a in args], _p(M), M.stride(1), M.stride(0), _stream()),
           "slab_leaf_assemble_nodes")
    return M


def leaf_residual(uc, ub, tables, coefs, p, uu=None):
    """FP64 leaf residuals of the un-condensed system for all leaves: uc
    (nboxes, (p-2)^2, nrhs) interior values, ub (nboxes, 4(p-2), nrhs) face
    values -> (rI = -(A_loc u) at the interior nodes, fl = N_x u); uu
    (nboxes, p^2, nrhs), if given, receives all node values.  tables: dict
    with Intmap, Jc, Jxreo, D, Dsq, nx_idx, nx_val (device)."""
    nb, ni, nrhs = uc.shape
    f = p - 2
    _require(uc.dtype == torch.float64 and ub.dtype == torch.float64, "leaf_residual works in FP64")
    _require(ni == f * f and ub.shape == (nb, 4 * f, nrhs), "leaf_residual shapes")
    _require(uc.stride(2) == 1 and ub.stride(2) == 1, "leaf_residual layouts")
    for k in ("c11", "c22"):
        _require(coefs.get(k) is not None, "coefficient %s required" % k)
    args = [coefs.get(k) for k in ("c11", "c22", "c12", "c1", "c2", "c")]
    for a in args:
        _require(a is None or (a.dtype == torch.float64 and a.is_contiguous() and a.numel() == nb * p * p),
                 "coefficient arrays must be contiguous FP64 of nboxes*p*p")
    rI = torch.empty(nb, ni, nrhs, dtype=torch.float64, device=uc.device)
    fl = torch.empty(nb, 4 * f, nrhs, dtype=torch.float64, device=uc.device)
    if uu is not None:
        _require(uu.shape == (nb, p * p, nrhs) and uu.stride(2) == 1 and uu.dtype == torch.float64, "uu layout")
    t = tables
    _check(LIB.slab_leaf_residual(nb, p, nrhs, _p(uc), uc.stride(1), uc.stride(0), _p(ub), ub.stride(1),
                                  ub.stride(0), _p(t["Intmap"]), _p(t["Jc"]), _p(t["Jxreo"]), _p(t["D"]),
                                  _p(t["Dsq"]), *[_p(a) for a in args], _p(t["nx_idx"]), _p(t["nx_val"]), _p(rI),
                                  rI.stride(1), rI.stride(0), _p(fl), fl.stride(1), fl.stride(0), _p(uu),
                                  uu.stride(1) if uu is not None else 0, uu.stride(0) if uu is not None else 0,
                                  _stream()), "slab_leaf_residual")
    return rI, fl


PROBE_GEMM_F32, PROBE_GEMM_F64 = 1, 2


def probe_start(kernel, max_launches=4096):
    """Time every launch of one kernel (PROBE_GEMM_F32: k_gemm<float, int,
    true, true>; PROBE_GEMM_F64: k_gemm<double, int, false, false>) with HIP
    events on the stream it runs on, until probe_stop."""
    _check(LIB.slab_probe_start(int(kernel), int(max_launches)), "slab_probe_start")


def probe_stop():
    """dict(launches, dropped, total_ms, flops, bytes) of the probed launches
    (waits for their events)."""
    n, d = _int(0), _int(0)
    ms, fl, by = _dbl(0), _dbl(0), _dbl(0)
    _check(LIB.slab_probe_stop(ctypes.byref(n), ctypes.byref(d), ctypes.byref(ms), ctypes.byref(fl),
                               ctypes.byref(by)), "slab_probe_stop")
    return dict(launches=n.value, dropped=d.value, total_ms=ms.value, flops=fl.value, bytes=by.value)
