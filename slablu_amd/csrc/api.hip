// extern "C" entry points of libslabhip (declared in include/slab_hip.h)
// and the block-tridiagonal drivers built on the batched kernels.
#include "slab_internal.h"
#include "../../include/slab_hip.h"

using namespace slab;

namespace {

inline int status(hipError_t e) { return e == hipSuccess ? 0 : -(1000 + (int)e); }
inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

template <typename T>
View<T> blk(void* base, int64_t bstride, int64_t istride, int i, int64_t ld) {
  return View<T>{reinterpret_cast<T*>(base) + (int64_t)i * istride, ld, 1, bstride};
}

template <typename T>
hipError_t btri_factor(int b, int m, int bs, void* diag, int64_t dbs, int64_t dis, void* sub, int64_t lbs,
                       int64_t lis, void* sup, int64_t ubs, int64_t uis, int* ipiv, int64_t pbs, int64_t pis,
                       int* info, hipStream_t s) {
  if (btri_lu_fused_ok<T>(bs))
    return btri_lu_factor_fused<T>(b, m, bs, (T*)diag, dbs, dis, (T*)sub, lbs, lis, (T*)sup, ubs, uis, ipiv, pbs,
                                   pis, info, s);
  hipError_t e;
  e = getrf<T>(b, bs, bs, bs, bs, blk<T>(diag, dbs, dis, 0, bs), ipiv, pbs, info, s);
  if (e != hipSuccess) return e;
  for (int i = 1; i < m; ++i) {
    View<T> Dp = blk<T>(diag, dbs, dis, i - 1, bs);
    View<T> Up = blk<T>(sup, ubs, uis, i - 1, bs);
    View<T> Lp = blk<T>(sub, lbs, lis, i - 1, bs);
    const int* pv = ipiv + (int64_t)(i - 1) * pis;
    // Ufac = L_{i-1}^{-1} P^T sup_{i-1}
    if ((e = laswp<T>(b, bs, 0, bs, true, Up, pv, pbs, s)) != hipSuccess) return e;
    if ((e = trsm_left<T>(b, bs, bs, true, true, Dp, Up, s)) != hipSuccess) return e;
    // Lfac = sub_{i-1} U_{i-1}^{-1}  <=>  U^T Lfac^T = sub^T
    if ((e = trsm_left<T>(b, bs, bs, true, false, Dp.t(), Lp.t(), s)) != hipSuccess) return e;
    View<T> Di = blk<T>(diag, dbs, dis, i, bs);
    if ((e = gemm<T>(b, bs, bs, bs, T(-1), Lp, Up, T(1), Di, s)) != hipSuccess) return e;
    if ((e = getrf<T>(b, bs, bs, bs, bs, Di, ipiv + (int64_t)i * pis, pbs, info, s)) != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename T>
hipError_t btri_solve(int b, int m, int bs, const void* diag, int64_t dbs, int64_t dis, const void* sub,
                      int64_t lbs, int64_t lis, const void* sup, int64_t ubs, int64_t uis, const int* ipiv,
                      int64_t pbs, int64_t pis, void* f, int64_t fbs, int64_t ldf, int nrhs, bool trans,
                      hipStream_t s) {
  auto D = [&](int i) { return blk<T>(const_cast<void*>(diag), dbs, dis, i, bs); };
  auto L = [&](int i) { return blk<T>(const_cast<void*>(sub), lbs, lis, i, bs); };
  auto U = [&](int i) { return blk<T>(const_cast<void*>(sup), ubs, uis, i, bs); };
  auto F = [&](int i) { return View<T>{reinterpret_cast<T*>(f) + (int64_t)i * bs * ldf, ldf, 1, fbs}; };
  auto P = [&](int i) { return ipiv + (int64_t)i * pis; };
  hipError_t e;
  if (btri_lu_fused_ok<T>(bs))
    return btri_lu_solve_fused<T>(b, m, bs, (T*)diag, dbs, dis, (T*)sub, lbs, lis, (T*)sup, ubs, uis, ipiv, pbs,
                                  pis, (T*)f, fbs, ldf, nrhs, s, trans);
  if (!trans) {
    for (int i = 0; i < m; ++i) {
      if (i > 0 && (e = gemm<T>(b, bs, nrhs, bs, T(-1), L(i - 1), F(i - 1), T(1), F(i), s)) != hipSuccess) return e;
      if ((e = laswp<T>(b, nrhs, 0, bs, true, F(i), P(i), pbs, s)) != hipSuccess) return e;
      if ((e = trsm_left<T>(b, bs, nrhs, true, true, D(i), F(i), s)) != hipSuccess) return e;
    }
    for (int i = m - 1; i >= 0; --i) {
      if (i < m - 1 && (e = gemm<T>(b, bs, nrhs, bs, T(-1), U(i), F(i + 1), T(1), F(i), s)) != hipSuccess) return e;
      if ((e = trsm_left<T>(b, bs, nrhs, false, false, D(i), F(i), s)) != hipSuccess) return e;
    }
  } else {
    // U_blk^T y = f : U_i^T y_i = f_i - Ufac_{i-1}^T y_{i-1}
    for (int i = 0; i < m; ++i) {
      if (i > 0 && (e = gemm<T>(b, bs, nrhs, bs, T(-1), U(i - 1).t(), F(i - 1), T(1), F(i), s)) != hipSuccess)
        return e;
      if ((e = trsm_left<T>(b, bs, nrhs, true, false, D(i).t(), F(i), s)) != hipSuccess) return e;
    }
    // L_blk^T x = y : (P_i L_i)^T x_i = y_i - Lfac_i^T x_{i+1}
    for (int i = m - 1; i >= 0; --i) {
      if (i < m - 1 && (e = gemm<T>(b, bs, nrhs, bs, T(-1), L(i).t(), F(i + 1), T(1), F(i), s)) != hipSuccess)
        return e;
      if ((e = trsm_left<T>(b, bs, nrhs, false, true, D(i).t(), F(i), s)) != hipSuccess) return e;
      if ((e = laswp<T>(b, nrhs, 0, bs, false, F(i), P(i), pbs, s)) != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

}  // namespace

extern "C" {

int slab_version(void) { return 100; }

int slab_gemm(int dtype, int batch, int m, int n, int k, double alpha, const void* A, int64_t a_rs, int64_t a_cs,
              int64_t a_bs, const void* B, int64_t b_rs, int64_t b_cs, int64_t b_bs, double beta, void* C,
              int64_t c_rs, int64_t c_cs, int64_t c_bs, void* stream) {
  if (batch < 0 || m < 0 || n < 0 || k < 0) return -1;
  if (dtype == SLAB_F64)
    return status(gemm<double>(batch, m, n, k, alpha, view<double>(const_cast<void*>(A), a_rs, a_cs, a_bs),
                               view<double>(const_cast<void*>(B), b_rs, b_cs, b_bs), beta,
                               view<double>(C, c_rs, c_cs, c_bs), S(stream)));
  if (dtype == SLAB_F32)
    return status(gemm<float>(batch, m, n, k, (float)alpha, view<float>(const_cast<void*>(A), a_rs, a_cs, a_bs),
                              view<float>(const_cast<void*>(B), b_rs, b_cs, b_bs), (float)beta,
                              view<float>(C, c_rs, c_cs, c_bs), S(stream)));
  return -1;
}

int slab_trsm(int dtype, int batch, int m, int n, int lower, int unit, const void* A, int64_t a_rs, int64_t a_cs,
              int64_t a_bs, void* B, int64_t b_rs, int64_t b_cs, int64_t b_bs, void* stream) {
  if (batch < 0 || m < 0 || n < 0) return -1;
  if (dtype == SLAB_F64)
    return status(trsm_left<double>(batch, m, n, lower != 0, unit != 0,
                                     view<double>(const_cast<void*>(A), a_rs, a_cs, a_bs),
                                     view<double>(B, b_rs, b_cs, b_bs), S(stream)));
  if (dtype == SLAB_F32)
    return status(trsm_left<float>(batch, m, n, lower != 0, unit != 0,
                                    view<float>(const_cast<void*>(A), a_rs, a_cs, a_bs),
                                    view<float>(B, b_rs, b_cs, b_bs), S(stream)));
  return -1;
}

int slab_laswp(int dtype, int batch, int ncols, int k1, int k2, int forward, void* A, int64_t a_rs, int64_t a_cs,
               int64_t a_bs, const int* ipiv, int64_t p_bs, void* stream) {
  if (batch < 0 || ncols < 0 || k1 < 0 || k2 < k1) return -1;
  if (dtype == SLAB_F64)
    return status(laswp<double>(batch, ncols, k1, k2, forward != 0, view<double>(A, a_rs, a_cs, a_bs), ipiv, p_bs,
                                S(stream)));
  if (dtype == SLAB_F32)
    return status(laswp<float>(batch, ncols, k1, k2, forward != 0, view<float>(A, a_rs, a_cs, a_bs), ipiv, p_bs,
                               S(stream)));
  return -1;
}

int slab_getrf(int dtype, int batch, int nrows, int ncols, int nelim, int npiv, void* A, int64_t lda, int64_t a_bs,
               int* ipiv, int64_t p_bs, int* info, void* stream) {
  if (batch < 0 || nelim < 0 || nelim > npiv || npiv > nrows || nelim > ncols || lda < ncols) return -1;
  if (dtype == SLAB_F64)
    return status(getrf<double>(batch, nrows, ncols, nelim, npiv, view<double>(A, lda, 1, a_bs), ipiv, p_bs, info,
                                S(stream)));
  if (dtype == SLAB_F32)
    return status(getrf<float>(batch, nrows, ncols, nelim, npiv, view<float>(A, lda, 1, a_bs), ipiv, p_bs, info,
                               S(stream)));
  return -1;
}

int slab_piv_to_perm(int batch, int n, const int* ipiv, int64_t p_bs, int* perm, int64_t q_bs, void* stream) {
  if (batch < 0 || n < 0) return -1;
  return status(piv_to_perm_batched(batch, n, ipiv, p_bs, perm, q_bs, S(stream)));
}

size_t slab_getri_ws(int dtype, int batch, int n) {
  if (batch < 0 || n < 0 || (dtype != SLAB_F64 && dtype != SLAB_F32)) return 0;
  return getri_ws_bytes(batch, n, dtype == SLAB_F64 ? sizeof(double) : sizeof(float));
}

int slab_getri(int dtype, int batch, int n, const void* LU, int64_t lda, int64_t lu_bs, const int* ipiv,
               int64_t p_bs, void* X, int64_t ldx, int64_t x_bs, void* ws, void* stream) {
  if (batch < 0 || n < 0 || (batch > 0 && n > 0 && ws == nullptr)) return -1;
  if (dtype == SLAB_F64)
    return status(getri<double>(batch, n, view<double>(const_cast<void*>(LU), lda, 1, lu_bs), ipiv, p_bs,
                                view<double>(X, ldx, 1, x_bs), ws, S(stream)));
  if (dtype == SLAB_F32)
    return status(getri<float>(batch, n, view<float>(const_cast<void*>(LU), lda, 1, lu_bs), ipiv, p_bs,
                               view<float>(X, ldx, 1, x_bs), ws, S(stream)));
  return -1;
}

int slab_orth(int dtype, int batch, int m, int k, void* A, int64_t lda, int64_t a_bs, void* Q, int64_t ldq,
              int64_t q_bs, void* stream) {
  if (batch < 0 || k < 0 || m < k || k > 256) return -1;
  if (dtype == SLAB_F64)
    return status(house_orth<double>(batch, m, k, view<double>(A, lda, 1, a_bs), view<double>(Q, ldq, 1, q_bs),
                                     S(stream)));
  if (dtype == SLAB_F32)
    return status(house_orth<float>(batch, m, k, view<float>(A, lda, 1, a_bs), view<float>(Q, ldq, 1, q_bs),
                                    S(stream)));
  return -1;
}

int slab_getrs_perm(int dtype, int batch, int n, int nrhs, const void* LU, int64_t lda, int64_t lu_bs,
                    const int* perm, int64_t q_bs, void* B, int64_t ldb, int64_t b_bs, void* stream) {
  if (batch < 0 || n < 1 || nrhs < 1 || nrhs > 8) return -1;
  if (batch == 0) return 0;
  if (dtype == SLAB_F64)
    return status(getrs_perm<double>(batch, n, nrhs, view<double>(const_cast<void*>(LU), lda, 1, lu_bs), perm, q_bs,
                                     view<double>(B, ldb, 1, b_bs), S(stream)));
  if (dtype == SLAB_F32)
    return status(getrs_perm<float>(batch, n, nrhs, view<float>(const_cast<void*>(LU), lda, 1, lu_bs), perm, q_bs,
                                    view<float>(B, ldb, 1, b_bs), S(stream)));
  return -1;
}

int slab_getrs(int dtype, int batch, int n, int nrhs, int trans, const void* LU, int64_t lda, int64_t lu_bs,
               const int* ipiv, int64_t p_bs, void* B, int64_t ldb, int64_t b_bs, void* stream) {
  if (batch < 0 || n < 0 || nrhs < 0) return -1;
  if (dtype == SLAB_F64)
    return status(getrs<double>(batch, n, nrhs, trans != 0, view<double>(const_cast<void*>(LU), lda, 1, lu_bs), ipiv,
                                p_bs, view<double>(B, ldb, 1, b_bs), S(stream)));
  if (dtype == SLAB_F32)
    return status(getrs<float>(batch, n, nrhs, trans != 0, view<float>(const_cast<void*>(LU), lda, 1, lu_bs), ipiv,
                               p_bs, view<float>(B, ldb, 1, b_bs), S(stream)));
  return -1;
}

int slab_btri_factor(int dtype, int b, int m, int bs, void* diag, int64_t d_bstride, int64_t d_istride, void* sub,
                     int64_t l_bstride, int64_t l_istride, void* sup, int64_t u_bstride, int64_t u_istride,
                     int* ipiv, int64_t p_bstride, int64_t p_istride, int* info, void* stream) {
  if (b < 0 || m < 1 || bs < 1) return -1;
  if (dtype == SLAB_F64)
    return status(btri_factor<double>(b, m, bs, diag, d_bstride, d_istride, sub, l_bstride, l_istride, sup,
                                      u_bstride, u_istride, ipiv, p_bstride, p_istride, info, S(stream)));
  if (dtype == SLAB_F32)
    return status(btri_factor<float>(b, m, bs, diag, d_bstride, d_istride, sub, l_bstride, l_istride, sup,
                                     u_bstride, u_istride, ipiv, p_bstride, p_istride, info, S(stream)));
  return -1;
}

int slab_btri_solve(int dtype, int b, int m, int bs, const void* diag, int64_t d_bstride, int64_t d_istride,
                    const void* sub, int64_t l_bstride, int64_t l_istride, const void* sup, int64_t u_bstride,
                    int64_t u_istride, const int* ipiv, int64_t p_bstride, int64_t p_istride, void* f,
                    int64_t f_bstride, int64_t ldf, int nrhs, int transpose, void* stream) {
  if (b < 0 || m < 1 || bs < 1 || nrhs < 0 || ldf < nrhs) return -1;
  if (dtype == SLAB_F64)
    return status(btri_solve<double>(b, m, bs, diag, d_bstride, d_istride, sub, l_bstride, l_istride, sup,
                                     u_bstride, u_istride, ipiv, p_bstride, p_istride, f, f_bstride, ldf, nrhs,
                                     transpose != 0, S(stream)));
  if (dtype == SLAB_F32)
    return status(btri_solve<float>(b, m, bs, diag, d_bstride, d_istride, sub, l_bstride, l_istride, sup, u_bstride,
                                    u_istride, ipiv, p_bstride, p_istride, f, f_bstride, ldf, nrhs, transpose != 0,
                                    S(stream)));
  return -1;
}

int slab_btri_fused(int dtype, int bs) {
  if (bs < 1) return 0;
  if (dtype == SLAB_F64) return btri_lu_fused_ok<double>(bs) ? 1 : 0;
  if (dtype == SLAB_F32) return btri_lu_fused_ok<float>(bs) ? 1 : 0;
  return -1;
}

int slab_btri_factor_twisted(int dtype, int b, int m, int bs, void* diag, int64_t d_bstride, int64_t d_istride,
                             void* sub, int64_t l_bstride, int64_t l_istride, void* sup, int64_t u_bstride,
                             int64_t u_istride, int* ipiv, int64_t p_bstride, int64_t p_istride, int* info,
                             void* stream) {
  if (b < 0 || m < 1 || bs < 1 || slab_btri_fused(dtype, bs) != 1) return -1;
  if (dtype == SLAB_F64)
    return status(btri_lu_factor_twisted<double>(b, m, bs, (double*)diag, d_bstride, d_istride, (double*)sub,
                                                 l_bstride, l_istride, (double*)sup, u_bstride, u_istride, ipiv,
                                                 p_bstride, p_istride, info, S(stream)));
  return status(btri_lu_factor_twisted<float>(b, m, bs, (float*)diag, d_bstride, d_istride, (float*)sub, l_bstride,
                                              l_istride, (float*)sup, u_bstride, u_istride, ipiv, p_bstride,
                                              p_istride, info, S(stream)));
}

int slab_btri_solve_twisted(int dtype, int b, int m, int bs, const void* diag, int64_t d_bstride,
                            int64_t d_istride, const void* sub, int64_t l_bstride, int64_t l_istride,
                            const void* sup, int64_t u_bstride, int64_t u_istride, const int* ipiv,
                            int64_t p_bstride, int64_t p_istride, void* f, int64_t f_bstride, int64_t ldf, int nrhs,
                            int transpose, void* stream) {
  if (b < 0 || m < 1 || bs < 1 || nrhs < 0 || ldf < nrhs || slab_btri_fused(dtype, bs) != 1) return -1;
  if (nrhs == 0 || b == 0) return 0;
  if (dtype == SLAB_F64)
    return status(btri_lu_solve_twisted<double>(b, m, bs, (double*)diag, d_bstride, d_istride, (double*)sub,
                                                l_bstride, l_istride, (double*)sup, u_bstride, u_istride, ipiv,
                                                p_bstride, p_istride, (double*)f, f_bstride, ldf, nrhs, S(stream),
                                                nullptr, transpose != 0));
  return status(btri_lu_solve_twisted<float>(b, m, bs, (float*)diag, d_bstride, d_istride, (float*)sub, l_bstride,
                                             l_istride, (float*)sup, u_bstride, u_istride, ipiv, p_bstride,
                                             p_istride, (float*)f, f_bstride, ldf, nrhs, S(stream), nullptr,
                                             transpose != 0));
}

int slab_btri_invert_diag(int dtype, int b, int m, int bs, void* diag, int64_t d_bstride, int64_t d_istride,
                          const int* ipiv, int64_t p_bstride, int64_t p_istride, void* linv, void* ws,
                          void* stream) {
  if (b < 0 || m < 1 || bs < 1) return -1;
  // all blocks must form one batch of contiguous bs x bs blocks
  if (d_istride != (int64_t)bs * bs || (b > 1 && d_bstride != (int64_t)m * bs * bs)) return -1;
  if (p_istride != bs || (b > 1 && p_bstride != (int64_t)m * bs)) return -1;
  if (b == 0) return 0;
  if (linv == nullptr || ws == nullptr || (int64_t)b * m > 32767) return -1;
  const int nbk = b * m;
  const int64_t bb = (int64_t)bs * bs;
  if (dtype == SLAB_F64)
    return status(getri_factors<double>(nbk, bs, view<double>(diag, bs, 1, bb), ipiv, bs, view<double>(linv, bs, 1, bb),
                                        view<double>(diag, bs, 1, bb), ws, S(stream)));
  if (dtype == SLAB_F32)
    return status(getri_factors<float>(nbk, bs, view<float>(diag, bs, 1, bb), ipiv, bs, view<float>(linv, bs, 1, bb),
                                       view<float>(diag, bs, 1, bb), ws, S(stream)));
  return -1;
}

int slab_btri_solve_twisted_inv(int dtype, int b, int m, int bs, const void* diag, int64_t d_bstride,
                                int64_t d_istride, const void* linv, const void* sub, int64_t l_bstride,
                                int64_t l_istride, const void* sup, int64_t u_bstride, int64_t u_istride, void* f,
                                int64_t f_bstride, int64_t ldf, int nrhs, int transpose, void* stream) {
  if (b < 0 || m < 1 || bs < 1 || nrhs < 0 || ldf < nrhs || linv == nullptr || slab_btri_fused(dtype, bs) != 1)
    return -1;
  if (nrhs == 0 || b == 0) return 0;
  if (dtype == SLAB_F64)
    return status(btri_lu_solve_twisted<double>(b, m, bs, (double*)diag, d_bstride, d_istride, (double*)sub,
                                                l_bstride, l_istride, (double*)sup, u_bstride, u_istride, nullptr, 0,
                                                0, (double*)f, f_bstride, ldf, nrhs, S(stream), (const double*)linv,
                                                transpose != 0));
  return status(btri_lu_solve_twisted<float>(b, m, bs, (float*)diag, d_bstride, d_istride, (float*)sub, l_bstride,
                                             l_istride, (float*)sup, u_bstride, u_istride, nullptr, 0, 0, (float*)f,
                                             f_bstride, ldf, nrhs, S(stream), (const float*)linv, transpose != 0));
}

int slab_btri_inv_fused(int dtype, int bs) {
  if (dtype == SLAB_F64) return btri_inv_fused_ok<double>(bs) ? 1 : 0;
  if (dtype == SLAB_F32) return btri_inv_fused_ok<float>(bs) ? 1 : 0;
  return -1;
}

int slab_btri_inv_factor(int dtype, int b, int m, int bs, void* diag, int64_t d_bstride, int64_t d_istride,
                         void* sub, int64_t l_bstride, int64_t l_istride, void* sup, int64_t u_bstride,
                         int64_t u_istride, int* info, void* ws, int* iws, void* stream) {
  if (b < 0 || m < 1 || bs < 1) return -1;
  if (dtype == SLAB_F64)
    return status(btri_inv_factor<double>(b, m, bs, (double*)diag, d_bstride, d_istride, (double*)sub, l_bstride,
                                          l_istride, (double*)sup, u_bstride, u_istride, info, (double*)ws, iws,
                                          S(stream)));
  if (dtype == SLAB_F32)
    return status(btri_inv_factor<float>(b, m, bs, (float*)diag, d_bstride, d_istride, (float*)sub, l_bstride,
                                         l_istride, (float*)sup, u_bstride, u_istride, info, (float*)ws, iws,
                                         S(stream)));
  return -1;
}

int slab_btri_inv_solve(int dtype, int b, int m, int bs, const void* diag, int64_t d_bstride, int64_t d_istride,
                        const void* sub, int64_t l_bstride, int64_t l_istride, const void* sup, int64_t u_bstride,
                        int64_t u_istride, void* f, int64_t f_bstride, int64_t ldf, int nrhs, void* ws,
                        void* stream) {
  if (b < 0 || m < 1 || bs < 1 || nrhs < 0 || ldf < nrhs) return -1;
  if (nrhs == 0) return 0;
  if (dtype == SLAB_F64)
    return status(btri_inv_solve<double>(b, m, bs, (double*)diag, d_bstride, d_istride, (double*)sub, l_bstride,
                                         l_istride, (double*)sup, u_bstride, u_istride, (double*)f, f_bstride, ldf,
                                         nrhs, (double*)ws, S(stream)));
  if (dtype == SLAB_F32)
    return status(btri_inv_solve<float>(b, m, bs, (float*)diag, d_bstride, d_istride, (float*)sub, l_bstride,
                                        l_istride, (float*)sup, u_bstride, u_istride, (float*)f, f_bstride, ldf,
                                        nrhs, (float*)ws, S(stream)));
  return -1;
}

int slab_leaf_assemble(int dtype, int nboxes, int p, int border, const double* D, const double* Dsq,
                       const double* c11, const double* c22, const double* c12, const double* c1, const double* c2,
                       const double* c, const double* Nxtot, void* M, int64_t ldm, int64_t m_bs, void* stream) {
  if (nboxes < 0 || p < 3 || !D || !Dsq || !c11 || !c22 || (border && !Nxtot)) return -1;
  const int f = p - 2;
  if (ldm < (int64_t)f * f + 4 * f) return -1;
  if (dtype == SLAB_F64)
    return status(leaf_assemble<double>(nboxes, p, border, D, Dsq, c11, c22, c12, c1, c2, c, Nxtot,
                                        reinterpret_cast<double*>(M), ldm, m_bs, S(stream)));
  if (dtype == SLAB_F32)
    return status(leaf_assemble<float>(nboxes, p, border, D, Dsq, c11, c22, c12, c1, c2, c, Nxtot,
                                       reinterpret_cast<float*>(M), ldm, m_bs, S(stream)));
  return -1;
}

int slab_leaf_assemble_nodes(int dtype, int nboxes, int p, int nxc, const int* xnodes, const double* D,
                             const double* Dsq, const double* c11, const double* c22, const double* c12,
                             const double* c1, const double* c2, const double* c, void* M, int64_t ldm, int64_t m_bs,
                             void* stream) {
  if (nboxes < 0 || p < 3 || nxc < 1 || nxc > p * p || !xnodes || !D || !Dsq || !c11 || !c22) return -1;
  const int f = p - 2;
  if (ldm < (int64_t)f * f + nxc || m_bs < (int64_t)f * f * ldm) return -1;
  if (dtype == SLAB_F64)
    return status(leaf_assemble<double>(nboxes, p, 0, D, Dsq, c11, c22, c12, c1, c2, c, nullptr,
                                        reinterpret_cast<double*>(M), ldm, m_bs, S(stream), xnodes, nxc));
  if (dtype == SLAB_F32)
    return status(leaf_assemble<float>(nboxes, p, 0, D, Dsq, c11, c22, c12, c1, c2, c, nullptr,
                                       reinterpret_cast<float*>(M), ldm, m_bs, S(stream), xnodes, nxc));
  return -1;
}

int slab_set_sweep_tile(int dtype, int bs, int inv_factors, int nr) {
  return set_sweep_tile(dtype, bs, inv_factors, nr);
}

size_t slab_mdot_ws(int nc, int k, int64_t n) {
  if (nc < 0 || k < 0 || n < 0) return 0;
  return mdot_ws_elems(nc, k, n) * sizeof(double);
}

int slab_mdot(int nc, int k, int64_t n, const double* V, int64_t v_cs, int64_t v_ks, const double* w, int64_t w_cs,
              double* out, int64_t out_cs, void* ws, void* stream) {
  if (nc < 0 || k < 0 || k > 32 || n < 0 || (nc > 0 && k > 0 && (!V || !w || !out || !ws))) return -1;
  return status(mdot(nc, k, n, V, v_cs, v_ks, w, w_cs, out, out_cs, reinterpret_cast<double*>(ws), S(stream)));
}

int slab_maxpy(int nc, int k, int64_t n, double alpha, const double* V, int64_t v_cs, int64_t v_ks,
               const double* coef, int64_t coef_cs, double* w, int64_t w_cs, void* stream) {
  if (nc < 0 || k < 0 || k > 32 || n < 0 || (nc > 0 && k > 0 && n > 0 && (!V || !coef || !w))) return -1;
  return status(maxpy(nc, k, n, alpha, V, v_cs, v_ks, coef, coef_cs, w, w_cs, S(stream)));
}

int slab_leaf_residual(int nboxes, int p, int nrhs, const double* uc, int64_t uc_ld, int64_t uc_bs,
                       const double* ub, int64_t ub_ld, int64_t ub_bs, const double* Intmap, const int* Jc,
                       const int* Jxreo, const double* D, const double* Dsq, const double* c11, const double* c22,
                       const double* c12, const double* c1, const double* c2, const double* c, const int* nx_idx,
                       const double* nx_val, double* rI, int64_t rI_ld, int64_t rI_bs, double* fl, int64_t fl_ld,
                       int64_t fl_bs, double* uu, int64_t uu_ld, int64_t uu_bs, void* stream) {
  if (nboxes < 0 || nrhs < 0 || p < 3 || p > 64) return -1;
  if (!uc || !ub || !Intmap || !Jc || !Jxreo || !D || !Dsq || !c11 || !c22 || !nx_idx || !nx_val || !rI || !fl)
    return -1;
  const int f = p - 2;
  if (uc_ld < nrhs || ub_ld < nrhs || rI_ld < nrhs || fl_ld < nrhs || (uu && uu_ld < nrhs)) return -1;
  if (uc_bs < (int64_t)f * f * uc_ld || ub_bs < (int64_t)4 * f * ub_ld || rI_bs < (int64_t)f * f * rI_ld ||
      fl_bs < (int64_t)4 * f * fl_ld || (uu && uu_bs < (int64_t)p * p * uu_ld))
    return -1;
  return status(leaf_residual(nboxes, p, nrhs, uc, uc_ld, uc_bs, ub, ub_ld, ub_bs, Intmap, Jc, Jxreo, D, Dsq, c11,
                              c22, c12, c1, c2, c, nx_idx, nx_val, rI, rI_ld, rI_bs, fl, fl_ld, fl_bs, uu, uu_ld,
                              uu_bs, S(stream)));
}

}  // extern "C"
