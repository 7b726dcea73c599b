// Block-tridiagonal LU (build_sweep of src_solver/block_tridiag_solver_dense.py
// :152-177): one workgroup per chain runs the whole factorization with each
// diagonal block resident in LDS.
#include "lds_dev.h"

namespace slab {

namespace {

// One workgroup per chain for the whole factorization.  Per block: stage
// D_i, subtract Lfac_{i-1} Ufac_{i-1} (products from global in KC-wide LDS
// chunks), blocked LU with partial pivoting in LDS, then Ufac_i (rows gathered
// through the composed pivot permutation) and Lfac_i (transposed) in column
// chunks of XC by triangular solves against the LDS-resident factors.
template <typename T, int RT>
__global__ __launch_bounds__(256) void k_btri_lu_factor(int nb, int m, int bs, BtriArgs<T> a, int* ipiv,
                                                        int64_t pbs, int64_t pis, int* info, int XC) {
  constexpr int BP = 16 * RT;
  extern __shared__ unsigned char smem_raw[];
  T* S = reinterpret_cast<T*>(smem_raw);  // BP x (BP+1), padded diagonal block
  const int ldx = lu_factor_ldx(XC);
  T* X = S + BP * (BP + 1);                // bs x ldx right-hand-side chunk
  T* TA = X + bs * ldx;                    // BP x (KC+1)
  T* TB = TA + BP * (KC + 1);              // KC x (BP+1)
  int* piv = reinterpret_cast<int*>(TB + KC * (BP + 1));
  int* Q = piv + bs;
  __shared__ __align__(16) int wsp[6 * 16 + 2 * 16 * (int)(sizeof(T) / sizeof(int))];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int ld = BP + 1;
  const float rbs = 1.0f / (float)bs;
  const int nn = bs * bs;
  auto Lat = [&](int r, int k) -> T { return S[r * ld + k]; };
  auto UTat = [&](int r, int k) -> T { return S[k * ld + r]; };
  for (int64_t b = blockIdx.x; b < nb; b += gridDim.x) {
    for (int i = 0; i < m; ++i) {
      T* Di = a.diag + b * a.dbs + (int64_t)i * a.dis;
      stage_block<T>(S, Di, bs, rbs, ld);
      if (i > 0) {
        T acc[RT][RT];
#pragma unroll
        for (int x = 0; x < RT; ++x)
#pragma unroll
          for (int y = 0; y < RT; ++y) acc[x][y] = T(0);
        const T* Lp = a.sub + b * a.lbs + (int64_t)(i - 1) * a.lis;
        const T* Up = a.sup + b * a.ubs + (int64_t)(i - 1) * a.uis;
        mm_global_global<T, RT>(Lp, Up, bs, TA, TB, acc);
#pragma unroll
        for (int x = 0; x < RT; ++x)
#pragma unroll
          for (int y = 0; y < RT; ++y) S[(ty + 16 * x) * ld + tx + 16 * y] -= acc[x][y];
        __syncthreads();
      }
      lu_lds_blocked<T, RT>(S, bs, ld, piv, wsp, info, b);
      for (int e = tid; e < nn; e += 256) {
        const int r = idiv(e, rbs);
        Di[e] = S[r * ld + e - r * bs];
      }
      int* Pi = ipiv + b * pbs + (int64_t)i * pis;
      for (int r = tid; r < bs; r += 256) Pi[r] = piv[r] + 1;
      if (i < m - 1) {
        T* Ui = a.sup + b * a.ubs + (int64_t)i * a.uis;
        T* Li = a.sub + b * a.lbs + (int64_t)i * a.lis;
        piv_to_perm(piv, bs, Q);
        for (int c0 = 0; c0 < bs; c0 += XC) {
          const int nc = (bs - c0) < XC ? (bs - c0) : XC;
          const float rnc = 1.0f / (float)nc;
          // Ufac chunk: L^{-1} (P sup_i)[:, c0:c0+nc]
          for (int e = tid; e < bs * nc; e += 256) {
            const int r = idiv(e, rnc), c = e - r * nc;
            X[r * ldx + c] = Ui[(int64_t)Q[r] * bs + c0 + c];
          }
          __syncthreads();
          tri_solve<T, false, true>(Lat, bs, X, ldx, nc);
          for (int e = tid; e < bs * nc; e += 256) {
            const int r = idiv(e, rnc), c = e - r * nc;
            Ui[(int64_t)r * bs + c0 + c] = X[r * ldx + c];
          }
          __syncthreads();
          // Lfac chunk: rows c0..c0+nc of sub_i U^{-1}, i.e. U^T Z = sub_i^T
          for (int e = tid; e < bs * nc; e += 256) {
            const int c = idiv(e, rbs), r = e - c * bs;
            X[r * ldx + c] = Li[(int64_t)(c0 + c) * bs + r];
          }
          __syncthreads();
          tri_solve<T, false, false>(UTat, bs, X, ldx, nc);
          for (int e = tid; e < bs * nc; e += 256) {
            const int c = idiv(e, rbs), r = e - c * bs;
            Li[(int64_t)(c0 + c) * bs + r] = X[r * ldx + c];
          }
          __syncthreads();
        }
      }
      __threadfence();
      __syncthreads();
    }
  }
}

template <typename T, int RT>
hipError_t launch_lu_factor(int nb, int m, int bs, BtriArgs<T> a, int* ipiv, int64_t pbs, int64_t pis, int* info,
                            hipStream_t s) {
  const int xc = lu_factor_xc<T>(bs);
  const size_t lds = lu_factor_lds<T>(bs, xc);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_btri_lu_factor<T, RT>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_btri_lu_factor<T, RT>), dim3(nb < 65535 ? nb : 65535), dim3(256), lds, s, nb, m, bs, a,
                     ipiv, pbs, pis, info, xc);
  return hipGetLastError();
}

}  // namespace

template <typename T>
bool btri_lu_fused_ok(int bs) {
  return (bs + 15) / 16 <= 10 && bs <= 512 && lu_factor_lds<T>(bs, lu_factor_xc<T>(bs)) <= kLdsMax &&
         lu_solve_lds<T>(bs, 1) <= kLdsMax;
}

template <typename T>
hipError_t btri_lu_factor_fused(int nb, int m, int bs, T* diag, int64_t dbs, int64_t dis, T* sub, int64_t lbs,
                                int64_t lis, T* sup, int64_t ubs, int64_t uis, int* ipiv, int64_t pbs, int64_t pis,
                                int* info, hipStream_t s) {
  BtriArgs<T> a{diag, dbs, dis, sub, lbs, lis, sup, ubs, uis};
  switch (tile_rt(bs)) {
    case 1: return launch_lu_factor<T, 1>(nb, m, bs, a, ipiv, pbs, pis, info, s);
    case 2: return launch_lu_factor<T, 2>(nb, m, bs, a, ipiv, pbs, pis, info, s);
    case 3: return launch_lu_factor<T, 3>(nb, m, bs, a, ipiv, pbs, pis, info, s);
    case 4: return launch_lu_factor<T, 4>(nb, m, bs, a, ipiv, pbs, pis, info, s);
    case 5: return launch_lu_factor<T, 5>(nb, m, bs, a, ipiv, pbs, pis, info, s);
    case 6: return launch_lu_factor<T, 6>(nb, m, bs, a, ipiv, pbs, pis, info, s);
    case 7: return launch_lu_factor<T, 7>(nb, m, bs, a, ipiv, pbs, pis, info, s);
    case 8: return launch_lu_factor<T, 8>(nb, m, bs, a, ipiv, pbs, pis, info, s);
    case 9: return launch_lu_factor<T, 9>(nb, m, bs, a, ipiv, pbs, pis, info, s);
    case 10: return launch_lu_factor<T, 10>(nb, m, bs, a, ipiv, pbs, pis, info, s);
    default: return hipErrorInvalidValue;
  }
}

#define SLAB_BTRI_LU_FACTOR_INST(T)                                                                               \
  template bool btri_lu_fused_ok<T>(int);                                                                          \
  template hipError_t btri_lu_factor_fused<T>(int, int, int, T*, int64_t, int64_t, T*, int64_t, int64_t, T*,      \
                                              int64_t, int64_t, int*, int64_t, int64_t, int*, hipStream_t);
SLAB_BTRI_LU_FACTOR_INST(double)
SLAB_BTRI_LU_FACTOR_INST(float)

}  // namespace slab
