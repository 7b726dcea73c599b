// Block-tridiagonal LU (build_sweep of src_solver/block_tridiag_solver_dense.py
// :152-177): one workgroup per chain runs the whole factorization with each
// diagonal block resident in LDS.
#pragma once

#include "lds_dev.h"

namespace slab {

namespace {

// Elements 0..n-1 from global memory into LDS, eight loads in flight per
// thread: load(e) reads element e (e < n; the tail reads element n-1
// again), store(e, v) writes it.
template <typename T, typename Load, typename Store>
__device__ inline void gather8(int n, Load load, Store store) {
  for (int e0 = threadIdx.x; e0 < n; e0 += 256 * 8) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + 256 * u;
      v[u] = load(e < n ? e : n - 1);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + 256 * u;
      if (e < n) store(e, v[u]);
    }
  }
}

// One workgroup per chain for the whole factorization.  Per block: stage
// D_i, subtract Lfac_{i-1} Ufac_{i-1} (products from global in KC-wide LDS
// chunks), blocked LU with partial pivoting in LDS, then Ufac_i (rows gathered
// through the composed pivot permutation) and Lfac_i (transposed) in column
// chunks of XC by triangular solves against the LDS-resident factors.
template <typename T, int RT>
__global__ __launch_bounds__(256) void k_btri_lu_factor(int bs, ChainSet<T> c0, ChainSet<T> c1, int XC) {
  constexpr int BP = 16 * RT;
  extern __shared__ unsigned char smem_raw[];
  T* S = reinterpret_cast<T*>(smem_raw);  // BP x (BP+1), padded diagonal block
  const int ldx = lu_factor_ldx<T>(XC);
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
  const int nt = c0.nchains + c1.nchains;
  for (int g = blockIdx.x; g < nt; g += gridDim.x) {
    const bool sec = g >= c0.nchains;
    const ChainSet<T> a = sec ? c1 : c0;
    const int64_t b = sec ? g - c0.nchains : g;
    const int m = a.m;
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
      lu_lds_blocked<T, RT>(S, bs, ld, piv, wsp, a.info, b);
      for (int e = tid; e < nn; e += 256) {
        const int r = idiv(e, rbs);
        Di[e] = S[r * ld + e - r * bs];
      }
      int* Pi = a.ipiv + b * a.pbs + (int64_t)i * a.pis;
      for (int r = tid; r < bs; r += 256) Pi[r] = piv[r] + 1;
      if (i < m - 1 || a.open) {
        T* Ui = a.sup + b * a.ubs + (int64_t)i * a.uis;
        T* Li = a.sub + b * a.lbs + (int64_t)i * a.lis;
        piv_to_perm(piv, bs, Q);
        for (int c0 = 0; c0 < bs; c0 += XC) {
          const int nc = (bs - c0) < XC ? (bs - c0) : XC;
          const float rnc = 1.0f / (float)nc;
          // Ufac chunk: L^{-1} (P sup_i)[:, c0:c0+nc]
          gather8<T>(bs * nc, [&](int e) {
            const int r = idiv(e, rnc), c = e - r * nc;
            return Ui[(int64_t)Q[r] * bs + c0 + c];
          }, [&](int e, T v) {
            const int r = idiv(e, rnc), c = e - r * nc;
            X[r * ldx + c] = v;
          });
          __syncthreads();
          tri_solve_wide<T, false, true>(Lat, bs, X, ldx, nc);
          for (int e = tid; e < bs * nc; e += 256) {
            const int r = idiv(e, rnc), c = e - r * nc;
            Ui[(int64_t)r * bs + c0 + c] = X[r * ldx + c];
          }
          __syncthreads();
          // Lfac chunk: rows c0..c0+nc of sub_i U^{-1}, i.e. U^T Z = sub_i^T
          gather8<T>(bs * nc, [&](int e) {
            const int c = idiv(e, rbs), r = e - c * bs;
            return Li[(int64_t)(c0 + c) * bs + r];
          }, [&](int e, T v) {
            const int c = idiv(e, rbs), r = e - c * bs;
            X[r * ldx + c] = v;
          });
          __syncthreads();
          tri_solve_wide<T, false, false>(UTat, bs, X, ldx, nc);
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
hipError_t launch_lu_factor(int bs, const ChainSet<T>& c0, const ChainSet<T>& c1, hipStream_t s) {
  const int xc = lu_factor_xc<T>(bs);
  const size_t lds = lu_factor_lds<T>(bs, xc);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_btri_lu_factor<T, RT>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  const int nt = c0.nchains + c1.nchains;
  if (nt <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_btri_lu_factor<T, RT>), dim3(nt < 65535 ? nt : 65535), dim3(256), lds, s, bs, c0, c1, xc);
  return hipGetLastError();
}

template <typename T>
hipError_t factor_sets(int bs, const ChainSet<T>& c0, const ChainSet<T>& c1, hipStream_t s) {
  switch (tile_rt(bs)) {
    case 1: return launch_lu_factor<T, 1>(bs, c0, c1, s);
    case 2: return launch_lu_factor<T, 2>(bs, c0, c1, s);
    case 3: return launch_lu_factor<T, 3>(bs, c0, c1, s);
    case 4: return launch_lu_factor<T, 4>(bs, c0, c1, s);
    case 5: return launch_lu_factor<T, 5>(bs, c0, c1, s);
    case 6: return launch_lu_factor<T, 6>(bs, c0, c1, s);
    case 7: return launch_lu_factor<T, 7>(bs, c0, c1, s);
    case 8: return launch_lu_factor<T, 8>(bs, c0, c1, s);
    case 9: return launch_lu_factor<T, 9>(bs, c0, c1, s);
    case 10: return launch_lu_factor<T, 10>(bs, c0, c1, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

template <typename T>
bool btri_lu_fused_ok(int bs) {
  return (bs + 15) / 16 <= 10 && bs <= (sizeof(T) == 8 ? 112 : 160) &&
         lu_factor_lds<T>(bs, lu_factor_xc<T>(bs)) <= kLdsMax &&
         lu_solve_lds<T>(bs, 1) <= kLdsMax;
}

template <typename T>
hipError_t btri_lu_factor_fused(int nb, int m, int bs, T* diag, int64_t dbs, int64_t dis, T* sub, int64_t lbs,
                                int64_t lis, T* sup, int64_t ubs, int64_t uis, int* ipiv, int64_t pbs, int64_t pis,
                                int* info, hipStream_t s) {
  ChainSet<T> c0{diag, dbs, dis, sub, lbs, lis, sup, ubs, uis, ipiv, pbs, pis, info, nullptr, 0, 0, m, nb, 0};
  ChainSet<T> none{};
  return factor_sets<T>(bs, c0, none, s);
}

// Twisted block LU: blocks 0..k-1 eliminated top-down and blocks m-1..k+1
// bottom-up (k = m / 2), both chains of every system in one launch, then
// the middle block D_k - Lfac_{k-1} Ufac_{k-1} - Lfac'_{k+1} Ufac'_{k+1}.
// The top chain's last Ufac/Lfac land in sup_{k-1}/sub_{k-1}, the bottom
// chain's in sub_k/sup_k.
template <typename T>
hipError_t btri_lu_factor_twisted(int nb, int m, int bs, T* diag, int64_t dbs, int64_t dis, T* sub, int64_t lbs,
                                  int64_t lis, T* sup, int64_t ubs, int64_t uis, int* ipiv, int64_t pbs,
                                  int64_t pis, int* info, hipStream_t s) {
  const int k = m / 2, lb = m - 1 - k;
  hipError_t e;
  if (k > 0) {
    ChainSet<T> top{diag, dbs, dis, sub, lbs, lis, sup, ubs, uis, ipiv, pbs, pis, info, nullptr, 0, 0, k, nb, 1};
    ChainSet<T> bot = mirrored<T>(nb, m, lb, bs, diag, dbs, dis, sub, lbs, lis, sup, ubs, uis, ipiv, pbs, pis, info,
                                  nullptr, 0, 0);
    if ((e = factor_sets<T>(bs, top, bot, s)) != hipSuccess) return e;
    View<T> Dk{diag + (int64_t)k * dis, bs, 1, dbs};
    View<T> Lt{sub + (int64_t)(k - 1) * lis, bs, 1, lbs}, Ut{sup + (int64_t)(k - 1) * uis, bs, 1, ubs};
    if ((e = gemm<T>(nb, bs, bs, bs, T(-1), Lt, Ut, T(1), Dk, s)) != hipSuccess) return e;
    if (lb > 0) {
      View<T> Lb{sup + (int64_t)k * uis, bs, 1, ubs}, Ub{sub + (int64_t)k * lis, bs, 1, lbs};
      if ((e = gemm<T>(nb, bs, bs, bs, T(-1), Lb, Ub, T(1), Dk, s)) != hipSuccess) return e;
    }
  }
  ChainSet<T> mid{diag + (int64_t)k * dis, dbs, dis, sub, lbs, lis, sup, ubs, uis,
                  ipiv + (int64_t)k * pis, pbs, pis, info, nullptr, 0, 0, 1, nb, 0};
  ChainSet<T> none{};
  return factor_sets<T>(bs, mid, none, s);
}

#define SLAB_BTRI_LU_FACTOR_INST(T)                                                                               \
  template bool btri_lu_fused_ok<T>(int);                                                                          \
  template hipError_t btri_lu_factor_fused<T>(int, int, int, T*, int64_t, int64_t, T*, int64_t, int64_t, T*,      \
                                              int64_t, int64_t, int*, int64_t, int64_t, int*, hipStream_t);        \
  template hipError_t btri_lu_factor_twisted<T>(int, int, int, T*, int64_t, int64_t, T*, int64_t, int64_t, T*,    \
                                                int64_t, int64_t, int*, int64_t, int64_t, int*, hipStream_t);
}  // namespace slab
