// Block-tridiagonal factorization in block-Thomas form with explicit
// diagonal inverses, and the matching sweeps, for gfx950.
//
//   S_0 = D_0,  S_i = D_i - sub_{i-1} G_{i-1},  G_i = S_i^{-1} sup_i
//   forward:   y_i = S_i^{-1} (f_i - sub_{i-1} y_{i-1})
//   backward:  x_i = y_i - G_i x_{i+1}
//
// which is the Schur recursion of src_solver/block_tridiag_solver_structured.py
// :34-76 (and the block LU of block_tridiag_solver_dense.py:152-177 up to
// how the diagonal factor is kept).  Keeping S_i^{-1} instead of its LU
// turns every sweep step into matrix products: no sequential triangular
// solves and no row interchanges in the solve.  It serves the FP64
// interface system; the slab chains use the backward-stable LU form.  Storage in place: diag_i <-
// S_i^{-1}, sup_i <- G_i, sub_i unchanged.
//
// Chains whose blocks fit in LDS (bs <= 128 in FP64, bs <= 160 in FP32) run
// one workgroup per chain for the whole factorization (Gauss-Jordan with
// partial pivoting in LDS) and one workgroup per (chain, right-hand-side
// tile) for the whole sweep.  Larger blocks go through batched getrf / getrs
// / gemm launches.
#include "lds_dev.h"

namespace slab {

namespace {

// In-place Gauss-Jordan inverse with partial (row) pivoting of the n x n
// matrix S in LDS (row stride n+1).  Row interchanges are undone at the end
// by the matching column interchanges in reverse order (as getri does).
template <typename T, int RT>
__device__ void gj_inverse(T* S, int n, int* piv, T* red_v, int* red_i, int* s_piv, int* info, int64_t b) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tx = tid & 15, ty = tid >> 4;
  const int ld = n + 1;
  for (int k = 0; k < n; ++k) {
    T best = T(-1);
    int bi = INT_MAX;
    for (int r = k + tid; r < n; r += 256) {
      const T v = absv(S[r * ld + k]);
      if (v > best) { best = v; bi = r; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const T ov = __shfl_xor(best, off);
      const int oi = __shfl_xor(bi, off);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if (lane == 0) { red_v[wave] = best; red_i[wave] = bi; }
    __syncthreads();
    if (tid == 0) {
      T bv = red_v[0];
      int bx = red_i[0];
      for (int q = 1; q < 4; ++q)
        if (red_v[q] > bv || (red_v[q] == bv && red_i[q] < bx)) { bv = red_v[q]; bx = red_i[q]; }
      if (bx < k || bx >= n) bx = k;
      *s_piv = bx;
      piv[k] = bx;
    }
    __syncthreads();
    const int p = *s_piv;
    if (p != k) {
      for (int c = tid; c < n; c += 256) {
        const T t = S[k * ld + c];
        S[k * ld + c] = S[p * ld + c];
        S[p * ld + c] = t;
      }
    }
    __syncthreads();
    T pv = S[k * ld + k];
    if (pv == T(0)) {
      if (tid == 0 && info != nullptr && info[b] == 0) info[b] = k + 1;
    }
    const T rp = T(1) / pv;
    // entries outside row k and column k
    T rowk[RT], colk[RT];
#pragma unroll
    for (int q = 0; q < RT; ++q) {
      const int c = tx + 16 * q;
      rowk[q] = ((c < n) & (c != k)) ? S[k * ld + c] * rp : T(0);
    }
#pragma unroll
    for (int a = 0; a < RT; ++a) {
      const int r = ty + 16 * a;
      colk[a] = ((r < n) & (r != k)) ? S[r * ld + k] : T(0);
    }
#pragma unroll
    for (int a = 0; a < RT; ++a) {
      const int r = ty + 16 * a;
      if (r < n && r != k) {
#pragma unroll
        for (int q = 0; q < RT; ++q) {
          const int c = tx + 16 * q;
          if (c < n && c != k) S[r * ld + c] -= colk[a] * rowk[q];
        }
      }
    }
    __syncthreads();
    for (int c = tid; c < n; c += 256)
      if (c != k) S[k * ld + c] *= rp;
    for (int r = tid; r < n; r += 256)
      if (r != k) S[r * ld + k] = -S[r * ld + k] * rp;
    if (tid == 0) S[k * ld + k] = rp;
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; --k) {
    const int p = piv[k];
    if (p != k) {
      for (int r = tid; r < n; r += 256) {
        const T t = S[r * ld + k];
        S[r * ld + k] = S[r * ld + p];
        S[r * ld + p] = t;
      }
    }
    __syncthreads();
  }
}

template <typename T, int RT>
__global__ __launch_bounds__(256) void k_btri_inv_factor(int nb, int m, int bs, BtriArgs<T> a, int* info) {
  extern __shared__ unsigned char smem_raw[];
  T* S = reinterpret_cast<T*>(smem_raw);
  constexpr int BP = 16 * RT;
  T* TA = S + bs * (bs + 1);   // BP x (KC+1)
  T* TB = TA + BP * (KC + 1);  // KC x (BP+1)
  int* piv = reinterpret_cast<int*>(TB + KC * (BP + 1));
  __shared__ T red_v[4];
  __shared__ int red_i[4];
  __shared__ int s_piv;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int ld = bs + 1;
  for (int64_t b = blockIdx.x; b < nb; b += gridDim.x) {
    for (int i = 0; i < m; ++i) {
      T* Di = a.diag + b * a.dbs + (int64_t)i * a.dis;
      for (int e = tid; e < bs * bs; e += 256) {
        const int r = e / bs, c = e - r * bs;
        S[r * ld + c] = Di[e];
      }
      __syncthreads();
      if (i > 0) {
        T acc[RT][RT];
#pragma unroll
        for (int x = 0; x < RT; ++x)
#pragma unroll
          for (int y = 0; y < RT; ++y) acc[x][y] = T(0);
        const T* Lp = a.sub + b * a.lbs + (int64_t)(i - 1) * a.lis;
        const T* Gp = a.sup + b * a.ubs + (int64_t)(i - 1) * a.uis;
        mm_global_global<T, RT>(Lp, Gp, bs, TA, TB, acc);
#pragma unroll
        for (int x = 0; x < RT; ++x) {
          const int r = ty + 16 * x;
#pragma unroll
          for (int y = 0; y < RT; ++y) {
            const int c = tx + 16 * y;
            if (r < bs && c < bs) S[r * ld + c] -= acc[x][y];
          }
        }
        __syncthreads();
      }
      gj_inverse<T, RT>(S, bs, piv, red_v, red_i, &s_piv, info, b);
      for (int e = tid; e < bs * bs; e += 256) {
        const int r = e / bs, c = e - r * bs;
        Di[e] = S[r * ld + c];
      }
      if (i < m - 1) {
        T acc[RT][RT];
#pragma unroll
        for (int x = 0; x < RT; ++x)
#pragma unroll
          for (int y = 0; y < RT; ++y) acc[x][y] = T(0);
        T* Ui = a.sup + b * a.ubs + (int64_t)i * a.uis;
        mm_lds_global<T, RT>(S, Ui, bs, TB, acc);
        // every read of sup_i is done (barrier inside): overwrite with G_i
#pragma unroll
        for (int x = 0; x < RT; ++x) {
          const int r = ty + 16 * x;
#pragma unroll
          for (int y = 0; y < RT; ++y) {
            const int c = tx + 16 * y;
            if (r < bs && c < bs) Ui[(int64_t)r * bs + c] = acc[x][y];
          }
        }
      }
      // G_i and S_i^{-1} are read back by this workgroup in the next step
      __threadfence();
      __syncthreads();
    }
  }
}

// y (rows of this thread) = M (bs x bs, global) @ V (LDS, bs x NR, row
// stride NR) for the NR columns of the tile; thread owns column tc and rows
// tr + RG*a.
template <typename T, int RPT>
__device__ void mv_global_lds(const T* M, const T* V, int bs, int nr, T* TA, T (&acc)[RPT]) {
  const int tid = threadIdx.x;
  const int tc = tid % nr, tr = tid / nr, RG = 256 / nr;
  for (int k0 = 0; k0 < bs; k0 += KC) {
    for (int e = tid; e < bs * KC; e += 256) {
      const int r = e / KC, kk = e - r * KC;
      TA[r * (KC + 1) + kk] = (k0 + kk < bs) ? M[(int64_t)r * bs + k0 + kk] : T(0);
    }
    __syncthreads();
    const int kn = (bs - k0) < KC ? (bs - k0) : KC;
    if (tr < RG) {
      for (int kk = 0; kk < kn; ++kk) {
        const T v = V[(k0 + kk) * nr + tc];
#pragma unroll
        for (int x = 0; x < RPT; ++x) {
          const int r = tr + RG * x;
          if (r < bs) acc[x] += TA[r * (KC + 1) + kk] * v;
        }
      }
    }
    __syncthreads();
  }
}

template <typename T, int RPT>
__global__ __launch_bounds__(256) void k_btri_inv_solve(int nb, int m, int bs, BtriArgs<T> a, T* f, int64_t fbs,
                                                        int64_t ldf, int nrhs, int NR) {
  extern __shared__ unsigned char smem_raw[];
  T* Y = reinterpret_cast<T*>(smem_raw);  // bs x NR
  T* Rv = Y + bs * NR;                    // bs x NR
  T* TA = Rv + bs * NR;                   // bs x (KC+1)
  const int tid = threadIdx.x;
  const int nr = NR;
  const int tc = tid % nr, tr = tid / nr, RG = 256 / nr;
  const int c0 = blockIdx.y * NR;
  const int ncols = (nrhs - c0) < NR ? (nrhs - c0) : NR;
  for (int64_t b = blockIdx.x; b < nb; b += gridDim.x) {
    T* fb = f + b * fbs;
    auto F = [&](int i, int r, int c) -> T* { return fb + (int64_t)(i * bs + r) * ldf + c0 + c; };
    // forward sweep
    for (int i = 0; i < m; ++i) {
      for (int e = tid; e < bs * nr; e += 256) {
        const int r = e / nr, c = e - r * nr;
        Rv[e] = (c < ncols) ? *F(i, r, c) : T(0);
      }
      __syncthreads();
      if (i > 0) {
        T acc[RPT];
#pragma unroll
        for (int x = 0; x < RPT; ++x) acc[x] = T(0);
        mv_global_lds<T, RPT>(a.sub + b * a.lbs + (int64_t)(i - 1) * a.lis, Y, bs, nr, TA, acc);
        if (tr < RG) {
#pragma unroll
          for (int x = 0; x < RPT; ++x) {
            const int r = tr + RG * x;
            if (r < bs) Rv[r * nr + tc] -= acc[x];
          }
        }
        __syncthreads();
      }
      T acc[RPT];
#pragma unroll
      for (int x = 0; x < RPT; ++x) acc[x] = T(0);
      mv_global_lds<T, RPT>(a.diag + b * a.dbs + (int64_t)i * a.dis, Rv, bs, nr, TA, acc);
      if (tr < RG) {
#pragma unroll
        for (int x = 0; x < RPT; ++x) {
          const int r = tr + RG * x;
          if (r < bs) {
            Y[r * nr + tc] = acc[x];
            if (tc < ncols) *F(i, r, tc) = acc[x];
          }
        }
      }
      __syncthreads();
    }
    // backward sweep: Y holds x_{i+1}
    for (int i = m - 2; i >= 0; --i) {
      T acc[RPT];
#pragma unroll
      for (int x = 0; x < RPT; ++x) acc[x] = T(0);
      mv_global_lds<T, RPT>(a.sup + b * a.ubs + (int64_t)i * a.uis, Y, bs, nr, TA, acc);
      // all reads of Y done (barrier inside mv)
      if (tr < RG) {
#pragma unroll
        for (int x = 0; x < RPT; ++x) {
          const int r = tr + RG * x;
          if (r < bs) {
            const T v = (tc < ncols ? *F(i, r, tc) : T(0)) - acc[x];
            Y[r * nr + tc] = v;
            if (tc < ncols) *F(i, r, tc) = v;
          }
        }
      }
      __syncthreads();
    }
  }
}

template <typename T, int RT>
hipError_t launch_factor(int nb, int m, int bs, BtriArgs<T> a, int* info, hipStream_t s) {
  const size_t lds = factor_lds<T>(bs, RT);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_btri_inv_factor<T, RT>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_btri_inv_factor<T, RT>), dim3(nb < 65535 ? nb : 65535), dim3(256), lds, s, nb, m, bs, a,
                     info);
  return hipGetLastError();
}

template <typename T, int RPT>
hipError_t launch_solve(int nb, int m, int bs, BtriArgs<T> a, T* f, int64_t fbs, int64_t ldf, int nrhs, int nr,
                        hipStream_t s) {
  const size_t lds = solve_lds<T>(bs, nr);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_btri_inv_solve<T, RPT>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  dim3 grid(nb < 65535 ? nb : 65535, (nrhs + nr - 1) / nr);
  hipLaunchKernelGGL((k_btri_inv_solve<T, RPT>), grid, dim3(256), lds, s, nb, m, bs, a, f, fbs, ldf, nrhs, nr);
  return hipGetLastError();
}

// RT instantiated for a block size
int inv_rt(int bs) {
  const int rt = tile_rt(bs);
  return rt <= 2 ? rt : (rt <= 4 ? 4 : (rt <= 6 ? 6 : (rt <= 8 ? 8 : 10)));
}

}  // namespace

template <typename T>
bool btri_inv_fused_ok(int bs) {
  return tile_rt(bs) <= 10 && factor_lds<T>(bs, inv_rt(bs)) <= kLdsMax;
}

template <typename T>
hipError_t btri_inv_factor(int nb, int m, int bs, T* diag, int64_t dbs, int64_t dis, T* sub, int64_t lbs,
                           int64_t lis, T* sup, int64_t ubs, int64_t uis, int* info, T* ws, int* iws,
                           hipStream_t s) {
  BtriArgs<T> a{diag, dbs, dis, sub, lbs, lis, sup, ubs, uis};
  if (btri_inv_fused_ok<T>(bs)) {
    const int rt = (bs + 15) / 16;
    switch (rt) {
      case 1: return launch_factor<T, 1>(nb, m, bs, a, info, s);
      case 2: return launch_factor<T, 2>(nb, m, bs, a, info, s);
      case 3: case 4: return launch_factor<T, 4>(nb, m, bs, a, info, s);
      case 5: case 6: return launch_factor<T, 6>(nb, m, bs, a, info, s);
      case 7: case 8: return launch_factor<T, 8>(nb, m, bs, a, info, s);
      default: return launch_factor<T, 10>(nb, m, bs, a, info, s);
    }
  }
  // large blocks: batched getrf + getrs(identity) for S_i^{-1}, gemm for the
  // Schur updates and G_i; ws holds 2 * nb * bs^2 elements, iws nb * bs.
  if (ws == nullptr || iws == nullptr) return hipErrorInvalidValue;
  const int64_t bb = (int64_t)bs * bs;
  View<T> W1{ws, bs, 1, bb}, W2{ws + nb * bb, bs, 1, bb};
  hipError_t e;
  for (int i = 0; i < m; ++i) {
    View<T> Di{diag + (int64_t)i * dis, bs, 1, dbs};
    if (i > 0) {
      View<T> Lp{sub + (int64_t)(i - 1) * lis, bs, 1, lbs};
      View<T> Gp{sup + (int64_t)(i - 1) * uis, bs, 1, ubs};
      if ((e = gemm<T>(nb, bs, bs, bs, T(-1), Lp, Gp, T(1), Di, s)) != hipSuccess) return e;
    }
    if ((e = copy_batched<T>(nb, bs, bs, Di, W1, s)) != hipSuccess) return e;
    if ((e = getrf<T>(nb, bs, bs, bs, bs, W1, iws, bs, info, s)) != hipSuccess) return e;
    if ((e = set_identity<T>(nb, bs, Di, s)) != hipSuccess) return e;
    if ((e = getrs<T>(nb, bs, bs, false, W1, iws, bs, Di, s)) != hipSuccess) return e;
    if (i < m - 1) {
      View<T> Ui{sup + (int64_t)i * uis, bs, 1, ubs};
      if ((e = gemm<T>(nb, bs, bs, bs, T(1), Di, Ui, T(0), W2, s)) != hipSuccess) return e;
      if ((e = copy_batched<T>(nb, bs, bs, W2, Ui, s)) != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

template <typename T>
hipError_t btri_inv_solve(int nb, int m, int bs, T* diag, int64_t dbs, int64_t dis, T* sub, int64_t lbs,
                          int64_t lis, T* sup, int64_t ubs, int64_t uis, T* f, int64_t fbs, int64_t ldf, int nrhs,
                          T* ws, hipStream_t s) {
  BtriArgs<T> a{diag, dbs, dis, sub, lbs, lis, sup, ubs, uis};
  const int nr = nrhs >= 32 ? 32 : (nrhs >= 16 ? 16 : (nrhs >= 8 ? 8 : (nrhs >= 4 ? 4 : (nrhs >= 2 ? 2 : 1))));
  const int rg = 256 / nr;
  const int rpt = (bs + rg - 1) / rg;
  if (btri_inv_fused_ok<T>(bs) && solve_lds<T>(bs, nr) <= kLdsMax && rpt <= 24) {
    if (rpt <= 1) return launch_solve<T, 1>(nb, m, bs, a, f, fbs, ldf, nrhs, nr, s);
    if (rpt <= 2) return launch_solve<T, 2>(nb, m, bs, a, f, fbs, ldf, nrhs, nr, s);
    if (rpt <= 4) return launch_solve<T, 4>(nb, m, bs, a, f, fbs, ldf, nrhs, nr, s);
    if (rpt <= 8) return launch_solve<T, 8>(nb, m, bs, a, f, fbs, ldf, nrhs, nr, s);
    if (rpt <= 13) return launch_solve<T, 13>(nb, m, bs, a, f, fbs, ldf, nrhs, nr, s);
    return launch_solve<T, 24>(nb, m, bs, a, f, fbs, ldf, nrhs, nr, s);
  }
  // large blocks: gemm per step through a workspace of nb * bs * nrhs
  if (ws == nullptr) return hipErrorInvalidValue;
  View<T> W{ws, nrhs, 1, (int64_t)bs * nrhs};
  auto Fi = [&](int i) { return View<T>{f + (int64_t)i * bs * ldf, ldf, 1, fbs}; };
  hipError_t e;
  for (int i = 0; i < m; ++i) {
    if (i > 0) {
      View<T> Lp{sub + (int64_t)(i - 1) * lis, bs, 1, lbs};
      if ((e = gemm<T>(nb, bs, nrhs, bs, T(-1), Lp, Fi(i - 1), T(1), Fi(i), s)) != hipSuccess) return e;
    }
    View<T> Di{diag + (int64_t)i * dis, bs, 1, dbs};
    if ((e = gemm<T>(nb, bs, nrhs, bs, T(1), Di, Fi(i), T(0), W, s)) != hipSuccess) return e;
    if ((e = copy_batched<T>(nb, bs, nrhs, W, Fi(i), s)) != hipSuccess) return e;
  }
  for (int i = m - 2; i >= 0; --i) {
    View<T> Gi{sup + (int64_t)i * uis, bs, 1, ubs};
    if ((e = gemm<T>(nb, bs, nrhs, bs, T(-1), Gi, Fi(i + 1), T(1), Fi(i), s)) != hipSuccess) return e;
  }
  return hipSuccess;
}

#define SLAB_BTRI_INV_INST(T)                                                                                     \
  template bool btri_inv_fused_ok<T>(int);                                                                         \
  template hipError_t btri_inv_factor<T>(int, int, int, T*, int64_t, int64_t, T*, int64_t, int64_t, T*, int64_t,  \
                                         int64_t, int*, T*, int*, hipStream_t);                                    \
  template hipError_t btri_inv_solve<T>(int, int, int, T*, int64_t, int64_t, T*, int64_t, int64_t, T*, int64_t,   \
                                        int64_t, T*, int64_t, int64_t, int, T*, hipStream_t);
SLAB_BTRI_INV_INST(double)
SLAB_BTRI_INV_INST(float)

}  // namespace slab
