// Block-tridiagonal factorization in block-Thomas form with explicit
// diagonal inverses, and the matching sweeps, for gfx950.
//
//   S_0 = D_0,  S_i = D_i - sub_{i-1} G_{i-1},  G_i = S_i^{-1} sup_i
//   forward:   y_i = S_i^{-1} (f_i - sub_{i-1} y_{i-1})
//   backward:  x_i = y_i - G_i x_{i+1}
//
// which is the Schur recursion of src_solver/block_tridiag_solver_structured.py
// :34-76 (and the block LU of block_tridiag_solver_dense.py:152-222 up to
// how the diagonal factor is kept).  Keeping S_i^{-1} instead of its LU
// turns every sweep step into matrix products: no sequential triangular
// solves and no row interchanges in the solve.  Storage in place: diag_i <-
// S_i^{-1}, sup_i <- G_i, sub_i unchanged.
//
// Chains whose blocks fit in LDS (bs <= 128 in FP64, bs <= 160 in FP32) run
// one workgroup per chain for the whole factorization (Gauss-Jordan with
// partial pivoting in LDS) and one workgroup per (chain, right-hand-side
// tile) for the whole sweep.  Larger blocks go through batched getrf / getrs
// / gemm launches.
#include "slab_internal.h"

#include <climits>

namespace slab {

namespace {

constexpr int KC = 16;  // k-chunk staged through LDS in the matrix products

template <typename T>
__device__ inline T absv(T v) { return v < T(0) ? -v : v; }

template <typename T>
struct BtriArgs {
  T* diag; int64_t dbs, dis;
  T* sub; int64_t lbs, lis;
  T* sup; int64_t ubs, uis;
};

// acc[a][q] += sum_k A(r_a, k) B(k, c_q) for a bs x bs A and B, both read
// from global memory in KC-wide chunks staged in LDS; thread (ty, tx) of a
// 16 x 16 arrangement owns rows ty + 16a, columns tx + 16q.
template <typename T, int RT>
__device__ void mm_global_global(const T* A, const T* B, int bs, T* TA, T* TB, T (&acc)[RT][RT]) {
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  for (int k0 = 0; k0 < bs; k0 += KC) {
    for (int e = tid; e < bs * KC; e += 256) {
      const int r = e / KC, kk = e - r * KC;
      TA[r * (KC + 1) + kk] = (k0 + kk < bs) ? A[(int64_t)r * bs + k0 + kk] : T(0);
    }
    for (int e = tid; e < KC * bs; e += 256) {
      const int kk = e / bs, c = e - kk * bs;
      TB[kk * (bs + 1) + c] = (k0 + kk < bs) ? B[(int64_t)(k0 + kk) * bs + c] : T(0);
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < KC; ++kk) {
      T av[RT], bv[RT];
#pragma unroll
      for (int a = 0; a < RT; ++a) {
        const int r = ty + 16 * a;
        av[a] = (r < bs) ? TA[r * (KC + 1) + kk] : T(0);
      }
#pragma unroll
      for (int q = 0; q < RT; ++q) {
        const int c = tx + 16 * q;
        bv[q] = (c < bs) ? TB[kk * (bs + 1) + c] : T(0);
      }
#pragma unroll
      for (int a = 0; a < RT; ++a)
#pragma unroll
        for (int q = 0; q < RT; ++q) acc[a][q] += av[a] * bv[q];
    }
    __syncthreads();
  }
}

// acc[a][q] += sum_k S(r_a, k) B(k, c_q): S in LDS (row stride bs+1), B global.
template <typename T, int RT>
__device__ void mm_lds_global(const T* S, const T* B, int bs, T* TB, T (&acc)[RT][RT]) {
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int ld = bs + 1;
  for (int k0 = 0; k0 < bs; k0 += KC) {
    for (int e = tid; e < KC * bs; e += 256) {
      const int kk = e / bs, c = e - kk * bs;
      TB[kk * (bs + 1) + c] = (k0 + kk < bs) ? B[(int64_t)(k0 + kk) * bs + c] : T(0);
    }
    __syncthreads();
    const int kn = (bs - k0) < KC ? (bs - k0) : KC;
    for (int kk = 0; kk < kn; ++kk) {
      T av[RT], bv[RT];
#pragma unroll
      for (int a = 0; a < RT; ++a) {
        const int r = ty + 16 * a;
        av[a] = (r < bs) ? S[r * ld + k0 + kk] : T(0);
      }
#pragma unroll
      for (int q = 0; q < RT; ++q) {
        const int c = tx + 16 * q;
        bv[q] = (c < bs) ? TB[kk * (bs + 1) + c] : T(0);
      }
#pragma unroll
      for (int a = 0; a < RT; ++a)
#pragma unroll
        for (int q = 0; q < RT; ++q) acc[a][q] += av[a] * bv[q];
    }
    __syncthreads();
  }
}

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
      rowk[q] = (c < n && c != k) ? S[k * ld + c] * rp : T(0);
    }
#pragma unroll
    for (int a = 0; a < RT; ++a) {
      const int r = ty + 16 * a;
      colk[a] = (r < n && r != k) ? S[r * ld + k] : T(0);
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
  T* TA = S + bs * (bs + 1);
  T* TB = TA + bs * (KC + 1);
  int* piv = reinterpret_cast<int*>(TB + KC * (bs + 1));
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

// Batched LU solve for a few right-hand sides with the factors of getrf
// (row interchanges, unit lower, upper), one workgroup per matrix: the
// right-hand side lives in LDS, each 64-row diagonal block is solved by one
// wavefront with shuffles, the rows below/above are updated by all threads.
template <typename T>
__global__ __launch_bounds__(256) void k_getrs_small(int batch, int n, int nrhs, View<T> LU, const int* ipiv,
                                                     int64_t pbs, View<T> B) {
  extern __shared__ unsigned char smem_raw[];
  T* X = reinterpret_cast<T*>(smem_raw);  // n x nrhs
  T* Dg = X + n * nrhs;                   // 64 x 65
  int* P = reinterpret_cast<int*>(Dg + 64 * 65);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int b = blockIdx.x; b < batch; b += gridDim.x) {
    for (int e = tid; e < n * nrhs; e += 256) {
      const int r = e / nrhs, c = e - r * nrhs;
      X[e] = *B.at(b, r, c);
    }
    for (int r = tid; r < n; r += 256) P[r] = ipiv[(int64_t)b * pbs + r] - 1;
    __syncthreads();
    if (tid < nrhs) {
      for (int k = 0; k < n; ++k) {
        const int p = P[k];
        if (p != k) {
          const T t = X[k * nrhs + tid];
          X[k * nrhs + tid] = X[p * nrhs + tid];
          X[p * nrhs + tid] = t;
        }
      }
    }
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
      const bool lower = (pass == 0);
      const int nblk = (n + 63) / 64;
      for (int bi = 0; bi < nblk; ++bi) {
        const int blk = lower ? bi : nblk - 1 - bi;
        const int d0 = blk * 64;
        const int tb = (n - d0) < 64 ? (n - d0) : 64;
        for (int e = tid; e < 64 * 64; e += 256) {
          const int r = e >> 6, c = e & 63;
          Dg[r * 65 + c] = (r < tb && c < tb) ? *LU.at(b, d0 + r, d0 + c) : T(0);
        }
        __syncthreads();
        if (wave == 0) {
          for (int c = 0; c < nrhs; ++c) {
            T x = (lane < tb) ? X[(d0 + lane) * nrhs + c] : T(0);
            if (lower) {
              for (int j = 0; j < tb; ++j) {
                const T xj = __shfl(x, j);
                if (lane > j) x -= Dg[lane * 65 + j] * xj;
              }
            } else {
              for (int j = tb - 1; j >= 0; --j) {
                if (lane == j) x = x / Dg[j * 65 + j];
                const T xj = __shfl(x, j);
                if (lane < j) x -= Dg[lane * 65 + j] * xj;
              }
            }
            if (lane < tb) X[(d0 + lane) * nrhs + c] = x;
          }
        }
        __syncthreads();
        // rows outside the block: below it (lower) or above it (upper); one
        // wavefront per row, lane j reads LU(r, d0+j) (coalesced)
        const int r0 = lower ? d0 + tb : 0;
        const int r1 = lower ? n : d0;
        for (int r = r0 + wave; r < r1; r += 4) {
          const T l = (lane < tb) ? *LU.at(b, r, d0 + lane) : T(0);
          for (int c = 0; c < nrhs; ++c) {
            T v = (lane < tb) ? l * X[(d0 + lane) * nrhs + c] : T(0);
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if (lane == 0) X[r * nrhs + c] -= v;
          }
        }
        __syncthreads();
      }
    }
    for (int e = tid; e < n * nrhs; e += 256) {
      const int r = e / nrhs, c = e - r * nrhs;
      *B.at(b, r, c) = X[e];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// LU form (backward stable; the reference's block LU): diag_i <- packed LU of
// D_i - Lfac_i Ufac_i, sub_i <- Lfac = sub_i U_i^{-1}, sup_i <- Ufac =
// L_i^{-1} P_i^T sup_i, ipiv (b, m, bs) 1-based.
// ---------------------------------------------------------------------------

// Lower triangular solve M X = X in LDS, X n x nc (row stride nc), M(r, k)
// read through `at` (LDS or global, see the callers); 64-row blocks: each
// wavefront substitutes its columns with shuffles (lane = row), then all
// threads update the rows below.  `upper` runs the mirrored algorithm from
// the bottom.  Dg (64 x 65, LDS) stages the diagonal block when M is global.
template <typename T, typename At>
__device__ void tri_solve_lds(At at, int n, bool upper, bool unit, T* X, int nc, T* Dg) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nblk = (n + 63) / 64;
  for (int bi = 0; bi < nblk; ++bi) {
    const int blk = upper ? nblk - 1 - bi : bi;
    const int d0 = blk * 64;
    const int tb = (n - d0) < 64 ? (n - d0) : 64;
    for (int e = tid; e < 64 * 64; e += 256) {
      const int r = e >> 6, c = e & 63;
      Dg[r * 65 + c] = (r < tb && c < tb) ? at(d0 + r, d0 + c) : T(0);
    }
    __syncthreads();
    for (int c = wave; c < nc; c += 4) {
      T x = (lane < tb) ? X[(d0 + lane) * nc + c] : T(0);
      if (!upper) {
        for (int j = 0; j < tb; ++j) {
          if (!unit && lane == j) x = x / Dg[j * 65 + j];
          const T xj = __shfl(x, j);
          if (lane > j) x -= Dg[lane * 65 + j] * xj;
        }
      } else {
        for (int j = tb - 1; j >= 0; --j) {
          if (!unit && lane == j) x = x / Dg[j * 65 + j];
          const T xj = __shfl(x, j);
          if (lane < j) x -= Dg[lane * 65 + j] * xj;
        }
      }
      if (lane < tb) X[(d0 + lane) * nc + c] = x;
    }
    __syncthreads();
    const int r0 = upper ? 0 : d0 + tb;
    const int r1 = upper ? d0 : n;
    const int nrow = r1 - r0;
    for (int e = tid; e < nrow * nc; e += 256) {
      const int rr = e / nc, c = e - rr * nc;
      const int r = r0 + rr;
      T acc = T(0);
      for (int j = 0; j < tb; ++j) acc += at(r, d0 + j) * X[(d0 + j) * nc + c];
      X[r * nc + c] -= acc;
    }
    __syncthreads();
  }
}

// Same as tri_solve_lds for a triangular matrix already resident in LDS
// (`at` reads LDS): no staging of the diagonal blocks.
template <typename T, typename At>
__device__ void tri_solve_direct(At at, int n, bool upper, bool unit, T* X, int nc) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nblk = (n + 63) / 64;
  for (int bi = 0; bi < nblk; ++bi) {
    const int blk = upper ? nblk - 1 - bi : bi;
    const int d0 = blk * 64;
    const int tb = (n - d0) < 64 ? (n - d0) : 64;
    const int rl = d0 + (lane < tb ? lane : 0);
    for (int c = wave; c < nc; c += 4) {
      T x = (lane < tb) ? X[(d0 + lane) * nc + c] : T(0);
      if (!upper) {
        for (int j = 0; j < tb; ++j) {
          if (!unit && lane == j) x = x / at(d0 + j, d0 + j);
          const T xj = __shfl(x, j);
          if (lane > j && lane < tb) x -= at(rl, d0 + j) * xj;
        }
      } else {
        for (int j = tb - 1; j >= 0; --j) {
          if (!unit && lane == j) x = x / at(d0 + j, d0 + j);
          const T xj = __shfl(x, j);
          if (lane < j) x -= at(rl, d0 + j) * xj;
        }
      }
      if (lane < tb) X[(d0 + lane) * nc + c] = x;
    }
    __syncthreads();
    const int r0 = upper ? 0 : d0 + tb;
    const int r1 = upper ? d0 : n;
    const int nrow = r1 - r0;
    for (int e = tid; e < nrow * nc; e += 256) {
      const int rr = e / nc, c = e - rr * nc;
      const int r = r0 + rr;
      T acc = T(0);
      for (int j = 0; j < tb; ++j) acc += at(r, d0 + j) * X[(d0 + j) * nc + c];
      X[r * nc + c] -= acc;
    }
    __syncthreads();
  }
}

// Row interchanges (forward) of a LU with pivots P (0-based, LDS) on X.
template <typename T>
__device__ void apply_piv_lds(const int* P, int n, T* X, int nc) {
  for (int c = threadIdx.x; c < nc; c += 256)
    for (int k = 0; k < n; ++k) {
      const int p = P[k];
      if (p != k) {
        const T t = X[k * nc + c];
        X[k * nc + c] = X[p * nc + c];
        X[p * nc + c] = t;
      }
    }
  __syncthreads();
}

// In-place LU with partial pivoting of S (n x n, LDS, row stride n+1).
template <typename T, int RT>
__device__ void lu_lds(T* S, int n, int* piv, T* red_v, int* red_i, int* s_piv, int* info, int64_t b) {
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
    const T pv = S[k * ld + k];
    const bool nz = pv != T(0);
    if (!nz && tid == 0 && info != nullptr && info[b] == 0) info[b] = k + 1;
    T rowk[RT], colk[RT];
#pragma unroll
    for (int q = 0; q < RT; ++q) {
      const int c = tx + 16 * q;
      rowk[q] = (c < n && c > k) ? S[k * ld + c] : T(0);
    }
#pragma unroll
    for (int a = 0; a < RT; ++a) {
      const int r = ty + 16 * a;
      colk[a] = (r < n && r > k) ? (nz ? S[r * ld + k] / pv : S[r * ld + k]) : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < RT; ++a) {
      const int r = ty + 16 * a;
      if (r < n && r > k) {
        if (tx == (k & 15) && (k >> 4) < RT) S[r * ld + k] = colk[a];
#pragma unroll
        for (int q = 0; q < RT; ++q) {
          const int c = tx + 16 * q;
          if (c < n && c > k) S[r * ld + c] -= colk[a] * rowk[q];
        }
      }
    }
    __syncthreads();
  }
}

template <typename T, int RT>
__global__ __launch_bounds__(256) void k_btri_lu_factor(int nb, int m, int bs, BtriArgs<T> a, int* ipiv,
                                                        int64_t pbs, int64_t pis, int* info) {
  extern __shared__ unsigned char smem_raw[];
  T* S = reinterpret_cast<T*>(smem_raw);   // bs x (bs+1)
  T* X = S + bs * (bs + 1);                 // bs x 32 (right-hand-side chunk)
  T* TA = X + bs * 32;                      // bs x (KC+1)
  T* TB = TA + bs * (KC + 1);               // KC x (bs+1); also the 64 x 65 diagonal stage
  int* piv = reinterpret_cast<int*>(TB + (KC * (bs + 1) > 64 * 65 ? KC * (bs + 1) : 64 * 65));
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
        const T* Up = a.sup + b * a.ubs + (int64_t)(i - 1) * a.uis;
        mm_global_global<T, RT>(Lp, Up, bs, TA, TB, acc);
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
      lu_lds<T, RT>(S, bs, piv, red_v, red_i, &s_piv, info, b);
      for (int e = tid; e < bs * bs; e += 256) {
        const int r = e / bs, c = e - r * bs;
        Di[e] = S[r * ld + c];
      }
      int* Pi = ipiv + b * pbs + (int64_t)i * pis;
      for (int r = tid; r < bs; r += 256) Pi[r] = piv[r] + 1;
      if (i < m - 1) {
        T* Ui = a.sup + b * a.ubs + (int64_t)i * a.uis;
        T* Li = a.sub + b * a.lbs + (int64_t)i * a.lis;
        auto Lat = [&](int r, int k) -> T { return S[r * ld + k]; };
        auto UTat = [&](int r, int k) -> T { return S[k * ld + r]; };
        for (int c0 = 0; c0 < bs; c0 += 32) {
          const int nc = (bs - c0) < 32 ? (bs - c0) : 32;
          // Ufac chunk: L^{-1} P sup_i[:, c0:c0+nc]
          for (int e = tid; e < bs * nc; e += 256) {
            const int r = e / nc, c = e - r * nc;
            X[e] = Ui[(int64_t)r * bs + c0 + c];
          }
          __syncthreads();
          apply_piv_lds<T>(piv, bs, X, nc);
          tri_solve_direct<T>(Lat, bs, false, true, X, nc);
          for (int e = tid; e < bs * nc; e += 256) {
            const int r = e / nc, c = e - r * nc;
            Ui[(int64_t)r * bs + c0 + c] = X[e];
          }
          __syncthreads();
          // Lfac chunk: rows c0..c0+nc of sub_i U^{-1}, i.e. U^T Z = sub_i^T
          for (int e = tid; e < bs * nc; e += 256) {
            const int r = e / nc, c = e - r * nc;
            X[e] = Li[(int64_t)(c0 + c) * bs + r];
          }
          __syncthreads();
          tri_solve_direct<T>(UTat, bs, false, false, X, nc);
          for (int e = tid; e < bs * nc; e += 256) {
            const int r = e / nc, c = e - r * nc;
            Li[(int64_t)(c0 + c) * bs + r] = X[e];
          }
          __syncthreads();
        }
      }
      __threadfence();
      __syncthreads();
    }
  }
}

// Stage the bs x bs row-major block src into LDS Mt (row stride bs+1).
template <typename T>
__device__ inline void stage_block(T* Mt, const T* src, int bs) {
  const int ld = bs + 1;
  for (int e = threadIdx.x; e < bs * bs; e += 256) {
    const int r = e / bs, c = e - r * bs;
    Mt[r * ld + c] = src[e];
  }
  __syncthreads();
}

// Xc[r][c] -= sum_k Mt[r][k] Y[k][c] (all operands in LDS).
template <typename T, int RPT>
__device__ inline void sub_mat_vec(const T* Mt, const T* Y, T* Xc, int bs, int nr) {
  const int tid = threadIdx.x;
  const int tc = tid % nr, tr = tid / nr, RG = 256 / nr;
  const int ld = bs + 1;
  T acc[RPT];
#pragma unroll
  for (int x = 0; x < RPT; ++x) acc[x] = T(0);
  for (int k = 0; k < bs; ++k) {
    const T y = Y[k * nr + tc];
#pragma unroll
    for (int x = 0; x < RPT; ++x) {
      const int r = tr + RG * x;
      if (r < bs) acc[x] += Mt[r * ld + k] * y;
    }
  }
#pragma unroll
  for (int x = 0; x < RPT; ++x) {
    const int r = tr + RG * x;
    if (r < bs) Xc[r * nr + tc] -= acc[x];
  }
  __syncthreads();
}

template <typename T, int RPT>
__global__ __launch_bounds__(256) void k_btri_lu_solve(int nb, int m, int bs, BtriArgs<T> a, const int* ipiv,
                                                       int64_t pbs, int64_t pis, T* f, int64_t fbs, int64_t ldf,
                                                       int nrhs, int NR) {
  extern __shared__ unsigned char smem_raw[];
  T* Mt = reinterpret_cast<T*>(smem_raw);  // bs x (bs+1): one block at a time
  T* Xc = Mt + bs * (bs + 1);              // bs x NR current block
  T* Y = Xc + bs * NR;                     // bs x NR previous block
  int* P = reinterpret_cast<int*>(Y + bs * NR);
  const int tid = threadIdx.x;
  const int nr = NR;
  const int ld = bs + 1;
  const int c0 = blockIdx.y * NR;
  const int ncols = (nrhs - c0) < NR ? (nrhs - c0) : NR;
  for (int64_t b = blockIdx.x; b < nb; b += gridDim.x) {
    T* fb = f + b * fbs;
    auto F = [&](int i, int r, int c) -> T* { return fb + (int64_t)(i * bs + r) * ldf + c0 + c; };
    auto Mat = [&](int r, int k) -> T { return Mt[r * ld + k]; };
    for (int i = 0; i < m; ++i) {
      for (int e = tid; e < bs * nr; e += 256) {
        const int r = e / nr, c = e - r * nr;
        Xc[e] = (c < ncols) ? *F(i, r, c) : T(0);
      }
      for (int r = tid; r < bs; r += 256) P[r] = ipiv[b * pbs + (int64_t)i * pis + r] - 1;
      if (i > 0) {
        stage_block<T>(Mt, a.sub + b * a.lbs + (int64_t)(i - 1) * a.lis, bs);
        sub_mat_vec<T, RPT>(Mt, Y, Xc, bs, nr);
      } else {
        __syncthreads();
      }
      apply_piv_lds<T>(P, bs, Xc, nr);
      stage_block<T>(Mt, a.diag + b * a.dbs + (int64_t)i * a.dis, bs);
      tri_solve_direct<T>(Mat, bs, false, true, Xc, nr);
      for (int e = tid; e < bs * nr; e += 256) {
        const int r = e / nr, c = e - r * nr;
        Y[e] = Xc[e];
        if (c < ncols) *F(i, r, c) = Xc[e];
      }
      __syncthreads();
    }
    for (int i = m - 1; i >= 0; --i) {
      if (i < m - 1) {
        for (int e = tid; e < bs * nr; e += 256) {
          const int r = e / nr, c = e - r * nr;
          Xc[e] = (c < ncols) ? *F(i, r, c) : T(0);
        }
        stage_block<T>(Mt, a.sup + b * a.ubs + (int64_t)i * a.uis, bs);
        sub_mat_vec<T, RPT>(Mt, Y, Xc, bs, nr);
      } else {
        for (int e = tid; e < bs * nr; e += 256) Xc[e] = Y[e];
      }
      stage_block<T>(Mt, a.diag + b * a.dbs + (int64_t)i * a.dis, bs);
      tri_solve_direct<T>(Mat, bs, true, false, Xc, nr);
      for (int e = tid; e < bs * nr; e += 256) {
        const int r = e / nr, c = e - r * nr;
        Y[e] = Xc[e];
        if (c < ncols) *F(i, r, c) = Xc[e];
      }
      __syncthreads();
    }
  }
}

template <typename T>
size_t lu_factor_lds(int bs) {
  const size_t tb = (size_t)KC * (bs + 1) > 64 * 65 ? (size_t)KC * (bs + 1) : 64 * 65;
  return sizeof(T) * ((size_t)bs * (bs + 1) + (size_t)bs * 32 + (size_t)bs * (KC + 1) + tb) + sizeof(int) * bs;
}

template <typename T>
size_t lu_solve_lds(int bs, int nr) {
  return sizeof(T) * ((size_t)bs * (bs + 1) + 2 * (size_t)bs * nr) + sizeof(int) * bs;
}

template <typename T>
size_t factor_lds(int bs) {
  return sizeof(T) * ((size_t)bs * (bs + 1) + (size_t)bs * (KC + 1) + (size_t)KC * (bs + 1)) + sizeof(int) * bs;
}

template <typename T>
size_t solve_lds(int bs, int nr) {
  return sizeof(T) * (2 * (size_t)bs * nr + (size_t)bs * (KC + 1));
}

constexpr size_t kLdsMax = 160 * 1024;

template <typename T, int RT>
hipError_t launch_factor(int nb, int m, int bs, BtriArgs<T> a, int* info, hipStream_t s) {
  const size_t lds = factor_lds<T>(bs);
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

template <typename T, int RT>
hipError_t launch_lu_factor(int nb, int m, int bs, BtriArgs<T> a, int* ipiv, int64_t pbs, int64_t pis, int* info,
                            hipStream_t s) {
  const size_t lds = lu_factor_lds<T>(bs);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_btri_lu_factor<T, RT>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_btri_lu_factor<T, RT>), dim3(nb < 65535 ? nb : 65535), dim3(256), lds, s, nb, m, bs, a,
                     ipiv, pbs, pis, info);
  return hipGetLastError();
}

template <typename T, int RPT>
hipError_t launch_lu_solve(int nb, int m, int bs, BtriArgs<T> a, const int* ipiv, int64_t pbs, int64_t pis, T* f,
                           int64_t fbs, int64_t ldf, int nrhs, int nr, hipStream_t s) {
  const size_t lds = lu_solve_lds<T>(bs, nr);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_btri_lu_solve<T, RPT>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  dim3 grid(nb < 65535 ? nb : 65535, (nrhs + nr - 1) / nr);
  hipLaunchKernelGGL((k_btri_lu_solve<T, RPT>), grid, dim3(256), lds, s, nb, m, bs, a, ipiv, pbs, pis, f, fbs, ldf,
                     nrhs, nr);
  return hipGetLastError();
}

}  // namespace

template <typename T>
bool btri_inv_fused_ok(int bs) {
  return (bs + 15) / 16 <= 10 && factor_lds<T>(bs) <= kLdsMax;
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

template <typename T>
bool btri_lu_fused_ok(int bs) {
  return (bs + 15) / 16 <= 10 && lu_factor_lds<T>(bs) <= kLdsMax;
}

template <typename T>
hipError_t btri_lu_factor_fused(int nb, int m, int bs, T* diag, int64_t dbs, int64_t dis, T* sub, int64_t lbs,
                                int64_t lis, T* sup, int64_t ubs, int64_t uis, int* ipiv, int64_t pbs, int64_t pis,
                                int* info, hipStream_t s) {
  BtriArgs<T> a{diag, dbs, dis, sub, lbs, lis, sup, ubs, uis};
  const int rt = (bs + 15) / 16;
  switch (rt) {
    case 1: return launch_lu_factor<T, 1>(nb, m, bs, a, ipiv, pbs, pis, info, s);
    case 2: return launch_lu_factor<T, 2>(nb, m, bs, a, ipiv, pbs, pis, info, s);
    case 3: case 4: return launch_lu_factor<T, 4>(nb, m, bs, a, ipiv, pbs, pis, info, s);
    case 5: case 6: return launch_lu_factor<T, 6>(nb, m, bs, a, ipiv, pbs, pis, info, s);
    case 7: case 8: return launch_lu_factor<T, 8>(nb, m, bs, a, ipiv, pbs, pis, info, s);
    default: return launch_lu_factor<T, 10>(nb, m, bs, a, ipiv, pbs, pis, info, s);
  }
}

template <typename T>
hipError_t btri_lu_solve_fused(int nb, int m, int bs, T* diag, int64_t dbs, int64_t dis, T* sub, int64_t lbs,
                               int64_t lis, T* sup, int64_t ubs, int64_t uis, const int* ipiv, int64_t pbs,
                               int64_t pis, T* f, int64_t fbs, int64_t ldf, int nrhs, hipStream_t s) {
  BtriArgs<T> a{diag, dbs, dis, sub, lbs, lis, sup, ubs, uis};
  int nr = nrhs >= 32 ? 32 : (nrhs >= 16 ? 16 : (nrhs >= 8 ? 8 : (nrhs >= 4 ? 4 : (nrhs >= 2 ? 2 : 1))));
  while (nr > 1 && lu_solve_lds<T>(bs, nr) > kLdsMax) nr >>= 1;
  const int rpt = (bs + 256 / nr - 1) / (256 / nr);
  if (lu_solve_lds<T>(bs, nr) > kLdsMax || rpt > 24) return hipErrorInvalidValue;
  if (rpt <= 1) return launch_lu_solve<T, 1>(nb, m, bs, a, ipiv, pbs, pis, f, fbs, ldf, nrhs, nr, s);
  if (rpt <= 2) return launch_lu_solve<T, 2>(nb, m, bs, a, ipiv, pbs, pis, f, fbs, ldf, nrhs, nr, s);
  if (rpt <= 4) return launch_lu_solve<T, 4>(nb, m, bs, a, ipiv, pbs, pis, f, fbs, ldf, nrhs, nr, s);
  if (rpt <= 8) return launch_lu_solve<T, 8>(nb, m, bs, a, ipiv, pbs, pis, f, fbs, ldf, nrhs, nr, s);
  if (rpt <= 13) return launch_lu_solve<T, 13>(nb, m, bs, a, ipiv, pbs, pis, f, fbs, ldf, nrhs, nr, s);
  return launch_lu_solve<T, 24>(nb, m, bs, a, ipiv, pbs, pis, f, fbs, ldf, nrhs, nr, s);
}

template <typename T>
hipError_t getrs_small(int batch, int n, int nrhs, View<T> LU, const int* ipiv, int64_t pbs, View<T> B,
                       hipStream_t s) {
  if (nrhs > 8) return hipErrorInvalidValue;
  const size_t lds = sizeof(T) * ((size_t)n * nrhs + 64 * 65) + sizeof(int) * n;
  if (lds > kLdsMax) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_getrs_small<T>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_getrs_small<T>, dim3(batch < 65535 ? batch : 65535), dim3(256), lds, s, batch, n, nrhs, LU,
                     ipiv, pbs, B);
  return hipGetLastError();
}

#define SLAB_BTRI_INST(T)                                                                                         \
  template bool btri_inv_fused_ok<T>(int);                                                                         \
  template hipError_t btri_inv_factor<T>(int, int, int, T*, int64_t, int64_t, T*, int64_t, int64_t, T*, int64_t,  \
                                         int64_t, int*, T*, int*, hipStream_t);                                    \
  template hipError_t btri_inv_solve<T>(int, int, int, T*, int64_t, int64_t, T*, int64_t, int64_t, T*, int64_t,   \
                                        int64_t, T*, int64_t, int64_t, int, T*, hipStream_t);                      \
  template hipError_t getrs_small<T>(int, int, int, View<T>, const int*, int64_t, View<T>, hipStream_t);          \
  template bool btri_lu_fused_ok<T>(int);                                                                          \
  template hipError_t btri_lu_factor_fused<T>(int, int, int, T*, int64_t, int64_t, T*, int64_t, int64_t, T*,      \
                                              int64_t, int64_t, int*, int64_t, int64_t, int*, hipStream_t);        \
  template hipError_t btri_lu_solve_fused<T>(int, int, int, T*, int64_t, int64_t, T*, int64_t, int64_t, T*,       \
                                             int64_t, int64_t, const int*, int64_t, int64_t, T*, int64_t, int64_t, \
                                             int, hipStream_t);
SLAB_BTRI_INST(double)
SLAB_BTRI_INST(float)

}  // namespace slab
