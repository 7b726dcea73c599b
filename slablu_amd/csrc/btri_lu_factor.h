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

// ---------------------------------------------------------------------------
// MFMA form of the factor kernel for blocks that fit LDS three times over
// (FP32, bs <= 112).  Ufac_i and Lfac_i stay in LDS after their solves, so
// block i+1's Schur product Lfac_i Ufac_i reads no global memory; both
// triangular solves run as blocked substitutions whose off-diagonal updates
// are MFMA tiles; every global transfer moves 16-byte vectors.
// ---------------------------------------------------------------------------

// Row stride of the LDS Ufac / Lfac^T blocks: at least bs and 16 mod 32, so
// the two rows a half-wavefront reads in one MFMA operand load fall in
// distinct banks.
__host__ __device__ inline int lu_mfma_ldx(int bs) { return bs <= 16 ? 16 : ((bs - 16 + 31) / 32) * 32 + 16; }

// inverses of the 16 x 16 diagonal blocks of both triangular factors, rows
// padded to 17
__host__ __device__ inline int lu_mfma_dinv(int bs) { return 2 * tile_rt(bs) * 16 * 17; }

template <typename T>
size_t lu_mfma_lds(int bs) {
  const size_t bp = 16 * (size_t)tile_rt(bs);
  return sizeof(T) * (bp * (bp + 1) + 2 * (size_t)bs * lu_mfma_ldx(bs) + lu_mfma_dinv(bs) + 64) +
         2 * sizeof(int) * bs;
}

template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  typedef float4 t;
  static constexpr int n = 4;
};
template <>
struct Vec16<double> {
  typedef double2 t;
  static constexpr int n = 2;
};

__device__ inline bool al16d(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// S (row stride ld) <- the bs x bs row-major block src
template <typename T>
__device__ inline void load_block_v(T* S, int ld, const T* src, int bs) {
  typedef typename Vec16<T>::t vt;
  constexpr int V = Vec16<T>::n;
  if ((bs % V) == 0 && al16d(src)) {
    const int nv = bs / V, n = bs * nv;
    const float rnv = 1.0f / (float)nv;
    const vt* s4 = reinterpret_cast<const vt*>(src);
    for (int e0 = threadIdx.x; e0 < n; e0 += 256 * 8) {
      vt v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + 256 * u;
        v[u] = s4[e < n ? e : n - 1];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + 256 * u;
        if (e < n) {
          const int r = idiv(e, rnv), c = V * (e - r * nv);
          const T* pv = reinterpret_cast<const T*>(&v[u]);
#pragma unroll
          for (int t = 0; t < V; ++t) S[r * ld + c + t] = pv[t];
        }
      }
    }
    __syncthreads();
  } else {
    stage_block<T>(S, src, bs, 1.0f / (float)bs, ld);
  }
}

// dst (bs x bs row-major) <- S (row stride ld)
template <typename T>
__device__ inline void store_block_v(T* dst, const T* S, int ld, int bs) {
  typedef typename Vec16<T>::t vt;
  constexpr int V = Vec16<T>::n;
  if ((bs % V) == 0 && al16d(dst)) {
    const int nv = bs / V, n = bs * nv;
    const float rnv = 1.0f / (float)nv;
    vt* d4 = reinterpret_cast<vt*>(dst);
    for (int e = threadIdx.x; e < n; e += 256) {
      const int r = idiv(e, rnv), c = V * (e - r * nv);
      vt v;
      T* pv = reinterpret_cast<T*>(&v);
#pragma unroll
      for (int t = 0; t < V; ++t) pv[t] = S[r * ld + c + t];
      d4[e] = v;
    }
  } else {
    const float rbs = 1.0f / (float)bs;
    for (int e = threadIdx.x; e < bs * bs; e += 256) {
      const int r = idiv(e, rbs);
      dst[e] = S[r * ld + e - r * bs];
    }
  }
}

// XU[r][c] = U[Q[r]][c] (rows of the coupling block through the pivot
// permutation) and XL[r][c] = L[c][r] (the coupling block transposed), row
// stride ldx
template <typename T>
__device__ inline void gather_couplings(T* XU, T* XL, int ldx, const T* U, const T* L, const int* Q, int bs) {
  typedef typename Vec16<T>::t vt;
  constexpr int V = Vec16<T>::n;
  if ((bs % V) == 0 && al16d(U) && al16d(L)) {
    const int nv = bs / V, n = bs * nv;
    const float rnv = 1.0f / (float)nv;
    for (int e0 = threadIdx.x; e0 < n; e0 += 256 * 4) {
      vt u[4], l[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = e0 + 256 * q;
        const int ec = e < n ? e : n - 1;
        const int r = idiv(ec, rnv), cv = ec - r * nv;
        u[q] = *reinterpret_cast<const vt*>(U + (int64_t)Q[r] * bs + V * cv);
        l[q] = *reinterpret_cast<const vt*>(L + (int64_t)r * bs + V * cv);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = e0 + 256 * q;
        if (e < n) {
          const int r = idiv(e, rnv), c = V * (e - r * nv);
          const T* pu = reinterpret_cast<const T*>(&u[q]);
          const T* pl = reinterpret_cast<const T*>(&l[q]);
#pragma unroll
          for (int t = 0; t < V; ++t) {
            XU[r * ldx + c + t] = pu[t];
            XL[(c + t) * ldx + r] = pl[t];
          }
        }
      }
    }
  } else {
    const float rbs = 1.0f / (float)bs;
    for (int e = threadIdx.x; e < bs * bs; e += 256) {
      const int r = idiv(e, rbs), c = e - r * bs;
      XU[r * ldx + c] = U[(int64_t)Q[r] * bs + c];
      XL[c * ldx + r] = L[e];
    }
  }
  __syncthreads();
}

// U[r][c] = XU[r][c], L[r][c] = XL[c][r]
template <typename T>
__device__ inline void store_couplings(T* U, T* L, const T* XU, const T* XL, int ldx, int bs) {
  typedef typename Vec16<T>::t vt;
  constexpr int V = Vec16<T>::n;
  if ((bs % V) == 0 && al16d(U) && al16d(L)) {
    const int nv = bs / V, n = bs * nv;
    const float rnv = 1.0f / (float)nv;
    for (int e = threadIdx.x; e < n; e += 256) {
      const int r = idiv(e, rnv), c = V * (e - r * nv);
      vt u, l;
      T* pu = reinterpret_cast<T*>(&u);
      T* pl = reinterpret_cast<T*>(&l);
#pragma unroll
      for (int t = 0; t < V; ++t) {
        pu[t] = XU[r * ldx + c + t];
        pl[t] = XL[(c + t) * ldx + r];
      }
      *reinterpret_cast<vt*>(U + (int64_t)r * bs + c) = u;
      *reinterpret_cast<vt*>(L + (int64_t)r * bs + c) = l;
    }
  } else {
    const float rbs = 1.0f / (float)bs;
    for (int e = threadIdx.x; e < bs * bs; e += 256) {
      const int r = idiv(e, rbs), c = e - r * bs;
      U[e] = XU[r * ldx + c];
      L[e] = XL[c * ldx + r];
    }
  }
}

// S[0:bs, 0:bs] -= XL^T XU (both bs x bs, row stride ldx, RT = ceil(bs/16)
// tiles per side): wavefront w owns column tiles w and w + 4 and every row
// tile, so each K step reads one XU fragment and RT XL fragments, then
// issues RT independent MFMAs.
template <typename T, int RT>
__device__ __attribute__((always_inline)) void schur_mfma(T* S, int ld, const T* XL, const T* XU, int ldx, int bs) {
  typedef typename Mfma<T>::acc_t acc_t;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int bm1 = bs - 1;
  const int i16 = lane & 15, kq = lane >> 4;
  int aoff[RT];
#pragma unroll
  for (int j = 0; j < RT; ++j) {
    const int r = j * 16 + i16;
    aoff[j] = r < bm1 ? r : bm1;
  }
  const int klast = ((bs - 1) >> 2) << 2;
  for (int ct = wave; ct < RT; ct += 4) {
    acc_t acc[RT];
#pragma unroll
    for (int j = 0; j < RT; ++j) acc[j] = acc_t{0, 0, 0, 0};
    const int col = ct * 16 + i16;
    const int colc = col < bm1 ? col : bm1;
    // K step k0's fragments (K past bs as zeros)
    auto fetch = [&](int k0, T (&av)[RT], T& b) __attribute__((always_inline)) {
      const int k = k0 + kq;
      const bool kin = k < bs;
      const int kr = (kin ? k : bm1) * ldx;
      const T bv = XU[kr + colc];
      b = kin ? bv : T(0);
#pragma unroll
      for (int j = 0; j < RT; ++j) {
        const T v = XL[kr + aoff[j]];
        av[j] = kin ? v : T(0);
      }
    };
    // the next step's fragments load while this step's MFMAs run
    T an[RT], bn;
    fetch(0, an, bn);
    for (int k0 = 0; k0 < bs; k0 += 4) {
      T av[RT];
#pragma unroll
      for (int j = 0; j < RT; ++j) av[j] = an[j];
      const T b = bn;
      fetch(k0 + 4 < klast ? k0 + 4 : klast, an, bn);
#pragma unroll
      for (int j = 0; j < RT; ++j) acc[j] = Mfma<T>::mma(av[j], b, acc[j]);
    }
    // rows and columns past bs land in the padding of S (never read)
#pragma unroll
    for (int j = 0; j < RT; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) S[(j * 16 + Mfma<T>::row(lane, q)) * ld + col] -= acc[j][q];
  }
  __syncthreads();
}

// Inverses of the 16 x 16 diagonal blocks of M1 (unit lower, through at1)
// and M2 (lower, through at2), both n x n: Dinv[(m * RT + kb) * 272 + r * 17
// + c], one thread per (matrix, block, column) solving for a unit vector by
// substitution.  Entries past the block's last row or column are zero.
template <typename T, int RT, typename A1, typename A2>
__device__ __attribute__((always_inline)) void diag_block_inverses(A1 at1, A2 at2, int n, T* Dinv) {
  const int nm1 = n - 1;
  for (int task = threadIdx.x; task < 2 * RT * 16; task += 256) {
    const bool second = task >= RT * 16;
    const int u = second ? task - RT * 16 : task;
    const int kb = u >> 4, c = u & 15;
    const int d0 = kb * 16;
    const int tb = (n - d0) < 16 ? (n - d0) : 16;
    T x[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) x[j] = (j == c) ? T(1) : T(0);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int jj = (d0 + j) < nm1 ? (d0 + j) : nm1;
      const T xj = second ? x[j] / at2(jj, jj) : x[j];
      x[j] = xj;
#pragma unroll
      for (int r = j + 1; r < 16; ++r) {
        const int rr = (d0 + r) < nm1 ? (d0 + r) : nm1;
        x[r] -= (second ? at2(rr, jj) : at1(rr, jj)) * xj;
      }
    }
    T* D = Dinv + (second ? RT + kb : kb) * 272;
#pragma unroll
    for (int r = 0; r < 16; ++r) D[r * 17 + c] = ((r < tb) & (c < tb)) ? x[r] : T(0);
  }
  __syncthreads();
}

// X <- M^{-1} X on column tiles ct0, ct0 + 2, ct0 + 4, ct0 + 6 (those below
// RT) of the n x n X (row stride ldx), M lower with diagonal-block inverses
// D (RT blocks of 16 x 17): per 16-row block, Y = X_k - M[k, 0:k) X[0:k)
// by MFMA (one M fragment, four X fragments, four independent products per
// K step), Y back into X's block rows, then X_k = Dinv_k Y by MFMA.  Only
// this wavefront's columns are read or written, so no workgroup barrier
// is needed.
template <typename T, int RT, typename At>
__device__ __attribute__((always_inline)) void tri_solve_cols_mfma(At at, const T* D, T* X, int ldx, int n,
                                                                   int ct0, T* junk) {
  typedef typename Mfma<T>::acc_t acc_t;
  const int lane = threadIdx.x & 63;
  const int i16 = lane & 15, kq = lane >> 4;
  const int nm1 = n - 1;
  int col[4], colc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    col[q] = (ct0 + 2 * q) * 16 + i16;
    colc[q] = col[q] < nm1 ? col[q] : nm1;
  }
  for (int kb = 0; kb < RT; ++kb) {
    const int d0 = kb << 4;
    const int rowc = (d0 + i16) < nm1 ? (d0 + i16) : nm1;
    acc_t acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = acc_t{0, 0, 0, 0};
    // d0 is a multiple of 16: 16 K values per round, all 20 fragments
    // loaded before the round's 16 MFMAs
    for (int k0 = 0; k0 < d0; k0 += 16) {
      T a[4], bv[4][4];
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const int k = k0 + 4 * ks + kq;
        a[ks] = at(rowc, k);
#pragma unroll
        for (int q = 0; q < 4; ++q) bv[ks][q] = X[k * ldx + colc[q]];
      }
#pragma unroll
      for (int ks = 0; ks < 4; ++ks)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = Mfma<T>::mma(a[ks], bv[ks][q], acc[q]);
    }
    // Y = X_k - acc into the block rows; entries past n - 1 go to a junk
    // slot, so the 16 read-modify-writes issue back to back
    T* dst[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) {
        const int r = d0 + Mfma<T>::row(lane, qq);
        dst[q][qq] = ((r < n) & (col[q] < n)) ? X + r * ldx + col[q] : junk;
      }
    }
    {
      T y[4][4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) y[q][qq] = *dst[q][qq];
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) *dst[q][qq] = y[q][qq] - acc[q][qq];
    }
    __builtin_amdgcn_wave_barrier();
    // X_k = Dinv_k Y (Y rows past n - 1 enter as zeros)
    const T* Dk = D + kb * 272;
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int k = 4 * ks + kq;
      const bool kin = d0 + k < n;
      const int kr = (kin ? d0 + k : nm1) * ldx;
      const T a = Dk[i16 * 17 + k];
      T bv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const T v = X[kr + colc[q]];
        bv[q] = kin ? v : T(0);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = Mfma<T>::mma(a, bv[q], acc[q]);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) *dst[q][qq] = acc[q][qq];
    __builtin_amdgcn_wave_barrier();
  }
}

// X1 <- M1^{-1} X1 (M1 unit lower, through at1) and X2 <- M2^{-1} X2 (M2
// lower, through at2), n x n each with row stride ldx, RT = ceil(n/16) <= 8
// tiles: the diagonal-block inverses, then wavefronts 0-1 solve X1 and 2-3
// X2 on alternate column tiles, each on its own columns only.
template <typename T, int RT, typename A1, typename A2>
__device__ __attribute__((always_inline)) void tri_solve2_mfma(A1 at1, A2 at2, int n, T* X1, T* X2, int ldx,
                                                               T* Dinv, T* junk) {
  diag_block_inverses<T, RT>(at1, at2, n, Dinv);
  const int wave = threadIdx.x >> 6;
  if (wave < 2) tri_solve_cols_mfma<T, RT>(at1, Dinv, X1, ldx, n, wave & 1, junk + 16 * wave);
  else tri_solve_cols_mfma<T, RT>(at2, Dinv + RT * 272, X2, ldx, n, wave & 1, junk + 16 * wave);
  __syncthreads();
}

template <typename T, int RT>
__global__ __launch_bounds__(256) void k_btri_lu_factor_mfma(int bs, ChainSet<T> c0, ChainSet<T> c1) {
  constexpr int BP = 16 * RT;
  static_assert(RT <= 7, "Schur product tiles");
  extern __shared__ unsigned char smem_raw[];
  T* S = reinterpret_cast<T*>(smem_raw);  // BP x (BP+1), padded diagonal block
  const int ldx = lu_mfma_ldx(bs);
  T* XU = S + BP * (BP + 1);              // Ufac of the current block
  T* XL = XU + bs * ldx;                  // Lfac^T of the current block
  T* Dinv = XL + bs * ldx;                // diagonal-block inverses of L and U^T
  T* junk = Dinv + lu_mfma_dinv(bs);      // 64 entries: targets of masked-off stores
  int* piv = reinterpret_cast<int*>(junk + 64);
  int* Q = piv + bs;
  __shared__ __align__(16) int wsp[6 * 16 + 2 * 16 * (int)(sizeof(T) / sizeof(int))];
  const int ld = BP + 1;
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
      load_block_v<T>(S, ld, Di, bs);
      // Lfac_{i-1} Ufac_{i-1}, both still in LDS from the previous block
      if (i > 0) schur_mfma<T, RT>(S, ld, XL, XU, ldx, bs);
      lu_lds_blocked<T, RT>(S, bs, ld, piv, wsp, a.info, b);
      store_block_v<T>(Di, S, ld, bs);
      int* Pi = a.ipiv + b * a.pbs + (int64_t)i * a.pis;
      for (int r = threadIdx.x; r < bs; r += 256) Pi[r] = piv[r] + 1;
      if (i < m - 1 || a.open) {
        T* Ui = a.sup + b * a.ubs + (int64_t)i * a.uis;
        T* Li = a.sub + b * a.lbs + (int64_t)i * a.lis;
        piv_to_perm(piv, bs, Q);
        gather_couplings<T>(XU, XL, ldx, Ui, Li, Q, bs);
        // Ufac = L^{-1} (P sup_i);  U^T Lfac^T = sub_i^T
        tri_solve2_mfma<T, RT>(Lat, UTat, bs, XU, XL, ldx, Dinv, junk);
        store_couplings<T>(Ui, Li, XU, XL, ldx, bs);
      }
    }
    __syncthreads();
  }
}

template <typename T>
bool lu_mfma_ok(int bs) {
  return sizeof(T) == 4 && tile_rt(bs) <= 7 && lu_mfma_lds<T>(bs) + 1024 <= kLdsMax;
}

template <typename T, int RT>
hipError_t launch_lu_factor(int bs, const ChainSet<T>& c0, const ChainSet<T>& c1, hipStream_t s) {
  if constexpr (RT <= 7 && sizeof(T) == 4) {
    if (lu_mfma_ok<T>(bs)) {
      const size_t lds = lu_mfma_lds<T>(bs);
      hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_btri_lu_factor_mfma<T, RT>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
      const int nt = c0.nchains + c1.nchains;
      if (nt <= 0) return hipSuccess;
      hipLaunchKernelGGL((k_btri_lu_factor_mfma<T, RT>), dim3(nt < 65535 ? nt : 65535), dim3(256), lds, s, bs, c0,
                         c1);
      return hipGetLastError();
    }
  }
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
