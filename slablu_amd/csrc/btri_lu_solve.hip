// Forward and backward block sweeps with the factors of the block LU (the
// lusolve_sweep of src_solver/block_tridiag_solver_dense.py:179-203).
#include "lds_dev.h"

namespace slab {

namespace {

// One workgroup per (chain, NR-column tile) for both sweeps.  The sweep is a
// fixed sequence of 4m - 2 stages, each with one block resident in LDS:
//   forward  DIAGL(0), SUB(0), DIAGL(1), ..., SUB(m-2), DIAGL(m-1)
//   backward DIAGU(m-1), SUP(m-2), DIAGU(m-2), ..., SUP(0), DIAGU(0)
// DIAGL(i): permute (composed pivots), unit-lower solve, store y_i;
// SUB(i): x_{i+1} = f_{i+1} - Lfac_i y_i; DIAGU(i): upper solve, store x_i;
// SUP(i): y_i - Ufac_i x_{i+1}.  While a stage computes, the next stage's
// block (PF elements per thread; any rest is loaded at hand-over), its
// right-hand-side rows and its pivots are already being loaded into
// registers, so global latency is off the critical path.
template <typename T, int RPT, int PF>
__global__ __launch_bounds__(256) void k_btri_lu_solve(int nb, int m, int bs, BtriArgs<T> a, const int* ipiv,
                                                       int64_t pbs, int64_t pis, T* f, int64_t fbs, int64_t ldf,
                                                       int nrhs, int NR) {
  extern __shared__ unsigned char smem_raw[];
  const int nr = NR;
  const int ldx = nr == 1 ? 1 : nr + 1;
  T* Mt = reinterpret_cast<T*>(smem_raw);  // bs x (bs+1): one block at a time
  T* B0 = Mt + bs * (bs + 1);              // bs x ldx
  T* B1 = B0 + bs * ldx;                   // bs x ldx
  T* W = B1 + bs * ldx;                    // bs: split-k partial sums
  int* P = reinterpret_cast<int*>(W + bs);
  int* Q = P + bs;
  const int tid = threadIdx.x;
  const int ld = bs + 1;
  const int lgnr = 31 - __clz(nr);
  const int c0 = blockIdx.y * NR;
  const int ncols = (nrhs - c0) < NR ? (nrhs - c0) : NR;
  const int nn = bs * bs, nf = bs * nr;
  const float rbs = 1.0f / (float)bs;
  const int nst = 4 * m - 2;
  auto Mat = [&](int r, int k) -> T { return Mt[r * ld + k]; };
  // stage s -> kind (0 DIAGL, 1 SUB, 2 DIAGU, 3 SUP) and block index i
  auto decode = [&](int s, int& i) -> int {
    if (s < 2 * m - 1) {
      if (s & 1) { i = (s - 1) >> 1; return 1; }
      i = s >> 1;
      return 0;
    }
    const int t = s - (2 * m - 1);
    if (t & 1) { i = m - 1 - ((t + 1) >> 1); return 3; }
    i = m - 1 - (t >> 1);
    return 2;
  };
  for (int64_t b = blockIdx.x; b < nb; b += gridDim.x) {
    T* fb = f + b * fbs;
    const int* pb = ipiv + b * pbs;
    T* Xc = B0;
    T* Y = B1;
    auto block_src = [&](int k, int i) -> const T* {
      if (k == 1) return a.sub + b * a.lbs + (int64_t)i * a.lis;
      if (k == 3) return a.sup + b * a.ubs + (int64_t)i * a.uis;
      return a.diag + b * a.dbs + (int64_t)i * a.dis;
    };
    auto f_row = [&](int k, int i) -> int {  // block row of f loaded into Xc
      if (k == 0) return i == 0 ? 0 : -1;
      if (k == 1) return i + 1;
      if (k == 3) return i;
      return -1;
    };
    T pf[PF], pfx[RPT];
    int pfp[2];
    auto prefetch = [&](int s) {
      int i;
      const int k = decode(s, i);
      if (!(k == 2 && i == m - 1)) {
        const T* src = block_src(k, i);
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const int e = tid + 256 * u;
          pf[u] = (e < nn) ? src[e] : T(0);
        }
      }
      const int fi = f_row(k, i);
      if (fi >= 0) {
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
          const int e = tid + 256 * u, r = e >> lgnr, c = e & (nr - 1);
          pfx[u] = (e < nf && c < ncols) ? fb[(int64_t)(fi * bs + r) * ldf + c0 + c] : T(0);
        }
      }
      if (k == 0) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int r = tid + 256 * u;
          pfp[u] = (r < bs) ? pb[(int64_t)i * pis + r] - 1 : 0;
        }
      }
    };
    auto commit = [&](int s) {
      int i;
      const int k = decode(s, i);
      if (!(k == 2 && i == m - 1)) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const int e = tid + 256 * u;
          if (e < nn) Mt[e + idiv(e, rbs)] = pf[u];
        }
        const T* src = block_src(k, i);
        for (int e = tid + 256 * PF; e < nn; e += 256) Mt[e + idiv(e, rbs)] = src[e];
      }
      if (f_row(k, i) >= 0) {
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
          const int e = tid + 256 * u, r = e >> lgnr, c = e & (nr - 1);
          if (e < nf) Xc[r * ldx + c] = pfx[u];
        }
      }
      if (k == 0) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int r = tid + 256 * u;
          if (r < bs) P[r] = pfp[u];
        }
      }
    };
    auto store_f = [&](int i, const T* Z) {
      for (int e = tid; e < nf; e += 256) {
        const int r = e >> lgnr, c = e & (nr - 1);
        if (c < ncols) fb[(int64_t)(i * bs + r) * ldf + c0 + c] = Z[r * ldx + c];
      }
    };
    prefetch(0);
    commit(0);
    __syncthreads();
    for (int s = 0; s < nst; ++s) {
      if (s + 1 < nst) prefetch(s + 1);
      int i;
      const int k = decode(s, i);
      if (k == 0) {
        piv_to_perm(P, bs, Q);
        for (int e = tid; e < nf; e += 256) {
          const int r = e >> lgnr, c = e & (nr - 1);
          Y[r * ldx + c] = Xc[Q[r] * ldx + c];
        }
        __syncthreads();
        tri_solve<T, false, true>(Mat, bs, Y, ldx, nr);
        store_f(i, Y);
      } else if (k == 1 || k == 3) {
        if (nr == 1)
          sub_mat_vec1<T>(Mt, Y, Xc, bs, W);
        else
          sub_mat_vec<T, RPT>(Mt, Y, Xc, bs, nr, ldx);
      } else {
        T* Z = (i == m - 1) ? Y : Xc;
        tri_solve<T, true, false>(Mat, bs, Z, ldx, nr);
        store_f(i, Z);
        if (i != m - 1) {
          T* t = Xc;
          Xc = Y;
          Y = t;
        }
      }
      __syncthreads();
      if (s + 1 < nst) commit(s + 1);
      __syncthreads();
    }
  }
}

template <typename T, int RPT, int PF>
hipError_t launch_lu_solve(int nb, int m, int bs, BtriArgs<T> a, const int* ipiv, int64_t pbs, int64_t pis, T* f,
                           int64_t fbs, int64_t ldf, int nrhs, int nr, hipStream_t s) {
  const size_t lds = lu_solve_lds<T>(bs, nr);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_btri_lu_solve<T, RPT, PF>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  dim3 grid(nb < 65535 ? nb : 65535, (nrhs + nr - 1) / nr);
  hipLaunchKernelGGL((k_btri_lu_solve<T, RPT, PF>), grid, dim3(256), lds, s, nb, m, bs, a, ipiv, pbs, pis, f, fbs,
                     ldf, nrhs, nr);
  return hipGetLastError();
}

template <typename T, int RPT>
hipError_t launch_lu_solve_pf(int nb, int m, int bs, BtriArgs<T> a, const int* ipiv, int64_t pbs, int64_t pis,
                              T* f, int64_t fbs, int64_t ldf, int nrhs, int nr, hipStream_t s) {
  if ((int64_t)bs * bs <= 256 * 8)
    return launch_lu_solve<T, RPT, 8>(nb, m, bs, a, ipiv, pbs, pis, f, fbs, ldf, nrhs, nr, s);
  return launch_lu_solve<T, RPT, 40>(nb, m, bs, a, ipiv, pbs, pis, f, fbs, ldf, nrhs, nr, s);
}

}  // namespace

template <typename T>
hipError_t btri_lu_solve_fused(int nb, int m, int bs, T* diag, int64_t dbs, int64_t dis, T* sub, int64_t lbs,
                               int64_t lis, T* sup, int64_t ubs, int64_t uis, const int* ipiv, int64_t pbs,
                               int64_t pis, T* f, int64_t fbs, int64_t ldf, int nrhs, hipStream_t s) {
  BtriArgs<T> a{diag, dbs, dis, sub, lbs, lis, sup, ubs, uis};
  int nr = nrhs >= 32 ? 32 : (nrhs >= 16 ? 16 : (nrhs >= 8 ? 8 : (nrhs >= 4 ? 4 : (nrhs >= 2 ? 2 : 1))));
  while (nr > 1 && lu_solve_lds<T>(bs, nr) > kLdsMax) nr >>= 1;
  const int rpt = (bs + 256 / nr - 1) / (256 / nr);
  if (lu_solve_lds<T>(bs, nr) > kLdsMax || rpt > 24 || bs > 512) return hipErrorInvalidValue;
  if (rpt <= 1) return launch_lu_solve_pf<T, 1>(nb, m, bs, a, ipiv, pbs, pis, f, fbs, ldf, nrhs, nr, s);
  if (rpt <= 2) return launch_lu_solve_pf<T, 2>(nb, m, bs, a, ipiv, pbs, pis, f, fbs, ldf, nrhs, nr, s);
  if (rpt <= 4) return launch_lu_solve_pf<T, 4>(nb, m, bs, a, ipiv, pbs, pis, f, fbs, ldf, nrhs, nr, s);
  if (rpt <= 8) return launch_lu_solve_pf<T, 8>(nb, m, bs, a, ipiv, pbs, pis, f, fbs, ldf, nrhs, nr, s);
  if (rpt <= 13) return launch_lu_solve_pf<T, 13>(nb, m, bs, a, ipiv, pbs, pis, f, fbs, ldf, nrhs, nr, s);
  return launch_lu_solve_pf<T, 24>(nb, m, bs, a, ipiv, pbs, pis, f, fbs, ldf, nrhs, nr, s);
}

#define SLAB_BTRI_LU_SOLVE_INST(T)                                                                                \
  template hipError_t btri_lu_solve_fused<T>(int, int, int, T*, int64_t, int64_t, T*, int64_t, int64_t, T*,       \
                                             int64_t, int64_t, const int*, int64_t, int64_t, T*, int64_t, int64_t, \
                                             int, hipStream_t);
SLAB_BTRI_LU_SOLVE_INST(double)
SLAB_BTRI_LU_SOLVE_INST(float)

}  // namespace slab
