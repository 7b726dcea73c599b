// Device building blocks shared by the block-tridiagonal and small-solve
// kernels: LDS-resident dense operations for one 256-thread workgroup
// (matrix products staged through LDS, LU with partial pivoting, triangular
// solves, pivot permutations) and the LDS size rules of the fused kernels.
#pragma once

#include "slab_internal.h"
#include "wave_ops.h"

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

// One set of chains for the fused LU-form kernels: base pointers with the
// stride between chains (*bs) and between consecutive blocks of a chain
// (*is; negative to walk a chain backwards), m blocks per chain.  sub_i
// couples block i+1 to block i and sup_i block i to block i+1 in chain
// order; open: the factorization also forms Ufac/Lfac of the last block
// (its couplings reach a block past the end of the chain).
template <typename T>
struct ChainSet {
  T* diag; int64_t dbs, dis;
  T* sub; int64_t lbs, lis;
  T* sup; int64_t ubs, uis;
  int* ipiv; int64_t pbs, pis;
  int* info;
  T* f; int64_t fbs, fis;
  int m, nchains, open;
  // inverse-factor sweeps when set: L^{-1} P of each diagonal block (diag's
  // strides), with diag holding U^{-1}
  const T* linv;
};

// The chain set that walks blocks m-1 .. m-len of each chain backwards:
// diag and pivots mirrored, sub'_t = sup_{m-2-t}, sup'_t = sub_{m-2-t}.
template <typename T>
inline ChainSet<T> mirrored(int nb, int m, int len, int bs, T* diag, int64_t dbs, int64_t dis, T* sub, int64_t lbs,
                     int64_t lis, T* sup, int64_t ubs, int64_t uis, int* ipiv, int64_t pbs, int64_t pis, int* info,
                     T* f, int64_t fbs, int64_t ldf, const T* linv = nullptr) {
  ChainSet<T> c{};
  c.diag = diag + (int64_t)(m - 1) * dis; c.dbs = dbs; c.dis = -dis;
  c.linv = linv != nullptr ? linv + (int64_t)(m - 1) * dis : nullptr;
  c.sub = sup + (int64_t)(m - 2) * uis; c.lbs = ubs; c.lis = -uis;
  c.sup = sub + (int64_t)(m - 2) * lis; c.ubs = lbs; c.uis = -lis;
  c.ipiv = ipiv + (int64_t)(m - 1) * pis; c.pbs = pbs; c.pis = -pis;
  c.info = info;
  if (f != nullptr) {
    c.f = f + (int64_t)(m - 1) * bs * ldf; c.fbs = fbs; c.fis = -(int64_t)bs * ldf;
  }
  c.m = len; c.nchains = len > 0 ? nb : 0; c.open = 1;
  return c;
}

// acc[a][q] += sum_k A(r_a, k) B(k, c_q) for bs x bs row-major A and B in
// global memory; thread (ty, tx) of a 16 x 16 arrangement owns rows
// ty + 16a and columns tx + 16q (a, q < RT).  K runs in KC-wide chunks
// staged through LDS (TA: 16RT x (KC+1), TB: KC x (16RT+1)), entries past
// bs staged as zeros so the inner loop has no bounds tests, and the next
// chunk is loaded into registers while the current one is multiplied.
template <typename T, int RT>
__device__ void mm_global_global(const T* A, const T* B, int bs, T* TA, T* TB, T (&acc)[RT][RT]) {
  constexpr int BP = 16 * RT;
  constexpr int NE = (BP * KC + 255) / 256;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int bm1 = bs - 1;
  T va[NE], vb[NE];
  // clamped addresses, masks applied when the chunk is written to LDS: the
  // next chunk's loads stay in flight across the current chunk's products
  auto load = [&](int k0) {
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      const int e = tid + 256 * u;
      const int ra = e / KC, ka = k0 + e % KC;
      va[u] = A[(int64_t)(ra < bm1 ? ra : bm1) * bs + (ka < bm1 ? ka : bm1)];
      const int kb = k0 + e / BP, cb = e % BP;
      vb[u] = B[(int64_t)(kb < bm1 ? kb : bm1) * bs + (cb < bm1 ? cb : bm1)];
    }
  };
  load(0);
  for (int k0 = 0; k0 < bs; k0 += KC) {
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      const int e = tid + 256 * u;
      if (e < BP * KC) {
        const int ra = e / KC, ka = k0 + e % KC, kb = k0 + e / BP, cb = e % BP;
        TA[(e / KC) * (KC + 1) + e % KC] = ((ra < bs) & (ka < bs)) ? va[u] : T(0);
        TB[(e / BP) * (BP + 1) + e % BP] = ((kb < bs) & (cb < bs)) ? vb[u] : T(0);
      }
    }
    __syncthreads();
    if (k0 + KC < bs) load(k0 + KC);
#pragma unroll 4
    for (int kk = 0; kk < KC; ++kk) {
      T av[RT], bv[RT];
#pragma unroll
      for (int a = 0; a < RT; ++a) av[a] = TA[(ty + 16 * a) * (KC + 1) + kk];
#pragma unroll
      for (int q = 0; q < RT; ++q) bv[q] = TB[kk * (BP + 1) + tx + 16 * q];
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

// e / d for 0 <= e < 2^21 with rd = 1.0f / d: (e + 0.5) / d lies at least
// 0.5 / d from an integer and the two single-precision roundings move it by
// less than 2^-22 (e + 0.5) / d, so the truncation is exact.
__device__ inline int idiv(int e, float rd) { return (int)(((float)e + 0.5f) * rd); }

// Q[r] = the row of X that lands in row r when the interchanges row k <->
// row P[k] (P[k] >= k, 0-based) are applied for k = 0..n-1 in order; thread
// v follows where row v ends up, so no step is sequential across threads.
__device__ inline void piv_to_perm(const int* P, int n, int* Q) {
  for (int v = threadIdx.x; v < n; v += 256) {
    int pos = v;
#pragma unroll 8
    for (int k = 0; k < n; ++k) {
      const int pk = P[k];
      pos = (pos == k) ? pk : ((pos == pk) ? k : pos);
    }
    Q[pos] = v;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// LU form (backward stable; the reference's block LU): diag_i <- packed LU of
// D_i - Lfac_{i-1} Ufac_{i-1}, sub_i <- Lfac_i = sub_i U_i^{-1}, sup_i <-
// Ufac_i = L_i^{-1} P_i^T sup_i, ipiv (b, m, bs) 1-based.
// ---------------------------------------------------------------------------

// Triangular solve M X = X in LDS for nc <= NC <= 8 columns (X row stride
// ldx): wavefront w substitutes columns w, w + 4, ... on each 64-row
// diagonal block with lane = row and readlane broadcasts (non-unit diagonals
// through their reciprocals), then all threads update the rows beyond the
// block, one thread per row.  M(r, k) is read through `at` (LDS).  Every LDS
// read in the loops is unconditional (clamped indices, masked results) so
// that reads are issued back to back.
template <typename T, bool UPPER, bool UNIT, int NC, typename At>
__device__ void tri_solve_wave(At at, int n, T* X, int ldx, int nc) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nblk = (n + 63) >> 6;
  const int cm1 = nc - 1;
  for (int bi = 0; bi < nblk; ++bi) {
    const int blk = UPPER ? nblk - 1 - bi : bi;
    const int d0 = blk << 6;
    const int tb = (n - d0) < 64 ? (n - d0) : 64;
    const bool act = lane < tb;
    const int rl = d0 + (act ? lane : tb - 1);
    for (int c = wave; c < nc; c += 4) {
      T x = X[rl * ldx + c];
      const T rinv = UNIT ? T(1) : T(1) / at(rl, rl);
      if (UPPER) {
#pragma unroll 8
        for (int j = tb - 1; j >= 0; --j) {
          const T a = at(rl, d0 + j);
          const T xj = UNIT ? rdlane(x, j) : rdlane(x, j) * rdlane(rinv, j);
          x = (lane == j) ? xj : ((lane < j) ? x - a * xj : x);
        }
      } else {
#pragma unroll 8
        for (int j = 0; j < tb; ++j) {
          const T a = at(rl, d0 + j);
          const T xj = UNIT ? rdlane(x, j) : rdlane(x, j) * rdlane(rinv, j);
          x = (lane == j) ? xj : ((lane > j) ? x - a * xj : x);
        }
      }
      if (act) X[rl * ldx + c] = x;
    }
    __syncthreads();
    const int r0 = UPPER ? 0 : d0 + tb;
    const int r1 = UPPER ? d0 : n;
    for (int r = r0 + tid; r < r1; r += 256) {
      T acc[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] = T(0);
#pragma unroll 8
      for (int j = 0; j < tb; ++j) {
        const T a = at(r, d0 + j);
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] += a * X[(d0 + j) * ldx + (c < cm1 ? c : cm1)];
      }
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if (c < nc) X[r * ldx + c] -= acc[c];
    }
    __syncthreads();
  }
}

// Triangular solve M X = X in LDS for 8 < nc <= 16 CQ columns with X row
// stride ldx >= 16 CQ (columns nc..16CQ-1 of X are scratch): each 16-row
// diagonal block is solved by one thread per column in registers, then
// thread (ty, tx) of a 16 x 16 arrangement updates rows ty + 16i beyond the
// block on columns tx + 16q (q < CQ) holding the block's solved values of
// its CQ columns in registers, so every read of M feeds CQ FMAs.
template <typename T, bool UPPER, bool UNIT, int CQ, typename At>
__device__ void tri_solve_tiled(At at, int n, T* X, int ldx, int nc) {
  constexpr int DB = 16;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int nblk = (n + DB - 1) / DB;
  const int nm1 = n - 1;
  for (int bi = 0; bi < nblk; ++bi) {
    const int blk = UPPER ? nblk - 1 - bi : bi;
    const int d0 = blk * DB;
    const int tb = (n - d0) < DB ? (n - d0) : DB;
    for (int c = tid; c < nc; c += 256) {
      T x[DB];
#pragma unroll
      for (int j = 0; j < DB; ++j) {
        const T v = X[(d0 + (j < tb ? j : tb - 1)) * ldx + c];
        x[j] = (j < tb) ? v : T(0);
      }
      if (UPPER) {
#pragma unroll
        for (int j = DB - 1; j >= 0; --j) {
          const int jj = (d0 + j) < nm1 ? (d0 + j) : nm1;
          T xj = UNIT ? x[j] : x[j] / at(jj, jj);
          xj = (j < tb) ? xj : T(0);
          x[j] = xj;
#pragma unroll
          for (int r = 0; r < j; ++r) x[r] -= at((d0 + r) < nm1 ? (d0 + r) : nm1, jj) * xj;
        }
      } else {
#pragma unroll
        for (int j = 0; j < DB; ++j) {
          const int jj = (d0 + j) < nm1 ? (d0 + j) : nm1;
          const T xj = UNIT ? x[j] : x[j] / at(jj, jj);
          x[j] = xj;
#pragma unroll
          for (int r = j + 1; r < DB; ++r) x[r] -= at((d0 + r) < nm1 ? (d0 + r) : nm1, jj) * xj;
        }
      }
#pragma unroll
      for (int j = 0; j < DB; ++j)
        if (j < tb) X[(d0 + j) * ldx + c] = x[j];
    }
    __syncthreads();
    const int r0 = UPPER ? 0 : d0 + tb;
    const int r1 = UPPER ? d0 : n;
    if (r1 > r0) {
      T xc[CQ][DB];
#pragma unroll
      for (int q = 0; q < CQ; ++q)
#pragma unroll
        for (int j = 0; j < DB; ++j) {
          const T v = X[(d0 + (j < tb ? j : tb - 1)) * ldx + tx + 16 * q];
          xc[q][j] = (j < tb) ? v : T(0);
        }
      for (int r = r0 + ty; r < r1; r += 16) {
        T acc[CQ];
#pragma unroll
        for (int q = 0; q < CQ; ++q) acc[q] = T(0);
#pragma unroll
        for (int j = 0; j < DB; ++j) {
          const T a = at(r, (d0 + j) < nm1 ? (d0 + j) : nm1);
#pragma unroll
          for (int q = 0; q < CQ; ++q) acc[q] += a * xc[q][j];
        }
#pragma unroll
        for (int q = 0; q < CQ; ++q) X[r * ldx + tx + 16 * q] -= acc[q];
      }
    }
    __syncthreads();
  }
}

// X (n x nc, row stride ldx) <- M^{-1} X for a triangular M read through
// `at`.  For nc > 8 the columns go in chunks of up to 64 (FP64) or 128
// (FP32) with 16 CQ columns per tile row, so ldx must be at least nc
// rounded up to 16; the padding columns are scratch.
template <typename T, bool UPPER, bool UNIT, typename At>
__device__ inline void tri_solve(At at, int n, T* X, int ldx, int nc) {
  if (nc == 1) {
    tri_solve_wave<T, UPPER, UNIT, 1>(at, n, X, ldx, nc);
    return;
  }
  if (nc == 2) {
    tri_solve_wave<T, UPPER, UNIT, 2>(at, n, X, ldx, nc);
    return;
  }
  if (nc <= 4) {
    tri_solve_wave<T, UPPER, UNIT, 4>(at, n, X, ldx, nc);
    return;
  }
  if (nc <= 8) {
    tri_solve_wave<T, UPPER, UNIT, 8>(at, n, X, ldx, nc);
    return;
  }
  constexpr int CQM = sizeof(T) == 8 ? 4 : 8;
  for (int c0 = 0; c0 < nc; c0 += 16 * CQM) {
    const int w = (nc - c0) < 16 * CQM ? (nc - c0) : 16 * CQM;
    switch ((w + 15) >> 4) {
      case 1: tri_solve_tiled<T, UPPER, UNIT, 1>(at, n, X + c0, ldx, w); break;
      case 2: tri_solve_tiled<T, UPPER, UNIT, 2>(at, n, X + c0, ldx, w); break;
      case 3: tri_solve_tiled<T, UPPER, UNIT, 3>(at, n, X + c0, ldx, w); break;
      case 4: tri_solve_tiled<T, UPPER, UNIT, 4>(at, n, X + c0, ldx, w); break;
      default:
        if constexpr (CQM > 4) {
          switch ((w + 15) >> 4) {
            case 5: tri_solve_tiled<T, UPPER, UNIT, 5>(at, n, X + c0, ldx, w); break;
            case 6: tri_solve_tiled<T, UPPER, UNIT, 6>(at, n, X + c0, ldx, w); break;
            case 7: tri_solve_tiled<T, UPPER, UNIT, 7>(at, n, X + c0, ldx, w); break;
            default: tri_solve_tiled<T, UPPER, UNIT, 8>(at, n, X + c0, ldx, w); break;
          }
        }
        break;
    }
  }
}

// tri_solve for a compile-time column count NC (1, 2, 4, 8, 16 or 32): only
// the one variant is instantiated.
template <typename T, bool UPPER, bool UNIT, int NC, typename At>
__device__ inline void tri_solve_n(At at, int n, T* X, int ldx) {
  if constexpr (NC <= 8)
    tri_solve_wave<T, UPPER, UNIT, NC>(at, n, X, ldx, NC);
  else
    tri_solve_tiled<T, UPPER, UNIT, NC / 16>(at, n, X, ldx, NC);
}

// Tiled triangular solve for any nc: chunks of 16 CQM columns, the last one
// padded to 16 times a power of two columns, so ldx must be at least
// lu_factor_ldx<T>(nc) - 1 (columns past nc are scratch).
template <typename T, bool UPPER, bool UNIT, typename At>
__device__ inline void tri_solve_wide(At at, int n, T* X, int ldx, int nc) {
  constexpr int CQM = sizeof(T) == 8 ? 4 : 8;
  for (int c0 = 0; c0 < nc; c0 += 16 * CQM) {
    const int w = (nc - c0) < 16 * CQM ? (nc - c0) : 16 * CQM;
    const int cq = (w + 15) >> 4;
    if (cq == 1) tri_solve_tiled<T, UPPER, UNIT, 1>(at, n, X + c0, ldx, w);
    else if (cq == 2) tri_solve_tiled<T, UPPER, UNIT, 2>(at, n, X + c0, ldx, w);
    else if (cq <= 4) tri_solve_tiled<T, UPPER, UNIT, 4>(at, n, X + c0, ldx, w);
    else if constexpr (CQM > 4) tri_solve_tiled<T, UPPER, UNIT, CQM>(at, n, X + c0, ldx, w);
  }
}

// Slots of the rows one LU panel touches (its rows k0..k0+nb-1 and the
// distinct pivot rows below it; piv holds 0-based absolute pivots) and, per
// slot, the slot whose entry lands there after the interchanges.  Run by one
// wavefront (lanes = panel rows); writes s_row, s_src (2nb each), *s_cnt.
__device__ inline void panel_slots(const int* piv, int k0, int nb, int* s_slot, int* s_row, int* s_src,
                                   int* s_cnt) {
  const int i = threadIdx.x & 63;
  const int p = (i < nb) ? piv[k0 + i] : -1;
  bool first = (i < nb) && (p >= k0 + nb);
  int src = i;
  for (int q = 0; q < nb; ++q) {
    const bool hit = first & (q < i) & (piv[k0 + q] == p);
    src = (hit && src == i) ? q : src;
  }
  first = first & (src == i);
  const unsigned long long mask = __ballot(first);
  const int rank = __popcll(mask & ((1ull << i) - 1ull));
  if (first) {
    s_slot[i] = nb + rank;
    s_row[nb + rank] = p;
  }
  if (i < nb) s_row[i] = k0 + i;
  const int cnt = nb + __popcll(mask);
  if (i == 0) *s_cnt = cnt;
  __builtin_amdgcn_wave_barrier();
  if (i < nb) {
    if (p < k0 + nb) s_slot[i] = p - k0;
    else if (!first) s_slot[i] = s_slot[src];
  }
  __builtin_amdgcn_wave_barrier();
  if (i < cnt) {
    int pos = i;
    for (int q = 0; q < nb; ++q) {
      const int c = s_slot[q];
      pos = (pos == q) ? c : ((pos == c) ? q : pos);
    }
    s_src[pos] = i;
  }
  __builtin_amdgcn_wave_barrier();
}

// S[kk.., kk..] -= S[kk.., k0..k0+nb) S[k0..k0+nb), kk..] with kk = k0 + nb a
// multiple of 16 (n <= 16RT, S padded to 16RT x 16RT): wavefront w takes column tiles
// w, w + 4, ... from kk's tile on and every row tile (rows above kk multiply
// zeros), so each K step reads one fragment of the panel's row block and RT
// of its column block, then issues RT independent MFMAs.
template <typename T, int RT>
__device__ __attribute__((always_inline)) void lds_panel_update(T* S, int n, int ld, int k0, int nb) {
  typedef typename Mfma<T>::acc_t acc_t;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i16 = lane & 15, kq = lane >> 4;
  const int kk = k0 + nb, t0 = kk >> 4, nm1 = n - 1;
  for (int ct = t0 + wave; ct < RT; ct += 4) {
    const int col = ct * 16 + i16;
    const int colc = col < nm1 ? col : nm1;
    acc_t acc[RT];
#pragma unroll
    for (int j = 0; j < RT; ++j) acc[j] = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int k = 4 * ks + kq;
      const bool kin = k < nb;
      const int kc = k0 + (kin ? k : nb - 1);
      const T bv = S[kc * ld + colc];
      T av[RT];
#pragma unroll
      for (int j = 0; j < RT; ++j) {
        const int r = j * 16 + i16;
        av[j] = S[(r < nm1 ? r : nm1) * ld + kc];
      }
      const T b = kin ? bv : T(0);
#pragma unroll
      for (int j = 0; j < RT; ++j) acc[j] = Mfma<T>::mma((kin & (j >= t0)) ? av[j] : T(0), b, acc[j]);
    }
    // rows above kk subtract zeros; rows and columns past n land in the
    // padding of S (never read)
#pragma unroll
    for (int j = 0; j < RT; ++j) {
#pragma unroll
      for (int q = 0; q < 4; ++q) S[(j * 16 + Mfma<T>::row(lane, q)) * ld + col] -= acc[j][q];
    }
  }
}

// Blocked in-place LU with partial pivoting of the n x n matrix S in LDS
// (row stride ld >= 16RT + 1, padded to 16RT rows and columns, n <= 16RT).
// Each 16-column panel is factored by wavefront 0 alone with the panel in
// registers (lane l holds panel rows l + 64s): rows never move, each keeps
// its logical position, the DPP argmax ranks (|a|, logical position), the
// pivot row reaches every lane by readlanes, then division by the pivot and
// the rank-1 update in registers; no LDS traffic or barrier per column, and
// the rows go back to LDS in logical order.  Then one thread per column
// outside the panel moves the entries the panel's interchanges touch (slots
// from panel_slots) and, right of the panel, solves with the unit-lower L11;
// then MFMA tiles update the trailing matrix.  Three barriers per panel.
// piv[k] is 0-based.  wsp: LDS workspace of 6 * 16 ints.
template <typename T, int RT>
__device__ __attribute__((always_inline)) void lu_lds_blocked(T* S, int n, int ld, int* piv, int* wsp, int* info, int64_t b) {
  constexpr int NB = 16;
  constexpr int RPW = (16 * RT + 63) / 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int* s_slot = wsp;                 // NB
  int* s_row = wsp + NB;             // 2 NB
  int* s_src = wsp + 3 * NB;         // 2 NB
  int* s_cnt = wsp + 5 * NB;         // 1
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int nb = (n - k0) < NB ? (n - k0) : NB;
    const int R = n - k0;
    if (wave == 0) {
      T x[RPW][NB];
      int lp[RPW];  // logical position of each register row (rows stay put)
#pragma unroll
      for (int s = 0; s < RPW; ++s) {
        const int rr = lane + 64 * s;
        lp[s] = rr;
        const int rc = rr < R - 1 ? rr : R - 1;
#pragma unroll
        for (int c = 0; c < NB; ++c) {
          const T v = S[(k0 + rc) * ld + k0 + c];
          x[s][c] = ((rr < R) & (c < nb)) ? v : T(0);
        }
      }
      int zc = 0;    // panel column (1-based) of the first exactly zero pivot
      int pivl = 0;  // lane j: the pivot row of column j
      // one column: a straight line of VALU/SALU work (no branch, no LDS)
      // when the panel is full, so consecutive columns can overlap
      auto column = [&](auto jc, auto fullc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        constexpr bool full = decltype(fullc)::value;
        if (full || j < nb) {
          // candidates: rows at logical positions j..R-1; the winner is the
          // largest |a|, ties to the first in order; a NaN ranks below every
          // number and above the non-candidates
          int p;
          if constexpr (sizeof(T) == 4) {
            unsigned long long best = 0ull;
#pragma unroll
            for (int s = 0; s < RPW; ++s) {
              const bool cand = (lp[s] >= j) & (lp[s] < R);
              const float a = absv(x[s][j]);
              const unsigned hi = (a == a) ? __float_as_uint(a) + 1u : 0u;
              const unsigned long long k = ((unsigned long long)hi << 32) | (0xFFFFFFFFu - (unsigned)lp[s]);
              best = umax64(best, cand ? k : 0ull);
            }
            best = wave_max_u64(best);
            p = (int)(0xFFFFFFFFu - (unsigned)best);
          } else {
            T best = T(-1);
            int bi = INT_MAX;
#pragma unroll
            for (int s = 0; s < RPW; ++s) {
              const bool cand = (lp[s] >= j) & (lp[s] < R);
              const T a = absv(x[s][j]);
              argmax_merge(best, bi, cand ? ((a == a) ? a : T(-0.5)) : T(-1), lp[s]);
            }
            wave_argmax(best, bi);
            p = bi;
          }
          // the register row holding logical position p (exactly one)
          int pl = 0, ps = 0;
          static_for<0, RPW>([&](auto sc) __attribute__((always_inline)) {
            constexpr int sl = decltype(sc)::value;
            const unsigned long long hit = __ballot(lp[sl] == p);
            ps = hit != 0ull ? sl : ps;
            pl = hit != 0ull ? (int)__builtin_ctzll(hit) : pl;
          });
          pivl = (lane == j) ? k0 + p : pivl;
          // the pivot row's columns j.. from its lane
          T pr[NB];
#pragma unroll
          for (int c = j; c < NB; ++c) {
            T v = x[0][c];
#pragma unroll
            for (int s = 1; s < RPW; ++s) v = bit_blend(v, x[s][c], ps == s);
            pr[c] = rdlane(v, pl);
          }
          // interchange j <-> p on the logical positions only
#pragma unroll
          for (int s = 0; s < RPW; ++s) lp[s] = (lp[s] == p) ? j : ((lp[s] == j) ? p : lp[s]);
          const T pv = pr[j];
          const bool nz = pv != T(0);
          zc = (!nz & (zc == 0)) ? j + 1 : zc;
          const T dv = nz ? pv : T(1);
#pragma unroll
          for (int s = 0; s < RPW; ++s) {
            const bool m = (lp[s] > j) & (lp[s] < R);
            const T l = x[s][j] / dv;
            const T lm = m ? l : T(0);
            x[s][j] = m ? l : x[s][j];
#pragma unroll
            for (int c = j + 1; c < NB; ++c) x[s][c] -= lm * pr[c];
          }
        }
      };
      if (nb == NB)
        static_for<0, NB>([&](auto jc) __attribute__((always_inline)) { column(jc, std::true_type{}); });
      else
        static_for<0, NB>([&](auto jc) __attribute__((always_inline)) { column(jc, std::false_type{}); });
      if (lane < nb) piv[k0 + lane] = pivl;
      if ((zc != 0) & (lane == 0) & (info != nullptr))
        if (info[b] == 0) info[b] = k0 + zc;
      // every row to its logical position
#pragma unroll
      for (int s = 0; s < RPW; ++s) {
        if (lp[s] < R) {
#pragma unroll
          for (int c = 0; c < NB; ++c)
            if (c < nb) S[(k0 + lp[s]) * ld + k0 + c] = x[s][c];
        }
      }
      __builtin_amdgcn_wave_barrier();
      panel_slots(piv, k0, nb, s_slot, s_row, s_src, s_cnt);
    }
    __syncthreads();
    // interchanges outside the panel and U12 = L11^{-1} A12
    {
      const int c = tid;
      const int cnt = *s_cnt;
      const int cc = c < n - 1 ? c : n - 1;
      const bool right = c >= k0 + nb;
      T u[NB], d[NB];
#pragma unroll
      for (int q = 0; q < NB; ++q) u[q] = S[s_row[s_src[q < nb - 1 ? q : nb - 1]] * ld + cc];
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        const int q = nb + t < cnt - 1 ? nb + t : cnt - 1;
        d[t] = S[s_row[s_src[q]] * ld + cc];
      }
      if (right) {
#pragma unroll
        for (int a = 0; a < NB; ++a) {
          const T ua = (a < nb) ? u[a] : T(0);
#pragma unroll
          for (int r = a + 1; r < NB; ++r) {
            const T l = S[(k0 + (r < nb ? r : nb - 1)) * ld + k0 + a];
            u[r] -= ((r < nb) ? l : T(0)) * ua;
          }
        }
      }
      // destination rows read up front (one LDS round trip), -1 = no store
      int ru[NB], rd[NB];
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const int r = s_row[q];
        ru[q] = (q < nb) ? r : -1;
      }
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        const int r = s_row[nb + t];
        rd[t] = (nb + t < cnt) ? r : -1;
      }
      if (c < n && (c < k0 || right)) {
#pragma unroll
        for (int q = 0; q < NB; ++q)
          if (ru[q] >= 0) S[ru[q] * ld + c] = u[q];
#pragma unroll
        for (int t = 0; t < NB; ++t)
          if (rd[t] >= 0) S[rd[t] * ld + c] = d[t];
      }
    }
    __syncthreads();
    // trailing update: rows and columns k0+nb .. n-1
    if (k0 + nb < n) lds_panel_update<T, RT>(S, n, ld, k0, nb);
    __syncthreads();
  }
}

// Stage the bs x bs row-major block src into LDS Mt (row stride ld), with
// eight loads in flight per thread.
template <typename T>
__device__ inline void stage_block(T* Mt, const T* src, int bs, float rbs, int ld) {
  const int nn = bs * bs;
  for (int e0 = threadIdx.x; e0 < nn; e0 += 256 * 8) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + 256 * u;
      v[u] = src[e < nn ? e : nn - 1];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + 256 * u;
      const int r = idiv(e, rbs);
      if (e < nn) Mt[r * ld + e - r * bs] = v[u];
    }
  }
  __syncthreads();
}

// Xc[r][c] -= sum_k Mt[r][k] Y[k][c] (all operands in LDS, Mt row stride
// bs+1, X row stride ldx); thread (tr, tc) owns column tc and rows
// tr + (256/nr) x, read from clamped rows so the k loop has no tests.
// SET: Xc = Mt Y instead of Xc -= Mt Y, with every read of Y done before
// the first write (Xc may be Y).
template <typename T, int RPT, bool SET = false>
__device__ inline void sub_mat_vec(const T* Mt, const T* Y, T* Xc, int bs, int nr, int ldx) {
  const int tid = threadIdx.x;
  const int tc = tid % nr, tr = tid / nr, RG = 256 / nr;
  const int ld = bs + 1;
  int off[RPT];
  T acc[RPT];
#pragma unroll
  for (int x = 0; x < RPT; ++x) {
    const int r = tr + RG * x;
    off[x] = (r < bs ? r : bs - 1) * ld;
    acc[x] = T(0);
  }
  const T* yp = Y + tc;
#pragma unroll 4
  for (int k = 0; k < bs; ++k) {
    const T y = yp[k * ldx];
#pragma unroll
    for (int x = 0; x < RPT; ++x) acc[x] += Mt[off[x] + k] * y;
  }
  if constexpr (SET) __syncthreads();
#pragma unroll
  for (int x = 0; x < RPT; ++x) {
    const int r = tr + RG * x;
    if (r < bs) {
      if constexpr (SET) Xc[r * ldx + tc] = acc[x];
      else Xc[r * ldx + tc] -= acc[x];
    }
  }
  __syncthreads();
}

// Xc -= Mt Y (SET: Xc = Mt Y, Xc may be Y) for 32 columns on the matrix
// cores (16 x 16 x 4 steps):
// wavefront w owns column tile w & 1 and row tiles (w >> 1) + 2 j, so the
// Y fragment of a K step is read once for all of its row tiles.  Rows of Mt
// past bs are read clamped and feed only rows that are never written; K
// past bs is zeroed.  bs <= 160.
template <typename T, bool SET = false>
__device__ inline void sub_mat_mfma32(const T* Mt, const T* Y, T* Xc, int bs, int ldx) {
  typedef typename Mfma<T>::acc_t acc_t;
  constexpr int MAXT = 5;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ld = bs + 1, bm1 = bs - 1;
  const int ct = wave & 1, rt0 = wave >> 1, nrt = (bs + 15) >> 4;
  const int cr = lane & 15, kq = lane >> 4;
  const int ycol = ct * 16 + cr;
  acc_t acc[MAXT];
  int aoff[MAXT];
#pragma unroll
  for (int j = 0; j < MAXT; ++j) {
    acc[j] = acc_t{0, 0, 0, 0};
    const int r = (rt0 + 2 * j) * 16 + cr;
    aoff[j] = (r < bm1 ? r : bm1) * ld;
  }
  for (int k0 = 0; k0 < bs; k0 += 4) {
    const int k = k0 + kq;
    const bool kin = k < bs;
    const int kc = kin ? k : bm1;
    const T yv = Y[kc * ldx + ycol];
    const T b = kin ? yv : T(0);
    // every tile's load is issued before the products (one LDS round trip
    // per k step); tiles past the block multiply zeros and are not stored
#pragma unroll
    for (int j = 0; j < MAXT; ++j) {
      const T av = Mt[aoff[j] + kc];
      acc[j] = Mfma<T>::mma((kin & (rt0 + 2 * j < nrt)) ? av : T(0), b, acc[j]);
    }
  }
  if constexpr (SET) __syncthreads();
#pragma unroll
  for (int j = 0; j < MAXT; ++j) {
    const int rt = rt0 + 2 * j;
    if (rt < nrt) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = rt * 16 + Mfma<T>::row(lane, q);
        if (r < bs) {
          if constexpr (SET) Xc[r * ldx + ycol] = acc[j][q];
          else Xc[r * ldx + ycol] -= acc[j][q];
        }
      }
    }
  }
  __syncthreads();
}

// Single-column Xc -= Mt y: for bs <= 128 two threads per row split k (the
// upper half's partial sums pass through W, bs elements of LDS).
template <typename T, bool SET = false>
__device__ inline void sub_mat_vec1(const T* Mt, const T* Y, T* Xc, int bs, T* W) {
  const int tid = threadIdx.x, ld = bs + 1;
  if (bs <= 128) {
    const int r = tid & 127, h = tid >> 7;
    const int rc = (r < bs ? r : bs - 1) * ld;
    const int kh = (bs + 1) >> 1;
    const int k0 = h * kh, k1 = (k0 + kh) < bs ? (k0 + kh) : bs;
    T acc = T(0);
#pragma unroll 8
    for (int k = k0; k < k1; ++k) acc += Mt[rc + k] * Y[k];
    if (h == 1 && r < bs) W[r] = acc;
    __syncthreads();
    if (h == 0 && r < bs) {
      if constexpr (SET) Xc[r] = acc + W[r];
      else Xc[r] -= acc + W[r];
    }
  } else {
    // bs <= 256: one row per thread
    const int r = tid;
    T acc = T(0);
    if (r < bs) {
#pragma unroll 8
      for (int k = 0; k < bs; ++k) acc += Mt[r * ld + k] * Y[k];
    }
    if constexpr (SET) __syncthreads();
    if (r < bs) {
      if constexpr (SET) Xc[r] = acc;
      else Xc[r] -= acc;
    }
  }
  __syncthreads();
}

constexpr size_t kLdsMax = 160 * 1024;

// 16 x 16 thread tiles per dimension of a bs x bs block (the RT of the LU
// and Gauss-Jordan kernels, which pad blocks to 16 RT).
__host__ __device__ inline int tile_rt(int bs) { return (bs + 15) / 16; }

// row stride of the factor kernel's right-hand-side chunk: tri_solve_wide
// runs full chunks of 16 CQM columns and pads the last one to 16 times a
// power of two columns; plus one for odd strides
template <typename T>
__host__ __device__ inline int lu_factor_ldx(int xc) {
  constexpr int CQM = sizeof(T) == 8 ? 4 : 8;
  const int full = (xc / (16 * CQM)) * (16 * CQM);
  const int q = (xc - full + 15) / 16;
  int p2 = q > 0 ? 1 : 0;
  while (p2 < q) p2 <<= 1;
  return full + 16 * p2 + 1;
}

template <typename T>
size_t lu_factor_lds(int bs, int xc) {
  const size_t bp = 16 * (size_t)tile_rt(bs);
  return sizeof(T) * (bp * (bp + 1) + (size_t)bs * lu_factor_ldx<T>(xc) + bp * (KC + 1) + (size_t)KC * (bp + 1)) +
         2 * sizeof(int) * bs;
}

// widest right-hand-side chunk of the factor kernel that fits in LDS
template <typename T>
int lu_factor_xc(int bs) {
  int xc = bs;
  while (xc > 8 && lu_factor_lds<T>(bs, xc) > kLdsMax) xc -= 8;
  return xc;
}

template <typename T>
size_t lu_solve_lds(int bs, int nr) {
  const size_t ldx = nr == 1 ? 1 : nr + 1;
  return sizeof(T) * ((size_t)bs * (bs + 1) + 2 * (size_t)bs * ldx + (size_t)bs) + 2 * sizeof(int) * bs;
}

// LDS of the Gauss-Jordan factor kernel instantiated with RT tiles
template <typename T>
size_t factor_lds(int bs, int rt) {
  const size_t bp = 16 * (size_t)rt;
  return sizeof(T) * ((size_t)bs * (bs + 1) + bp * (KC + 1) + (size_t)KC * (bp + 1)) + sizeof(int) * bs;
}

template <typename T>
size_t solve_lds(int bs, int nr) {
  return sizeof(T) * (2 * (size_t)bs * nr + (size_t)bs * (KC + 1));
}

}  // namespace

}  // namespace slab
