// Forward and backward block sweeps with the factors of the block LU (the
// lusolve_sweep of src_solver/block_tridiag_solver_dense.py:179-203).
#include "lds_dev.h"

#include <atomic>

namespace slab {

// Right-hand-side tile widths chosen by the autotuner, per (dtype, block
// size, inverse factors): 0 = the built-in rule.  Process-wide; read at every
// sweep launch.
namespace {
constexpr int kTileBsMax = 256;
std::atomic<int> g_sweep_tile[2][kTileBsMax + 1][2];
}  // namespace

int sweep_tile_override(int dtype, int bs, bool inv) {
  if (dtype < 0 || dtype > 1 || bs < 1 || bs > kTileBsMax) return 0;
  return g_sweep_tile[dtype][bs][inv ? 1 : 0].load(std::memory_order_relaxed);
}

int set_sweep_tile(int dtype, int bs, int inv, int nr) {
  if (dtype < 0 || dtype > 1 || bs < 1 || bs > kTileBsMax) return -1;
  if (nr != 0 && nr != 1 && nr != 2 && nr != 4 && nr != 8 && nr != 16 && nr != 32) return -1;
  return g_sweep_tile[dtype][bs][inv ? 1 : 0].exchange(nr);
}

namespace {

// One workgroup per (chain, NR-column tile).  A full sweep is a fixed
// sequence of 4m - 2 stages, each with one block resident in LDS:
//   forward  DIAGL(0), SUB(0), DIAGL(1), ..., SUB(m-2), DIAGL(m-1)
//   backward DIAGU(m-1), SUP(m-2), DIAGU(m-2), ..., SUP(0), DIAGU(0)
// DIAGL(i): permute (composed pivots), unit-lower solve, store y_i;
// SUB(i): x_{i+1} = f_{i+1} - Lfac_i y_i; DIAGU(i): upper solve, store x_i;
// SUP(i): y_i - Ufac_i x_{i+1}.  mode 1 runs the forward half only; mode 2
// runs the backward half of a chain whose block m (past its end) already
// holds its solution: SUP(m-1), DIAGU(m-1), ..., SUP(0), DIAGU(0).  While a
// stage computes, the next stage's block (PF elements per thread; any rest
// is loaded at hand-over), its right-hand-side rows and its pivots are
// already being loaded into registers, so global latency is off the
// critical path.  Workgroups past c0's chains take c1's chains.
//
// TRANS solves the transposed system (lusolve_sweep_transpose,
// block_tridiag_solver_dense.py:205-222).  With the factorization
// T = LL UU (LL: diagonal P_i^T L_i, sub-diagonal Lfac_i; UU: diagonal U_i,
// super-diagonal Ufac_i), T^T = UU^T LL^T, so the same stage sequence runs
// with every block staged transposed: DIAGL(i) solves with U_i^T (lower,
// non-unit), SUB(i) subtracts Ufac_i^T, DIAGU(i) solves with L_i^T (upper,
// unit) and then applies P_i^T to its result, SUP(i) subtracts Lfac_i^T.
// With inverse factors DIAGL multiplies by (U^{-1})^T and DIAGU by
// (L^{-1} P)^T.
template <typename T, int NR, int PF, bool INV, bool TRANS>
__global__ __launch_bounds__(256, (INV && NR == 32) ? 2 : 1) void k_btri_lu_solve(int bs, ChainSet<T> c0, ChainSet<T> c1, int64_t ldf,
                                                       int nrhs, int mode) {
  // rows per thread in the block products, for the largest block the fused
  // factorization accepts (112 rows in FP64, 160 in FP32)
  constexpr int BSMAX = sizeof(T) == 8 ? 112 : 160;
  constexpr int RPT = (BSMAX * NR + 255) / 256;
  extern __shared__ unsigned char smem_raw[];
  constexpr int nr = NR;
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
  const int col0 = blockIdx.y * NR;
  const int ncols = (nrhs - col0) < NR ? (nrhs - col0) : NR;
  const int cm1 = ncols - 1;
  const int nn = bs * bs, nf = bs * nr;
  const float rbs = 1.0f / (float)bs;
  auto Mat = [&](int r, int k) -> T { return Mt[r * ld + k]; };
  const int nt = c0.nchains + c1.nchains;
  for (int g = blockIdx.x; g < nt; g += gridDim.x) {
    const bool sec = g >= c0.nchains;
    const ChainSet<T> a = sec ? c1 : c0;
    const int64_t b = sec ? g - c0.nchains : g;
    const int m = a.m;
    const int nst = mode == 0 ? 4 * m - 2 : (mode == 1 ? 2 * m - 1 : 2 * m);
    // stage s -> kind (0 DIAGL, 1 SUB, 2 DIAGU, 3 SUP) and block index i
    auto decode = [&](int s, int& i) -> int {
      if (mode == 2) {
        if (s & 1) { i = m - 1 - ((s - 1) >> 1); return 2; }
        i = m - 1 - (s >> 1);
        return 3;
      }
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
    // inverse factors: DIAGL multiplies by L^{-1} P (linv), DIAGU by U^{-1}
    // (diag), both plain block products (transposed: DIAGL by the
    // transposed U^{-1}, DIAGU by the transposed L^{-1} P)
    // (compile-time, so each variant keeps only the registers its stages need)
    constexpr bool inv = INV;
    // the last block's DIAGU of a full sweep reuses the resident DIAGL block
    auto resident = [&](int k, int i) -> bool { return !inv && mode == 0 && k == 2 && i == m - 1; };
    // the stage whose block needs the pivots: the permutation enters before
    // the unit-lower solve, or (transposed) after the unit-upper one
    constexpr int kperm = TRANS ? 2 : 0;
    T* fb = a.f + b * a.fbs;
    const int* pb = a.ipiv + b * a.pbs;
    T* Xc = B0;
    T* Y = B1;
    auto frow = [&](int i, int r) -> T* { return fb + (int64_t)i * a.fis + (int64_t)r * ldf + col0; };
    auto block_src = [&](int k, int i) -> const T* {
      // transposed sweeps use Ufac in the forward and Lfac in the backward half
      if (k == (TRANS ? 3 : 1)) return a.sub + b * a.lbs + (int64_t)i * a.lis;
      if (k == (TRANS ? 1 : 3)) return a.sup + b * a.ubs + (int64_t)i * a.uis;
      if (k == (TRANS ? 2 : 0) && inv) return a.linv + b * a.dbs + (int64_t)i * a.dis;
      return a.diag + b * a.dbs + (int64_t)i * a.dis;
    };
    // LDS slot of element e = r * bs + c of a block: row-major, or
    // transposed for the transposed sweeps
    auto slot = [&](int e) -> int {
      const int r = idiv(e, rbs), c = e - r * bs;
      return TRANS ? c * ld + r : r * ld + c;
    };
    // forward sweeps start at the first block whose right-hand side tile
    // has a nonzero: the blocks above it solve to zero and already hold
    // zeros (a block-sparse right-hand side, as in the slab Schur
    // complements, skips most of its forward sweep)
    int i0 = 0;
    if (mode != 2) {
      for (; i0 < m; ++i0) {
        bool nz = false;
        for (int e = tid; e < nf; e += 256) {
          const int r = e >> lgnr, c = e & (nr - 1);
          if (c < ncols) nz |= *(frow(i0, r) + c) != T(0);
        }
        if (__syncthreads_or(nz)) break;
      }
      if (i0 == m) continue;  // zero right-hand side: zero solution, already in place
    }
    auto f_row = [&](int k, int i) -> int {  // block row of f loaded into Xc
      if (k == 0) return i == i0 ? i : -1;
      if (k == 1) return i + 1;
      if (k == 3) return i;
      return -1;
    };
    T pf[PF], pfx[RPT];
    int pfp[2];
    auto prefetch = [&](int s) {
      int i;
      const int k = decode(s, i);
      if (!resident(k, i)) {
        const T* src = block_src(k, i);
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const int e = tid + 256 * u;
          pf[u] = src[e < nn ? e : nn - 1];
        }
      }
      const int fi = f_row(k, i);
      if (fi >= 0) {
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
          const int e = tid + 256 * u, r = e >> lgnr, c = e & (nr - 1);
          // masked when committed, so the loads stay in flight meanwhile
          pfx[u] = *(frow(fi, r < bs ? r : bs - 1) + (c < cm1 ? c : cm1));
        }
      }
      if (k == kperm && !inv) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int r = tid + 256 * u;
          const int v = pb[(int64_t)i * a.pis + (r < bs ? r : bs - 1)];
          pfp[u] = v - 1;
        }
      }
    };
    auto commit = [&](int s) {
      int i;
      const int k = decode(s, i);
      if (!resident(k, i)) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const int e = tid + 256 * u;
          if (e < nn) Mt[slot(e)] = pf[u];
        }
        const T* src = block_src(k, i);
        // the rest of the block, eight loads in flight per thread: for the
        // FP32 inverse-factor sweeps 16-byte vectors when rows hold whole
        // vectors (a vector never crosses a row), else single elements
        constexpr int VN = 16 / sizeof(T);
        typedef T vt __attribute__((ext_vector_type(VN)));
        constexpr bool kVecLoad = INV && sizeof(T) == 4;
        if (kVecLoad && (bs % VN) == 0 && (reinterpret_cast<uintptr_t>(src) & 15) == 0) {
          const vt* s4 = reinterpret_cast<const vt*>(src);
          const int n4 = nn / VN;
          for (int q0 = tid + (256 * PF) / VN; q0 < n4; q0 += 256 * 8) {
            vt v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const int q = q0 + 256 * u;
              v[u] = s4[q < n4 ? q : n4 - 1];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const int q = q0 + 256 * u;
              if (q < n4) {
                const int o = slot(q * VN);
#pragma unroll
                for (int t = 0; t < VN; ++t) Mt[o + (TRANS ? t * ld : t)] = v[u][t];
              }
            }
          }
        } else {
          for (int e0 = tid + 256 * PF; e0 < nn; e0 += 256 * 8) {
            T v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const int e = e0 + 256 * u;
              v[u] = src[e < nn ? e : nn - 1];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const int e = e0 + 256 * u;
              if (e < nn) Mt[slot(e)] = v[u];
            }
          }
        }
      }
      if (f_row(k, i) >= 0) {
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
          const int e = tid + 256 * u, r = e >> lgnr, c = e & (nr - 1);
          if (e < nf) Xc[r * ldx + c] = (c < ncols) ? pfx[u] : T(0);
        }
      }
      if (k == kperm && !inv) {
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
        if (c < ncols) *(frow(i, r) + c) = Z[r * ldx + c];
      }
    };
    if (mode == 2) {
      // the block past the chain end holds its solution: x_m
      for (int e = tid; e < nf; e += 256) {
        const int r = e >> lgnr, c = e & (nr - 1);
        Y[r * ldx + c] = (c < ncols) ? *(frow(m, r) + c) : T(0);
      }
    }
    const int s0 = mode == 2 ? 0 : 2 * i0;
    prefetch(s0);
    commit(s0);
    __syncthreads();
    for (int s = s0; s < nst; ++s) {
      if (s + 1 < nst) prefetch(s + 1);
      int i;
      const int k = decode(s, i);
      // Out = Mt In (Out may be In) for the inverse-factor stages
      auto block_set = [&](const T* In, T* Out) {
        if constexpr (NR == 1)
          sub_mat_vec1<T, true>(Mt, In, Out, bs, W);
        else if constexpr (NR == 32)
          sub_mat_mfma32<T, true>(Mt, In, Out, bs, ldx);
        else
          sub_mat_vec<T, RPT, true>(Mt, In, Out, bs, nr, ldx);
      };
      if (k == 0) {
        if (inv) {
          block_set(Xc, Y);
        } else if constexpr (TRANS) {
          // U_i^T y = x: the lower triangle (with the diagonal) of the
          // transposed LU block
          for (int e = tid; e < nf; e += 256) {
            const int r = e >> lgnr, c = e & (nr - 1);
            Y[r * ldx + c] = Xc[r * ldx + c];
          }
          __syncthreads();
          tri_solve_n<T, false, false, NR>(Mat, bs, Y, ldx);
        } else {
          piv_to_perm(P, bs, Q);
          for (int e = tid; e < nf; e += 256) {
            const int r = e >> lgnr, c = e & (nr - 1);
            Y[r * ldx + c] = Xc[Q[r] * ldx + c];
          }
          __syncthreads();
          tri_solve_n<T, false, true, NR>(Mat, bs, Y, ldx);
        }
        store_f(i, Y);
      } else if (k == 1 || k == 3) {
        if constexpr (NR == 1)
          sub_mat_vec1<T>(Mt, Y, Xc, bs, W);
        else if constexpr (NR == 32)
          sub_mat_mfma32<T>(Mt, Y, Xc, bs, ldx);
        else
          sub_mat_vec<T, RPT>(Mt, Y, Xc, bs, nr, ldx);
      } else {
        const bool own = mode == 0 && i == m - 1;
        T* Z = own ? Y : Xc;
        if (!inv && TRANS) {
          // L_i^T w = z (the strict upper triangle of the transposed LU
          // block, unit diagonal), then x = P_i^T w: row r of w lands in
          // row Q[r] of x, in the buffer that is not Z
          tri_solve_n<T, true, true, NR>(Mat, bs, Z, ldx);
          T* O = own ? Xc : Y;
          piv_to_perm(P, bs, Q);
          for (int e = tid; e < nf; e += 256) {
            const int r = e >> lgnr, c = e & (nr - 1);
            O[Q[r] * ldx + c] = Z[r * ldx + c];
          }
          __syncthreads();
          store_f(i, O);
          // the result continues in Y
          if (own) {
            T* t = Xc;
            Xc = Y;
            Y = t;
          }
        } else {
          if (inv) block_set(Z, Z);
          else tri_solve_n<T, true, false, NR>(Mat, bs, Z, ldx);
          store_f(i, Z);
          if (!own) {
            T* t = Xc;
            Xc = Y;
            Y = t;
          }
        }
      }
      __syncthreads();
      if (s + 1 < nst) commit(s + 1);
      __syncthreads();
    }
  }
}

template <typename T, int NR, int PF, bool INV, bool TRANS>
hipError_t launch_lu_solve_k(int bs, const ChainSet<T>& c0, const ChainSet<T>& c1, int64_t ldf, int nrhs, int mode,
                             hipStream_t s) {
  const size_t lds = lu_solve_lds<T>(bs, NR);
  const void* fn = reinterpret_cast<const void*>(&k_btri_lu_solve<T, NR, PF, INV, TRANS>);
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  const int nt = c0.nchains + c1.nchains;
  if (nt <= 0) return hipSuccess;
  dim3 grid(nt < 65535 ? nt : 65535, (nrhs + NR - 1) / NR);
  hipLaunchKernelGGL((k_btri_lu_solve<T, NR, PF, INV, TRANS>), grid, dim3(256), lds, s, bs, c0, c1, ldf, nrhs, mode);
  return hipGetLastError();
}

template <typename T, int NR, int PF>
hipError_t launch_lu_solve(int bs, const ChainSet<T>& c0, const ChainSet<T>& c1, int64_t ldf, int nrhs, int mode,
                           bool trans, hipStream_t s) {
  const bool inv = c0.linv != nullptr;
  if (inv) {
    // 32-column inverse-factor sweeps always run with the small prefetch
    // (launch_lu_solve_pf); the large one would spill, so it is not built
    if constexpr (NR == 32 && PF != 8) return hipErrorInvalidValue;
    else
      return trans ? launch_lu_solve_k<T, NR, PF, true, true>(bs, c0, c1, ldf, nrhs, mode, s)
                   : launch_lu_solve_k<T, NR, PF, true, false>(bs, c0, c1, ldf, nrhs, mode, s);
  }
  return trans ? launch_lu_solve_k<T, NR, PF, false, true>(bs, c0, c1, ldf, nrhs, mode, s)
               : launch_lu_solve_k<T, NR, PF, false, false>(bs, c0, c1, ldf, nrhs, mode, s);
}

// PF: block elements prefetched per thread (the rest of a block is loaded
// at hand-over); FP64 with 16 or 32 columns keeps it small to stay within
// the register file
template <typename T, int NR>
hipError_t launch_lu_solve_pf(int bs, const ChainSet<T>& c0, const ChainSet<T>& c1, int64_t ldf, int nrhs, int mode,
                              bool trans, hipStream_t s) {
  constexpr int PFBIG = sizeof(T) == 8 ? (NR >= 32 ? 8 : (NR >= 16 ? 16 : 40)) : 40;
  // inverse factors with 32 columns: a small prefetch keeps two workgroups
  // per CU (their LDS fits twice), which hides more than a whole-block
  // prefetch in one
  if ((int64_t)bs * bs <= 256 * 8 || (NR == 32 && c0.linv != nullptr))
    return launch_lu_solve<T, NR, 8>(bs, c0, c1, ldf, nrhs, mode, trans, s);
  return launch_lu_solve<T, NR, PFBIG>(bs, c0, c1, ldf, nrhs, mode, trans, s);
}

// Single-column sweeps with inverse factors (FP32, bs <= 104, bs % 4 == 0):
// every stage is out = M in (DIAGL: L^{-1} P, DIAGU: U^{-1}) or
// out = f - M in (SUB: Lfac, SUP: Ufac), so no block passes through LDS.
// Thread (r, h) = (tid / 2, tid % 2) holds half of row r of the stage's
// block in registers as 16-byte vectors, loaded one stage ahead; the
// vectors in and out are 3 rotating LDS buffers, one barrier per stage.
// Same stage order, modes and chain sets as k_btri_lu_solve.
constexpr int kInv1MaxBs = 104;

__global__ __launch_bounds__(256) void k_btri_inv1(int bs, ChainSet<float> c0, ChainSet<float> c1, int64_t ldf,
                                                   int mode) {
  constexpr int NVMAX = kInv1MaxBs / 2 / 4;  // float4 per half row
  __shared__ __attribute__((aligned(16))) float vb[3][kInv1MaxBs];
  const int tid = threadIdx.x, r = tid >> 1, h = tid & 1;
  const int hw = ((bs + 7) / 8) * 4;  // first half: a multiple of 4 columns
  const int cb = h ? hw : 0, ce = h ? bs : hw;
  const int nv = (ce - cb) >> 2;      // bs % 4 == 0
  const int rc = r < bs ? r : bs - 1;
  const int nt = c0.nchains + c1.nchains;
  for (int g = blockIdx.x; g < nt; g += gridDim.x) {
    const bool sec = g >= c0.nchains;
    const ChainSet<float> a = sec ? c1 : c0;
    const int64_t b = sec ? g - c0.nchains : g;
    const int m = a.m;
    const int nst = mode == 0 ? 4 * m - 2 : (mode == 1 ? 2 * m - 1 : 2 * m);
    auto decode = [&](int s, int& i) __attribute__((always_inline)) -> int {
      if (mode == 2) {
        if (s & 1) { i = m - 1 - ((s - 1) >> 1); return 2; }
        i = m - 1 - (s >> 1);
        return 3;
      }
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
    float* fb = a.f + b * a.fbs;
    auto fptr = [&](int i) -> float* { return fb + (int64_t)i * a.fis; };
    auto block = [&](int k, int i) -> const float* {
      if (k == 1) return a.sub + b * a.lbs + (int64_t)i * a.lis;
      if (k == 3) return a.sup + b * a.ubs + (int64_t)i * a.uis;
      if (k == 0) return a.linv + b * a.dbs + (int64_t)i * a.dis;
      return a.diag + b * a.dbs + (int64_t)i * a.dis;
    };
    // three register buffers: the block of stage s + 2 is requested while
    // stage s computes (the chain is latency-bound: 20 workgroups)
    float4 q0[NVMAX], q1[NVMAX], q2[NVMAX];
    float f0 = 0.f, f1 = 0.f, f2 = 0.f;
    // branch-free and never consumed at load time: past the last stage the
    // last block is read again, past nv a vector is read clamped and masked
    // when used, and the f entry is read for every stage (used by SUB/SUP
    // only), so the waits stay counted instead of draining the prefetches
    auto fetch = [&](int s, float4 (&dst)[NVMAX], float& fv) __attribute__((always_inline)) {
      int i;
      const int k = decode(s < nst ? s : nst - 1, i);
      const float4* row = reinterpret_cast<const float4*>(block(k, i) + (int64_t)rc * bs + cb);
#pragma unroll
      for (int v = 0; v < NVMAX; ++v) dst[v] = row[v < nv ? v : nv - 1];
      fv = fptr(k == 1 ? i + 1 : i)[(int64_t)rc * ldf];
    };
    // LDS roles: X = input of the next DIAG stage, Y = last DIAG output
    int X = 0, Y = 1;
    if (tid < 3 * kInv1MaxBs) (&vb[0][0])[tid] = 0.f;  // entries past bs stay zero
    if (tid + 256 < 3 * kInv1MaxBs) (&vb[0][0])[tid + 256] = 0.f;
    __syncthreads();
    if (mode == 2) {
      if (tid < bs) vb[Y][tid] = fptr(m)[(int64_t)tid * ldf];  // the block past the end holds its solution
    } else {
      if (tid < bs) vb[X][tid] = fptr(0)[(int64_t)tid * ldf];
    }
    fetch(0, q0, f0);
    fetch(1, q1, f1);
    __syncthreads();
    // the LDS roles travel by value (st = X | Y << 2): captured by
    // reference they were kept in scratch, whose loads drain the prefetches
    auto stage = [&](int s, const float4 (&cur)[NVMAX], float fcur, int st) __attribute__((always_inline)) -> int {
      int X = st & 3, Y = st >> 2;
      int i;
      const int k = decode(s, i);
      const int in = (k == 0) ? X : ((k == 2) ? ((mode == 0 && i == m - 1) ? Y : X) : Y);
      const int o = 3 - X - Y;  // the buffer that is neither X nor Y
      const float4* iv = reinterpret_cast<const float4*>(&vb[in][cb]);
      float acc = 0.f;
#pragma unroll
      for (int v = 0; v < NVMAX; ++v) {
        const float4 x = iv[v];
        const float d = cur[v].x * x.x + cur[v].y * x.y + cur[v].z * x.z + cur[v].w * x.w;
        acc += (v < nv) ? d : 0.f;
      }
      acc += __shfl_xor(acc, 1);
      const float res = (k == 1 || k == 3) ? fcur - acc : acc;
      if (h == 0 && r < bs) {
        vb[o][r] = res;
        if (k == 0 || k == 2) fptr(i)[(int64_t)r * ldf] = res;
      }
      // DIAG stages produce Y, SUB/SUP stages produce X
      if (k == 0 || k == 2) Y = o;
      else X = o;
      __syncthreads();
      return X | (Y << 2);
    };
    int st = X | (Y << 2);
    for (int s = 0; s < nst; s += 3) {
      fetch(s + 2, q2, f2);
      st = stage(s, q0, f0, st);
      if (s + 1 >= nst) break;
      fetch(s + 3, q0, f0);
      st = stage(s + 1, q1, f1, st);
      if (s + 2 >= nst) break;
      fetch(s + 4, q1, f1);
      st = stage(s + 2, q2, f2, st);
    }
  }
}

template <typename T>
hipError_t solve_sets(int bs, const ChainSet<T>& c0, const ChainSet<T>& c1, int64_t ldf, int nrhs, int mode,
                      bool trans, hipStream_t s) {
  if constexpr (sizeof(T) == 4) {
    // 16-byte vectors: every block offset and base a multiple of 4 floats
    auto al = [](const ChainSet<T>& c) {
      if (c.nchains == 0) return true;
      const auto p4 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
      return c.linv != nullptr && p4(c.linv) && p4(c.diag) && p4(c.sub) && p4(c.sup) && c.dbs % 4 == 0 &&
             c.dis % 4 == 0 && c.lbs % 4 == 0 && c.lis % 4 == 0 && c.ubs % 4 == 0 && c.uis % 4 == 0;
    };
    if (!trans && nrhs == 1 && c0.linv != nullptr && al(c0) && al(c1) && bs <= kInv1MaxBs && bs % 4 == 0) {
      const int nt = c0.nchains + c1.nchains;
      if (nt <= 0) return hipSuccess;
      hipLaunchKernelGGL(k_btri_inv1, dim3(nt < 65535 ? nt : 65535), dim3(256), 0, s, bs,
                         reinterpret_cast<const ChainSet<float>&>(c0), reinterpret_cast<const ChainSet<float>&>(c1),
                         ldf, mode);
      return hipGetLastError();
    }
  }
  int nr = nrhs >= 32 ? 32 : (nrhs >= 16 ? 16 : (nrhs >= 8 ? 8 : (nrhs >= 4 ? 4 : (nrhs >= 2 ? 2 : 1))));
  // small blocks in substitution form: 8-column tiles (about 130 VGPRs, three
  // workgroups per CU) beat 32-column ones (one workgroup per CU)
  if (bs <= 32 && c0.linv == nullptr && nr > 8) nr = 8;
  // a tile width chosen by the autotuner for this block size, if any (never
  // wider than the columns there are)
  const int tuned = sweep_tile_override(sizeof(T) == 8 ? 0 : 1, bs, c0.linv != nullptr);
  if (tuned > 0 && nrhs > 1) {
    nr = tuned;
    while (nr > 1 && nr > nrhs) nr >>= 1;
  }
  while (nr > 1 && lu_solve_lds<T>(bs, nr) > kLdsMax) nr >>= 1;
  if (lu_solve_lds<T>(bs, nr) > kLdsMax || bs > (sizeof(T) == 8 ? 112 : 160)) return hipErrorInvalidValue;
  switch (nr) {
    case 1: return launch_lu_solve_pf<T, 1>(bs, c0, c1, ldf, nrhs, mode, trans, s);
    case 2: return launch_lu_solve_pf<T, 2>(bs, c0, c1, ldf, nrhs, mode, trans, s);
    case 4: return launch_lu_solve_pf<T, 4>(bs, c0, c1, ldf, nrhs, mode, trans, s);
    case 8: return launch_lu_solve_pf<T, 8>(bs, c0, c1, ldf, nrhs, mode, trans, s);
    case 16: return launch_lu_solve_pf<T, 16>(bs, c0, c1, ldf, nrhs, mode, trans, s);
    default: return launch_lu_solve_pf<T, 32>(bs, c0, c1, ldf, nrhs, mode, trans, s);
  }
}

}  // namespace

template <typename T>
hipError_t btri_lu_solve_fused(int nb, int m, int bs, T* diag, int64_t dbs, int64_t dis, T* sub, int64_t lbs,
                               int64_t lis, T* sup, int64_t ubs, int64_t uis, const int* ipiv, int64_t pbs,
                               int64_t pis, T* f, int64_t fbs, int64_t ldf, int nrhs, hipStream_t s, bool trans) {
  ChainSet<T> c0{diag, dbs, dis, sub, lbs, lis, sup, ubs, uis, const_cast<int*>(ipiv), pbs, pis, nullptr,
                 f, fbs, (int64_t)bs * ldf, m, nb, 0};
  ChainSet<T> none{};
  return solve_sets<T>(bs, c0, none, ldf, nrhs, 0, trans, s);
}

// Sweeps with the factors of btri_lu_factor_twisted: forward down both
// chains, the middle block's right-hand side less both chains' couplings,
// the middle block, then backward out along both chains.  The top chain's
// last Lfac/Ufac (towards the middle block) sit in sub_{k-1}/sup_{k-1}, the
// bottom chain's in sup_k/sub_k; the transposed sweep couples the middle
// block through the Ufac of both chains (transposed) on the way in.
template <typename T>
hipError_t btri_lu_solve_twisted(int nb, int m, int bs, T* diag, int64_t dbs, int64_t dis, T* sub, int64_t lbs,
                                 int64_t lis, T* sup, int64_t ubs, int64_t uis, const int* ipiv, int64_t pbs,
                                 int64_t pis, T* f, int64_t fbs, int64_t ldf, int nrhs, hipStream_t s,
                                 const T* linv, bool trans) {
  const int k = m / 2, lb = m - 1 - k;
  int* iv = const_cast<int*>(ipiv);
  const int64_t fis = (int64_t)bs * ldf;
  ChainSet<T> none{};
  ChainSet<T> top{diag, dbs, dis, sub, lbs, lis, sup, ubs, uis, iv, pbs, pis, nullptr, f, fbs, fis, k, nb, 1, linv};
  ChainSet<T> bot = mirrored<T>(nb, m, lb, bs, diag, dbs, dis, sub, lbs, lis, sup, ubs, uis, iv, pbs, pis, nullptr,
                                f, fbs, ldf, linv);
  hipError_t e;
  if (k > 0) {
    if ((e = solve_sets<T>(bs, top, bot, ldf, nrhs, 1, trans, s)) != hipSuccess) return e;
    View<T> Fk{f + (int64_t)k * fis, ldf, 1, fbs};
    // coupling of the top chain's last block into the middle block:
    // Lfac = sub_{k-1}, or Ufac^T = sup_{k-1}^T when transposed
    View<T> Ct = trans ? View<T>{sup + (int64_t)(k - 1) * uis, 1, bs, ubs} : View<T>{sub + (int64_t)(k - 1) * lis, bs, 1, lbs};
    if ((e = gemm<T>(nb, bs, nrhs, bs, T(-1), Ct, View<T>{f + (int64_t)(k - 1) * fis, ldf, 1, fbs}, T(1), Fk, s)) !=
        hipSuccess)
      return e;
    if (lb > 0) {
      // bottom chain: Lfac' = sup_k, or Ufac'^T = sub_k^T when transposed
      View<T> Cb = trans ? View<T>{sub + (int64_t)k * lis, 1, bs, lbs} : View<T>{sup + (int64_t)k * uis, bs, 1, ubs};
      if ((e = gemm<T>(nb, bs, nrhs, bs, T(-1), Cb, View<T>{f + (int64_t)(k + 1) * fis, ldf, 1, fbs}, T(1), Fk,
                       s)) != hipSuccess)
        return e;
    }
  }
  ChainSet<T> mid{diag + (int64_t)k * dis, dbs, dis, sub, lbs, lis, sup, ubs, uis, iv + (int64_t)k * pis, pbs, pis,
                  nullptr, f + (int64_t)k * fis, fbs, fis, 1, nb, 0, linv != nullptr ? linv + (int64_t)k * dis : nullptr};
  if ((e = solve_sets<T>(bs, mid, none, ldf, nrhs, 0, trans, s)) != hipSuccess) return e;
  if (k > 0) return solve_sets<T>(bs, top, bot, ldf, nrhs, 2, trans, s);
  return hipSuccess;
}

#define SLAB_BTRI_LU_SOLVE_INST(T)                                                                                \
  template hipError_t btri_lu_solve_fused<T>(int, int, int, T*, int64_t, int64_t, T*, int64_t, int64_t, T*,       \
                                             int64_t, int64_t, const int*, int64_t, int64_t, T*, int64_t, int64_t, \
                                             int, hipStream_t, bool);                                              \
  template hipError_t btri_lu_solve_twisted<T>(int, int, int, T*, int64_t, int64_t, T*, int64_t, int64_t, T*,     \
                                               int64_t, int64_t, const int*, int64_t, int64_t, T*, int64_t,        \
                                               int64_t, int, hipStream_t, const T*, bool);
SLAB_BTRI_LU_SOLVE_INST(double)
SLAB_BTRI_LU_SOLVE_INST(float)

}  // namespace slab
