// Batched dense linear algebra for gfx950 (CDNA4): MFMA-tiled GEMM,
// blocked triangular solves, row interchanges and a right-looking blocked
// LU whose panel is factored in LDS with a workgroup-wide pivot search.
// Every kernel takes strided views (row, column and batch strides) so that
// transposes, sub-blocks and broadcast operands need no copies.
#include "slab_internal.h"
#include "wave_ops.h"
#include "../../include/slab_hip.h"

#include <climits>

namespace slab {

static inline int ceil_div(int a, int b) { return (a + b - 1) / b; }
static inline int grid_batch(int batch) { return batch < 65535 ? batch : 65535; }

// ---------------------------------------------------------------------------
// GEMM: 64x64 output tile per 256-thread workgroup, 4 waves in a 2x2
// arrangement each owning 32x32 = 2x2 MFMA tiles; K staged through LDS in
// chunks of 16.
// ---------------------------------------------------------------------------
constexpr int GBM = 64, GBN = 64, GBK = 16;

// VEC: A and B row-major with 16-byte aligned rows and k, n multiples of
// the vector width: each thread moves 16-byte vectors of A (along k) and B
// (along n) instead of single elements.  VC (with VEC): C row-major and
// aligned too; the C tile is read as row vectors before the K loop and the
// result leaves through an LDS tile, so C moves as 16-byte row vectors.
// tiles > 0 (XCD-aware order, one-dimensional grid of tiles * 8 *
// ceil(batch / 8) workgroups): workgroups go to the eight XCDs round robin,
// so workgroup L runs on XCD L % 8; here the XCD's i-th workgroup (i = L / 8)
// takes tile i % tiles of matrix (L % 8) + 8 (i / tiles).  All tiles of one
// matrix then run on one XCD, one matrix after the other, and the operand
// rows and columns that several output tiles share are read from HBM once
// into that XCD's L2.  tiles = 0: tile blockIdx.x of matrices blockIdx.y,
// blockIdx.y + gridDim.y, ...
template <typename T, typename I, bool VEC, bool VC>
__global__ __launch_bounds__(256) void k_gemm(int batch, int m, int n, int k, T alpha, View<T> A, View<T> B,
                                              T beta, View<T> C, int tiles_n, int64_t ao, int64_t bo, int64_t co,
                                              int tiles) {
  typedef typename Mfma<T>::acc_t acc_t;
  constexpr int VW = VEC ? 16 / (int)sizeof(T) : 1;
  typedef T vec_t __attribute__((ext_vector_type(VW)));
  __shared__ T As[GBK][GBM + 1];
  __shared__ T Bs[GBK][GBN + 1];
  int tile = blockIdx.x, bfirst = blockIdx.y, bstep = gridDim.y;
  if (tiles > 0) {
    const int xi = blockIdx.x >> 3;
    tile = xi % tiles;
    bfirst = (blockIdx.x & 7) + 8 * (xi / tiles);
    bstep = batch;  // one matrix per workgroup
    if (bfirst >= batch) return;
  }
  const int m0 = (tile / tiles_n) * GBM, n0 = (tile % tiles_n) * GBN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const bool a_rowmajor = (A.cs == 1);
  const bool b_rowmajor = (B.cs == 1);
  const int mm1 = m - 1, nm1 = n - 1, km1 = k - 1;
  constexpr int NU = GBM * GBK / 256;
  constexpr int NV = NU / VW;  // loads per operand per thread
  constexpr int CPR = GBN / VW, NCV = GBM * GBN / VW / 256;  // C vectors per row / per thread
  __shared__ __attribute__((aligned(16))) T Cs[VC ? GBM : 1][VC ? GBN + VW : 1];
  // element offsets inside one matrix in I (32-bit whenever the host has
  // checked that they fit); row / column parts fixed per thread, so only the
  // K index moves between chunks.  Rows of A beyond m and columns of B
  // beyond n are read clamped and feed only C entries that are never
  // stored; K beyond k is zeroed.
  I a_off[NV], b_off[NV];
  int a_k[NV], b_k[NV];
#pragma unroll
  for (int u = 0; u < NV; ++u) {
    const int e = tid + 256 * u;
    int r, kk, c, kb;
    if (VEC) {
      r = e / (GBK / VW); kk = (e % (GBK / VW)) * VW;
      kb = e / (GBN / VW); c = (e % (GBN / VW)) * VW;
      a_off[u] = (I)(m0 + r < mm1 ? m0 + r : mm1) * (I)A.rs;
      b_off[u] = (I)(n0 + c < n ? n0 + c : n - VW);
    } else {
      if (a_rowmajor) { r = e / GBK; kk = e % GBK; } else { r = e % GBM; kk = e / GBM; }
      if (b_rowmajor) { kb = e / GBN; c = e % GBN; } else { kb = e % GBK; c = e / GBK; }
      a_off[u] = (I)(m0 + r < mm1 ? m0 + r : mm1) * (I)A.rs;
      b_off[u] = (I)(n0 + c < nm1 ? n0 + c : nm1) * (I)B.cs;
    }
    a_k[u] = kk;
    b_k[u] = kb;
  }
  I cvo[VC ? NCV : 1];
  if constexpr (VC) {
#pragma unroll
    for (int u = 0; u < NCV; ++u) {
      const int e = tid + 256 * u, cr = e / CPR, cc = (e % CPR) * VW;
      cvo[u] = (I)(m0 + cr < mm1 ? m0 + cr : mm1) * (I)C.rs + (I)(n0 + cc < n ? n0 + cc : n - VW);
    }
  }
  I c_row[2][4], c_col[2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = m0 + wm * 32 + i * 16 + Mfma<T>::row(lane, q);
      c_row[i][q] = (I)(row < mm1 ? row : mm1) * (I)C.rs;
    }
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int col = n0 + wn * 32 + j * 16 + (lane & 15);
    c_col[j] = (I)(col < nm1 ? col : nm1) * (I)C.cs;
  }
  for (int b = bfirst; b < batch; b += bstep) {
    // second batch level (grid z): whole-operand offsets
    const T* Ab = A.p + blockIdx.z * ao + b * A.bs;
    const T* Bb = B.p + blockIdx.z * bo + b * B.bs;
    T* Cb = C.p + blockIdx.z * co + b * C.bs;
    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
    // C is read before the K loop so that its latency overlaps the loop
    T cv[2][2][4];
    vec_t cvv[VC ? NCV : 1];
    if (VC && beta != T(0)) {
#pragma unroll
      for (int u = 0; u < (VC ? NCV : 1); ++u) cvv[u] = *reinterpret_cast<const vec_t*>(Cb + cvo[u]);
    } else if (!VC && beta != T(0)) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int q = 0; q < 4; ++q) cv[i][j][q] = Cb[c_row[i][q] + c_col[j]];
    }
    // operands of the next K chunk travel in registers while the current
    // chunk is multiplied
    T ra[NU], rb[NU];
    auto load = [&](int k0) {
#pragma unroll
      for (int u = 0; u < NV; ++u) {
        const int gk = k0 + a_k[u], gkb = k0 + b_k[u];
        if constexpr (VEC) {
          // k is a multiple of VW: a vector is wholly inside or past k
          const vec_t va = *reinterpret_cast<const vec_t*>(Ab + a_off[u] + (I)(gk < k ? gk : k - VW));
          const vec_t vb = *reinterpret_cast<const vec_t*>(Bb + b_off[u] + (I)(gkb < km1 ? gkb : km1) * (I)B.rs);
#pragma unroll
          for (int w = 0; w < VW; ++w) {
            ra[u * VW + w] = (gk < k) ? va[w] : T(0);
            rb[u * VW + w] = (gkb < k) ? vb[w] : T(0);
          }
        } else {
          const T va = Ab[a_off[u] + (I)(gk < km1 ? gk : km1) * (I)A.cs];
          ra[u] = (gk < k) ? va : T(0);
          const T vb = Bb[b_off[u] + (I)(gkb < km1 ? gkb : km1) * (I)B.rs];
          rb[u] = (gkb < k) ? vb : T(0);
        }
      }
    };
    if (k > 0) load(0);
    for (int k0 = 0; k0 < k; k0 += GBK) {
      if constexpr (VEC) {
#pragma unroll
        for (int u = 0; u < NV; ++u) {
          const int e = tid + 256 * u;
          const int r = e / (GBK / VW), kk = (e % (GBK / VW)) * VW;
          const int kb = e / (GBN / VW), c = (e % (GBN / VW)) * VW;
#pragma unroll
          for (int w = 0; w < VW; ++w) {
            As[kk + w][r] = ra[u * VW + w];
            Bs[kb][c + w] = rb[u * VW + w];
          }
        }
      } else {
#pragma unroll
        for (int u = 0; u < NU; ++u) {
          const int e = tid + 256 * u;
          if (a_rowmajor) As[e % GBK][e / GBK] = ra[u]; else As[e / GBM][e % GBM] = ra[u];
          if (b_rowmajor) Bs[e / GBN][e % GBN] = rb[u]; else Bs[e % GBK][e / GBK] = rb[u];
        }
      }
      __syncthreads();
      if (k0 + GBK < k) load(k0 + GBK);
#pragma unroll
      for (int kb = 0; kb < GBK; kb += 4) {
        T af[2], bf[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) af[i] = As[kb + (lane >> 4)][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < 2; ++j) bf[j] = Bs[kb + (lane >> 4)][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[i][j] = Mfma<T>::mma(af[i], bf[j], acc[i][j]);
      }
      __syncthreads();
    }
    if constexpr (VC) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int q = 0; q < 4; ++q)
            Cs[wm * 32 + i * 16 + Mfma<T>::row(lane, q)][wn * 32 + j * 16 + (lane & 15)] = acc[i][j][q];
      __syncthreads();
#pragma unroll
      for (int u = 0; u < NCV; ++u) {
        const int e = tid + 256 * u, cr = e / CPR, cc = (e % CPR) * VW;
        vec_t v = *reinterpret_cast<const vec_t*>(&Cs[cr][cc]);
#pragma unroll
        for (int w = 0; w < VW; ++w) v[w] = (beta != T(0)) ? alpha * v[w] + beta * cvv[u][w] : alpha * v[w];
        // n is a multiple of VW: a vector is wholly inside or past n
        if (m0 + cr < m && n0 + cc < n) *reinterpret_cast<vec_t*>(Cb + cvo[u]) = v;
      }
      __syncthreads();
    } else {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int col = n0 + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int row = m0 + wm * 32 + i * 16 + Mfma<T>::row(lane, q);
            if (row < m && col < n) {
              T v = alpha * acc[i][j][q];
              if (beta != T(0)) v += beta * cv[i][j][q];
              Cb[c_row[i][q] + c_col[j]] = v;
            }
          }
        }
    }
  }
}

// Few right-hand-side columns (n <= 8) with a row-major A: one wavefront per
// output row, lanes striding k (coalesced rows of A), shuffle reduction.
template <typename T, int NC>
__global__ __launch_bounds__(256) void k_gemv(int batch, int m, int n, int k, T alpha, View<T> A, View<T> B, T beta,
                                              View<T> C) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int b = blockIdx.y; b < batch; b += gridDim.y) {
    for (int row = blockIdx.x * 4 + wave; row < m; row += gridDim.x * 4) {
      T acc[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] = T(0);
      const T* arow = A.at(b, row, 0);
      const int nm1 = n - 1;
#pragma unroll 4
      for (int kk = lane; kk < k; kk += 64) {
        const T a = arow[kk];
        const T* brow = B.at(b, kk, 0);
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] += a * brow[(int64_t)(c < nm1 ? c : nm1) * B.cs];
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off);
      }
      if (lane < n) {
        T v = T(0);
#pragma unroll
        for (int c = 0; c < NC; ++c)
          if (c == lane) v = acc[c];
        T* cp = C.at(b, row, lane);
        T out = alpha * v;
        if (beta != T(0)) out += beta * (*cp);
        *cp = out;
      }
    }
  }
}

// Short rows (k <= 256): 16 lanes per output row, four rows per wavefront,
// so a row of 80 columns keeps every lane busy (the leaf recovery
// u_c -= S' d_x reads 400 x 80 per leaf).
template <typename T, int NC>
__global__ __launch_bounds__(256) void k_gemv16(int batch, int m, int n, int k, T alpha, View<T> A, View<T> B,
                                                T beta, View<T> C) {
  constexpr int LPR = 16;
  const int l = threadIdx.x & (LPR - 1), grp = threadIdx.x / LPR;
  for (int b = blockIdx.y; b < batch; b += gridDim.y) {
    for (int row0 = blockIdx.x * (256 / LPR); row0 < m; row0 += gridDim.x * (256 / LPR)) {
      const int row = row0 + grp;
      const int rc = row < m ? row : m - 1;
      T acc[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] = T(0);
      const T* arow = A.at(b, rc, 0);
      const int nm1 = n - 1;
#pragma unroll 4
      for (int kk = l; kk < k; kk += LPR) {
        const T a = arow[kk];
        const T* brow = B.at(b, kk, 0);
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] += a * brow[(int64_t)(c < nm1 ? c : nm1) * B.cs];
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
#pragma unroll
        for (int off = LPR / 2; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off);
      }
      if (row < m && l < n) {
        T v = T(0);
#pragma unroll
        for (int c = 0; c < NC; ++c)
          if (c == l) v = acc[c];
        T* cp = C.at(b, row, l);
        T out = alpha * v;
        if (beta != T(0)) out += beta * (*cp);
        *cp = out;
      }
    }
  }
}

// Two batch levels: `outer` groups of `batch` products, group g offset by
// g * (ao, bo, co) elements from the views.
template <typename T>
hipError_t gemm_outer(int outer, int64_t ao, int64_t bo, int64_t co, int batch, int m, int n, int k, T alpha,
                      View<T> A, View<T> B, T beta, View<T> C, hipStream_t s) {
  if (outer <= 0 || batch <= 0 || m <= 0 || n <= 0) return hipSuccess;
  if (outer > 65535) return hipErrorInvalidValue;
  if (k < 0) k = 0;
  const int tn = ceil_div(n, GBN), tm = ceil_div(m, GBM);
  dim3 grid(tm * tn, grid_batch(batch), outer);
  // large batches of several-tile products: XCD-aware order (a matrix's
  // tiles on one XCD); small batches keep every XCD busy in the plain order
  int xt = 0;
  if (outer == 1 && batch >= 16 && tm * tn > 1 && (int64_t)tm * tn * 8 * ceil_div(batch, 8) < INT_MAX) {
    xt = tm * tn;
    grid = dim3((unsigned)(xt * 8 * ceil_div(batch, 8)), 1, 1);
  }
  // 32-bit offsets inside each matrix when every element offset fits
  auto span = [](int rows, int cols, const View<T>& V) {
    const int64_t r = rows > 0 ? rows - 1 : 0, c = cols > 0 ? cols - 1 : 0;
    return r * (V.rs < 0 ? -V.rs : V.rs) + c * (V.cs < 0 ? -V.cs : V.cs);
  };
  const int64_t lim = INT_MAX;
  // 16-byte vectors of A and B: row-major, aligned rows and batch offsets,
  // k and n multiples of the vector width
  constexpr int VW = 16 / (int)sizeof(T);
  auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  const bool vec = A.cs == 1 && B.cs == 1 && A.rs % VW == 0 && B.rs % VW == 0 && A.bs % VW == 0 &&
                   B.bs % VW == 0 && ao % VW == 0 && bo % VW == 0 && al(A.p) && al(B.p) && k % VW == 0 &&
                   n % VW == 0 && k > 0 && (sizeof(T) == 4 || k >= 64);  // FP64 rank-16: measured slower
  // (FP64 products measured slower through the 33 KB C tile)
  const bool vc = vec && sizeof(T) == 4 && C.cs == 1 && C.rs % VW == 0 && C.bs % VW == 0 && co % VW == 0 &&
                  al(C.p);
  const bool small = span(m, k, A) < lim && span(k, n, B) < lim && span(m, n, C) < lim;
  // the launch probe times k_gemm<float, int, true, true> and
  // k_gemm<double, int, false, false>: 2 m n k flops per product, A, B and
  // C read once and C written once (C not read when beta is 0)
  const int pk = (sizeof(T) == 4 && small && vc) ? SLAB_PROBE_GEMM_F32
                 : (sizeof(T) == 8 && small && !vec) ? SLAB_PROBE_GEMM_F64 : 0;
  int slot = -1;
  if (pk != 0 && probe_on(pk)) slot = probe_begin(pk, s);
  if (small && vc)
    hipLaunchKernelGGL((k_gemm<T, int, true, true>), grid, dim3(256), 0, s, batch, m, n, k, alpha, A, B, beta, C, tn,
                       ao, bo, co, xt);
  else if (small && vec)
    hipLaunchKernelGGL((k_gemm<T, int, true, false>), grid, dim3(256), 0, s, batch, m, n, k, alpha, A, B, beta, C,
                       tn, ao, bo, co, xt);
  else if (small)
    hipLaunchKernelGGL((k_gemm<T, int, false, false>), grid, dim3(256), 0, s, batch, m, n, k, alpha, A, B, beta, C,
                       tn, ao, bo, co, xt);
  else
    hipLaunchKernelGGL((k_gemm<T, int64_t, false, false>), grid, dim3(256), 0, s, batch, m, n, k, alpha, A, B, beta,
                       C, tn, ao, bo, co, xt);
  const hipError_t e = hipGetLastError();
  if (slot >= 0) {
    const double prods = (double)outer * batch;
    probe_end(slot, s, 2.0 * prods * m * n * k,
              prods * sizeof(T) * ((double)m * k + (double)k * n + (beta != T(0) ? 2.0 : 1.0) * m * n));
  }
  return e;
}

// C[b] = alpha A B[b] + beta C[b] for one A shared by every matrix (FP32,
// m x k row-major with 16-byte rows, k % 4 == 0) and small outputs (m <=
// 16 MT, n <= 16 NTL), as the leaf DtN product N_x S' (80 x 400 by 400 x
// 80): one workgroup per matrix, the output's 16 x 16 tiles dealt
// round-robin to the four wavefronts (a slot past the last tile repeats it
// and is not stored), K in 16-wide chunks with the permuted K pairing of
// k_trsm_upper_mfma (a lane's A values of a chunk are one 16-byte load);
// the next chunk's operands load while this chunk's MFMAs run.
template <int MT, int NTL>
__global__ __launch_bounds__(256) void k_gemm_bcast_small(int batch, int m, int n, int k, float alpha, View<float> A,
                                                          View<float> B, float beta, View<float> C) {
  constexpr int NTILE = MT * NTL, TPW = (NTILE + 3) / 4;
  typedef float v4 __attribute__((ext_vector_type(4)));
  typedef Mfma<float>::acc_t acc_t;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i16 = lane & 15, kq = lane >> 4;
  int arow[TPW], bcol[TPW];
#pragma unroll
  for (int j = 0; j < TPW; ++j) {
    const int t0 = wave + 4 * j, t = t0 < NTILE ? t0 : NTILE - 1;
    const int r = (t / NTL) * 16 + i16, c = (t % NTL) * 16 + i16;
    arow[j] = r < m ? r : m - 1;
    bcol[j] = c < n ? c : n - 1;
  }
  for (int b = blockIdx.x; b < batch; b += gridDim.x) {
    acc_t acc[TPW];
#pragma unroll
    for (int j = 0; j < TPW; ++j) acc[j] = acc_t{0.f, 0.f, 0.f, 0.f};
    v4 an[TPW];
    float bn[4][TPW];
    auto fetch = [&](int k0) __attribute__((always_inline)) {
      const int kb = k0 + 4 * kq;
      const int kc = kb < k ? kb : k - 4;
#pragma unroll
      for (int j = 0; j < TPW; ++j) an[j] = *reinterpret_cast<const v4*>(A.at(0, arow[j], kc));
#pragma unroll
      for (int st = 0; st < 4; ++st)
#pragma unroll
        for (int j = 0; j < TPW; ++j) bn[st][j] = *B.at(b, kc + st, bcol[j]);
    };
    fetch(0);
    for (int k0 = 0; k0 < k; k0 += 16) {
      v4 av[TPW];
      float bv[4][TPW];
#pragma unroll
      for (int j = 0; j < TPW; ++j) {
        av[j] = an[j];
#pragma unroll
        for (int st = 0; st < 4; ++st) bv[st][j] = bn[st][j];
      }
      if (k0 + 16 < k) fetch(k0 + 16);
      const bool kin = k0 + 4 * kq < k;
#pragma unroll
      for (int st = 0; st < 4; ++st)
#pragma unroll
        for (int j = 0; j < TPW; ++j) acc[j] = Mfma<float>::mma(kin ? av[j][st] : 0.f, bv[st][j], acc[j]);
    }
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
      const int t = wave + 4 * j;
      if (t < NTILE) {
        const int c = (t % NTL) * 16 + i16;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = (t / NTL) * 16 + Mfma<float>::row(lane, q);
          if ((r < m) & (c < n)) {
            float* cp = C.at(b, r, c);
            *cp = beta != 0.f ? alpha * acc[j][q] + beta * *cp : alpha * acc[j][q];
          }
        }
      }
    }
  }
}

template <typename T>
static bool gemm_bcast_small(int batch, int m, int n, int k, T alpha, View<T> A, View<T> B, T beta, View<T> C,
                             hipStream_t s, hipError_t* err) {
  if constexpr (sizeof(T) != 4) {
    return false;
  } else {
    if (A.bs != 0 || batch < 64 || m < 49 || m > 80 || n < 49 || n > 80 || k < 64 || k % 4 != 0 || A.cs != 1 ||
        (A.rs & 3) != 0 || (reinterpret_cast<uintptr_t>(A.p) & 15) != 0)
      return false;
    const int mt = (m + 15) / 16, nt = (n + 15) / 16;
    const int grid = batch < 65535 ? batch : 65535;
#define SLAB_GEMM_BS(MT_, NT_)                                                                                    \
  if (mt == MT_ && nt == NT_) {                                                                                   \
    hipLaunchKernelGGL((k_gemm_bcast_small<MT_, NT_>), dim3(grid), dim3(256), 0, s, batch, m, n, k, alpha, A, B,  \
                       beta, C);                                                                                  \
    *err = hipGetLastError();                                                                                     \
    return true;                                                                                                  \
  }
    SLAB_GEMM_BS(4, 4)
    SLAB_GEMM_BS(4, 5)
    SLAB_GEMM_BS(5, 4)
    SLAB_GEMM_BS(5, 5)
#undef SLAB_GEMM_BS
    return false;
  }
}

template <typename T>
hipError_t gemm(int batch, int m, int n, int k, T alpha, View<T> A, View<T> B, T beta, View<T> C, hipStream_t s) {
  if (batch <= 0 || m <= 0 || n <= 0) return hipSuccess;
  hipError_t bs_err;
  if (gemm_bcast_small<T>(batch, m, n, k, alpha, A, B, beta, C, s, &bs_err)) return bs_err;
  if (n <= 8 && A.cs == 1 && k >= 32 && k <= 256) {
    const int gx = (m + 15) / 16 < 8192 ? (m + 15) / 16 : 8192;
    dim3 grid(gx, grid_batch(batch));
    if (n <= 1)
      hipLaunchKernelGGL((k_gemv16<T, 1>), grid, dim3(256), 0, s, batch, m, n, k, alpha, A, B, beta, C);
    else if (n <= 4)
      hipLaunchKernelGGL((k_gemv16<T, 4>), grid, dim3(256), 0, s, batch, m, n, k, alpha, A, B, beta, C);
    else
      hipLaunchKernelGGL((k_gemv16<T, 8>), grid, dim3(256), 0, s, batch, m, n, k, alpha, A, B, beta, C);
    return hipGetLastError();
  }
  if (n <= 8 && A.cs == 1 && k >= 32) {
    const int gx = (m + 3) / 4 < 8192 ? (m + 3) / 4 : 8192;
    dim3 grid(gx, grid_batch(batch));
    if (n <= 1)
      hipLaunchKernelGGL((k_gemv<T, 1>), grid, dim3(256), 0, s, batch, m, n, k, alpha, A, B, beta, C);
    else if (n <= 4)
      hipLaunchKernelGGL((k_gemv<T, 4>), grid, dim3(256), 0, s, batch, m, n, k, alpha, A, B, beta, C);
    else
      hipLaunchKernelGGL((k_gemv<T, 8>), grid, dim3(256), 0, s, batch, m, n, k, alpha, A, B, beta, C);
    return hipGetLastError();
  }
  // k <= 0: C = beta * C through a zero-length K loop
  return gemm_outer<T>(1, 0, 0, 0, batch, m, n, k, alpha, A, B, beta, C, s);
}

// ---------------------------------------------------------------------------
// Triangular solve of a <= 32 x 32 diagonal block: one thread per right-hand
// side column (one wavefront per workgroup), the block in LDS padded with
// the identity, the column in registers; all loads unconditional (clamped).
// ---------------------------------------------------------------------------
constexpr int TB = 32;

template <typename T>
__global__ __launch_bounds__(64) void k_trsm_diag(int batch, int n, int d0, int tb, int lower, int unit, View<T> A,
                                                  View<T> B) {
  __shared__ T Ls[TB][TB + 1];
  const int tid = threadIdx.x;
  const int col = blockIdx.x * 64 + tid;
  const int cl = col < n ? col : n - 1;
  const int tm1 = tb - 1;
  for (int b = blockIdx.y; b < batch; b += gridDim.y) {
#pragma unroll
    for (int u = 0; u < TB * TB / 64; ++u) {
      const int e = tid + 64 * u, r = e / TB, c = e % TB;
      const T v = *A.at(b, d0 + (r < tm1 ? r : tm1), d0 + (c < tm1 ? c : tm1));
      Ls[r][c] = ((r < tb) & (c < tb)) ? v : ((r == c) ? T(1) : T(0));
    }
    __syncthreads();
    T x[TB];
#pragma unroll
    for (int r = 0; r < TB; ++r) {
      const T v = *B.at(b, d0 + (r < tm1 ? r : tm1), cl);
      x[r] = (r < tb) ? v : T(0);
    }
    if (lower) {
#pragma unroll
      for (int i = 0; i < TB; ++i) {
        T v = x[i];
#pragma unroll
        for (int j = 0; j < i; ++j) v -= Ls[i][j] * x[j];
        x[i] = unit ? v : v / Ls[i][i];
      }
    } else {
#pragma unroll
      for (int i = TB - 1; i >= 0; --i) {
        T v = x[i];
#pragma unroll
        for (int j = i + 1; j < TB; ++j) v -= Ls[i][j] * x[j];
        x[i] = unit ? v : v / Ls[i][i];
      }
    }
    if (col < n) {
#pragma unroll
      for (int r = 0; r < TB; ++r)
        if (r < tb) *B.at(b, d0 + r, col) = x[r];
    }
    __syncthreads();
  }
}

// X <- U^{-1} X for large FP32 batches with U upper (non-unit diagonal,
// m x m, m % 4 == 0) and X m x n (n <= 16 NCT, n % 4 == 0, rows of both
// 16-byte aligned): one workgroup per matrix
// with all of X in LDS, 32-row blocks from the bottom, left-looking:
//   Y_k = X_k - U[k, k+1:] X[k+1:]   (MFMA, U read from HBM once),
//   X_k = U_kk^{-1} Y_k              (substitution, one thread per column).
// The K index of each 16 x 16 x 4 MFMA step is permuted so that a lane's
// four U values of a 16-wide K chunk are one 16-byte load: at step s, lane
// l pairs U[i][k0 + 4 (l / 16) + s] with X[k0 + 4 (l / 16) + s][j].  Wave w
// owns row tile w & 1 of the block and column tiles w / 2, w / 2 + 2, ...
// Rows past m are zero padding and stay zero (the padded diagonal block
// is the identity).  A lane's U values of a whole block are loaded at once,
// and the next block's during this block's substitution.
constexpr int kTrsmBlk = 32;
constexpr size_t kTrsmLdsMax = 160 * 1024;  // LDS per workgroup on gfx950
constexpr int kTrsmMaxCh = 26;              // 16-wide K chunks of a block's product: m <= 448

__host__ __device__ inline int trsm_up_ldx(int ntile) { return 16 * ntile; }
__host__ __device__ inline size_t trsm_up_lds(int m, int ntile) {
  const int mp = ((m + kTrsmBlk - 1) / kTrsmBlk) * kTrsmBlk;
  return sizeof(float) * ((size_t)mp * trsm_up_ldx(ntile) + kTrsmBlk * (kTrsmBlk + 1));
}

template <int NCT>
__global__ __launch_bounds__(256) void k_trsm_upper_mfma(int batch, int m, int n, View<float> A, View<float> B) {
  extern __shared__ float lds_f[];
  constexpr int ldx = 16 * NCT;
  constexpr int CPW = (NCT + 1) / 2;  // column tiles per wavefront, at most
  typedef float v4 __attribute__((ext_vector_type(4)));
  typedef Mfma<float>::acc_t acc_t;
  const int mp = ((m + kTrsmBlk - 1) / kTrsmBlk) * kTrsmBlk;
  float* Xs = lds_f;
  float* Us = Xs + mp * ldx;  // kTrsmBlk x (kTrsmBlk + 1)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i16 = lane & 15, kq = lane >> 4;
  const int rt = wave & 1, ct0 = wave >> 1;
  const int mm1 = m - 1;
  // this lane's U values of a block's product, all K chunks at once: rows
  // R0 + 16 rt + (l % 16), columns k0s + 16 ch + 4 (l / 16) .. + 3 (only the
  // block's chunks are loaded; the last chunk's columns past m read a
  // clamped address and are never used)
  v4 a[kTrsmMaxCh];
  for (int b = blockIdx.x; b < batch; b += gridDim.x) {
    // the block's diagonal block: thread t holds row t / 8, columns
    // 4 (t % 8) .. + 3 (clamped; masked when stored)
    v4 ud;
    auto fetch = [&](int R0) __attribute__((always_inline)) {
      const int row = R0 + rt * 16 + i16;
      const float* arow = A.at(b, row < mm1 ? row : mm1, 0);
      const int k0s = R0 + kTrsmBlk;
      const int nchf = k0s < m ? (m - k0s + 15) / 16 : 0;  // same for every lane
#pragma unroll
      for (int ch = 0; ch < kTrsmMaxCh; ++ch) {
        if (ch < nchf) {
          const int kb = k0s + ch * 16 + 4 * kq;
          a[ch] = *reinterpret_cast<const v4*>(arow + (kb < m ? kb : m - 4));
        }
      }
      const int dr = R0 + (tid >> 3), dc = R0 + 4 * (tid & 7);
      ud = *reinterpret_cast<const v4*>(A.at(b, dr < mm1 ? dr : mm1, dc < m ? dc : m - 4));
    };
    fetch(mp - kTrsmBlk);
    // X into LDS as 16-byte vectors, 8 loads in flight per thread (columns
    // past n and rows past m as zeros)
    {
      constexpr int Q = ldx / 4;
      const int nq = mp * Q;
      const int nv4 = (n + 3) / 4;
      for (int base = 0; base < nq; base += 8 * 256) {
        v4 t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int e = base + u * 256 + tid;
          const int r = e / Q, c4 = e % Q;
          const int rc = r < mm1 ? r : mm1, cc = c4 < nv4 - 1 ? c4 : nv4 - 1;
          t[u] = *reinterpret_cast<const v4*>(B.at(b, rc, 4 * cc));
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int e = base + u * 256 + tid;
          const int r = e / Q, c4 = e % Q;
          if (e < nq) {
#pragma unroll
            for (int q = 0; q < 4; ++q) Xs[r * ldx + 4 * c4 + q] = ((r < m) & (4 * c4 + q < n)) ? t[u][q] : 0.f;
          }
        }
      }
    }
    __syncthreads();
    for (int R0 = mp - kTrsmBlk; R0 >= 0; R0 -= kTrsmBlk) {
      // the diagonal block (fetched with the block's U values), identity
      // past m
      {
        const int r = tid >> 3, c0 = 4 * (tid & 7);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int rr = R0 + r, cc = R0 + c0 + q;
          Us[r * (kTrsmBlk + 1) + c0 + q] = ((rr < m) & (cc < m)) ? ud[q] : ((r == c0 + q) ? 1.f : 0.f);
        }
      }
      const int k0s = R0 + kTrsmBlk;
      if (k0s < m) {
        acc_t acc[CPW];
#pragma unroll
        for (int j = 0; j < CPW; ++j) acc[j] = acc_t{0.f, 0.f, 0.f, 0.f};
        const int nch = (m - k0s + 15) / 16;
        int ctc[CPW];
#pragma unroll
        for (int j = 0; j < CPW; ++j) ctc[j] = ct0 + 2 * j < NCT ? ct0 + 2 * j : NCT - 1;
#pragma unroll
        for (int ch = 0; ch < kTrsmMaxCh; ++ch) {
          if (ch < nch) {
            const int kb = k0s + ch * 16 + 4 * kq;
            const bool kin = kb < m;
            const int kr = kin ? kb : mm1;
#pragma unroll
            for (int st = 0; st < 4; ++st) {
              const float av = kin ? a[ch][st] : 0.f;
              const float* xr = Xs + (kr + (kin ? st : 0)) * ldx + i16;
              // every slot runs (a slot past the last column tile reads a
              // clamped tile and is not stored): no branch per MFMA
#pragma unroll
              for (int j = 0; j < CPW; ++j) acc[j] = Mfma<float>::mma(av, xr[ctc[j] * 16], acc[j]);
            }
          }
        }
#pragma unroll
        for (int j = 0; j < CPW; ++j) {
          const int ct = ct0 + 2 * j;
          if (ct < NCT) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int r = R0 + rt * 16 + Mfma<float>::row(lane, q);
              if (r < m) Xs[r * ldx + ct * 16 + i16] -= acc[j][q];
            }
          }
        }
      }
      // the next block's U values load during this block's substitution
      if (R0 > 0) fetch(R0 - kTrsmBlk);
      __syncthreads();
      if (tid < n) {
        float x[kTrsmBlk];
#pragma unroll
        for (int r = 0; r < kTrsmBlk; ++r) x[r] = Xs[(R0 + r) * ldx + tid];
        // column-oriented: x_j is final once divided, then leaves every row
        // above at once (the dependent chain is one division and one FMA per
        // row, not a row's whole dot product)
#pragma unroll
        for (int j = kTrsmBlk - 1; j >= 0; --j) {
          x[j] = x[j] / Us[j * (kTrsmBlk + 1) + j];
#pragma unroll
          for (int i = 0; i < j; ++i) x[i] -= Us[i * (kTrsmBlk + 1) + j] * x[j];
        }
#pragma unroll
        for (int r = 0; r < kTrsmBlk; ++r) Xs[(R0 + r) * ldx + tid] = x[r];
      }
      __syncthreads();
    }
    for (int e = tid; e < m * ldx; e += 256) {
      const int r = e / ldx, c = e % ldx;
      if (c < n) *B.at(b, r, c) = Xs[e];
    }
    __syncthreads();
  }
}

// the fused upper solve when it applies (FP32, many matrices, X in LDS)
template <typename T>
static bool trsm_upper_fused(int batch, int m, int n, View<T> A, View<T> B, hipStream_t s, hipError_t* err) {
  if constexpr (sizeof(T) != 4) {
    return false;
  } else {
    const int nct = (n + 15) / 16;
    if (batch < 64 || m % 4 != 0 || m < kTrsmBlk || m > kTrsmBlk + 16 * kTrsmMaxCh || n > 256 || A.cs != 1 || B.cs != 1 || (A.rs & 3) != 0 ||
        (reinterpret_cast<uintptr_t>(A.p) & 15) != 0 || (batch > 1 && (A.bs & 3) != 0) || nct < 4 || nct > 6 ||
        n % 4 != 0 || (B.rs & 3) != 0 || (reinterpret_cast<uintptr_t>(B.p) & 15) != 0 ||
        (batch > 1 && (B.bs & 3) != 0) ||
        trsm_up_lds(m, nct) > kTrsmLdsMax)
      return false;
    const size_t lds = trsm_up_lds(m, nct);
    const int grid = batch < 65535 ? batch : 65535;
#define SLAB_TRSM_UP(NCT_)                                                                                        \
  if (nct == NCT_) {                                                                                              \
    static bool attr = false;                                                                                     \
    if (!attr) {                                                                                                  \
      if ((*err = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_trsm_upper_mfma<NCT_>),                   \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTrsmLdsMax)) != hipSuccess) \
        return true;                                                                                              \
      attr = true;                                                                                                \
    }                                                                                                             \
    hipLaunchKernelGGL(k_trsm_upper_mfma<NCT_>, dim3(grid), dim3(256), lds, s, batch, m, n, A, B);               \
    *err = hipGetLastError();                                                                                     \
    return true;                                                                                                  \
  }
    SLAB_TRSM_UP(4)
    SLAB_TRSM_UP(5)
    SLAB_TRSM_UP(6)
#undef SLAB_TRSM_UP
    return false;
  }
}

template <typename T>
hipError_t trsm_left(int batch, int m, int n, bool lower, bool unit, View<T> A, View<T> B, hipStream_t s) {
  if (batch <= 0 || m <= 0 || n <= 0) return hipSuccess;
  dim3 grid(ceil_div(n, 64), grid_batch(batch));
  hipError_t e = hipSuccess;
  if (!lower && !unit && trsm_upper_fused<T>(batch, m, n, A, B, s, &e)) return e;
  if (lower) {
    for (int d0 = 0; d0 < m; d0 += TB) {
      const int tb = (m - d0) < TB ? (m - d0) : TB;
      hipLaunchKernelGGL(k_trsm_diag<T>, grid, dim3(64), 0, s, batch, n, d0, tb, 1, unit ? 1 : 0, A, B);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      if (d0 + tb < m) {
        e = gemm<T>(batch, m - d0 - tb, n, tb, T(-1), A.sub(d0 + tb, d0), B.sub(d0, 0), T(1), B.sub(d0 + tb, 0), s);
        if (e != hipSuccess) return e;
      }
    }
  } else {
    const int last = ((m - 1) / TB) * TB;
    for (int d0 = last; d0 >= 0; d0 -= TB) {
      const int tb = (m - d0) < TB ? (m - d0) : TB;
      hipLaunchKernelGGL(k_trsm_diag<T>, grid, dim3(64), 0, s, batch, n, d0, tb, 0, unit ? 1 : 0, A, B);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      if (d0 > 0) {
        e = gemm<T>(batch, d0, n, tb, T(-1), A.sub(0, d0), B.sub(d0, 0), T(1), B, s);
        if (e != hipSuccess) return e;
      }
    }
  }
  return hipSuccess;
}

// ---------------------------------------------------------------------------
// Row interchanges: one thread per column applies the pivot sequence.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_laswp(int batch, int ncols, int k1, int k2, int forward, View<T> A,
                                               const int* ipiv, int64_t pbs, int skip0, int skip1) {
  const int col = blockIdx.x * 256 + threadIdx.x;
  if (col >= ncols || (col >= skip0 && col < skip1)) return;
  for (int b = blockIdx.y; b < batch; b += gridDim.y) {
    const int* pv = ipiv + (int64_t)b * pbs;
    if (forward) {
      for (int k = k1; k < k2; ++k) {
        const int p = pv[k] - 1;
        if (p != k) {
          T* x = A.at(b, k, col);
          T* y = A.at(b, p, col);
          T t = *x; *x = *y; *y = t;
        }
      }
    } else {
      for (int k = k2 - 1; k >= k1; --k) {
        const int p = pv[k] - 1;
        if (p != k) {
          T* x = A.at(b, k, col);
          T* y = A.at(b, p, col);
          T t = *x; *x = *y; *y = t;
        }
      }
    }
  }
}

template <typename T>
hipError_t laswp(int batch, int ncols, int k1, int k2, bool forward, View<T> A, const int* ipiv, int64_t pbs,
                 hipStream_t s) {
  if (batch <= 0 || ncols <= 0 || k2 <= k1) return hipSuccess;
  dim3 grid(ceil_div(ncols, 256), grid_batch(batch));
  hipLaunchKernelGGL(k_laswp<T>, grid, dim3(256), 0, s, batch, ncols, k1, k2, forward ? 1 : 0, A, ipiv, pbs, 0, 0);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// LU panel: rows k0..nrows-1, columns k0..k0+w-1 of one matrix per
// workgroup, held column-major in LDS.  Pivot search: per-thread scan,
// wavefront butterfly reduction, then across the 4 waves; ties go to the
// smallest row (LAPACK idamax).
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline T tabs(T v) { return v < T(0) ? -v : v; }

template <typename T>
__global__ __launch_bounds__(256) void k_getf2_panel(int batch, int nrows, int k0, int w, int npiv, View<T> A,
                                                     int* ipiv, int64_t pbs, int* info) {
  extern __shared__ unsigned char smem_raw[];
  T* Ps = reinterpret_cast<T*>(smem_raw);
  __shared__ T red_v[4];
  __shared__ int red_i[4];
  __shared__ int s_piv;
  const int R = nrows - k0;
  const int Rp = npiv - k0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int b = blockIdx.x; b < batch; b += gridDim.x) {
    for (int e = tid; e < R * w; e += 256) {
      const int r = e / w, c = e - r * w;
      Ps[c * R + r] = *A.at(b, k0 + r, k0 + c);
    }
    __syncthreads();
    for (int j = 0; j < w; ++j) {
      T best = T(-1);
      int bi = INT_MAX;
      for (int r = j + tid; r < Rp; r += 256) {
        const T v = tabs(Ps[j * R + r]);
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
        if (bx < j || bx >= Rp) bx = j;  // no valid candidate (e.g. NaN column): keep the diagonal
        s_piv = bx;
        ipiv[(int64_t)b * pbs + k0 + j] = k0 + bx + 1;
      }
      __syncthreads();
      const int pv = s_piv;
      if (pv != j && tid < w) {
        T t = Ps[tid * R + j];
        Ps[tid * R + j] = Ps[tid * R + pv];
        Ps[tid * R + pv] = t;
      }
      __syncthreads();
      const T piv = Ps[j * R + j];
      const bool nz = piv != T(0);
      if (!nz && tid == 0 && info != nullptr && info[b] == 0) info[b] = k0 + j + 1;
      // each thread scales and updates whole rows: no index arithmetic in
      // the inner loop, the pivot row is a broadcast read
      for (int r = j + 1 + tid; r < R; r += 256) {
        T l = Ps[j * R + r];
        if (nz) {
          l = l / piv;
          Ps[j * R + r] = l;
        }
        for (int c = j + 1; c < w; ++c) Ps[c * R + r] -= l * Ps[c * R + j];
      }
      __syncthreads();
    }
    for (int e = tid; e < R * w; e += 256) {
      const int r = e / w, c = e - r * w;
      *A.at(b, k0 + r, k0 + c) = Ps[c * R + r];
    }
    __syncthreads();
  }
}

// Row interchanges of one LU panel (rows k0..k0+w-1 with 1-based pivots
// ipiv[k0..k0+w-1] >= row, w <= 32) on every column outside [skip0, skip1).
// The swap sequence is replayed once on row ids; each column then moves its
// (at most 2w) touched entries with independent loads and stores.
template <typename T>
__global__ __launch_bounds__(256) void k_laswp_panel(int batch, int ncols, int k0, int w, View<T> A,
                                                     const int* ipiv, int64_t pbs, int skip0, int skip1) {
  __shared__ int s_slot[32];
  __shared__ int s_row[64];
  __shared__ int s_src[64];
  __shared__ int s_cnt;
  const int tid = threadIdx.x;
  const int col = blockIdx.x * 256 + tid;
  for (int b = blockIdx.y; b < batch; b += gridDim.y) {
    // 1. slots: block rows k0+i -> i; pivot rows below the block -> w, w+1,
    //    ... in order of first occurrence
    bool first = false;
    int p = -1, src_lane = 0;
    unsigned long long mask = 0;
    if (tid < 64) {
      const int i = tid;
      p = (i < w) ? ipiv[(int64_t)b * pbs + k0 + i] - 1 : -1;
      first = (i < w) && (p >= k0 + w);
      src_lane = i;
      for (int q = 0; q < i && q < w; ++q) {
        const int pq = ipiv[(int64_t)b * pbs + k0 + q] - 1;
        if (first && pq == p) { first = false; src_lane = q; }
      }
      mask = __ballot(first);
      const int rank = __popcll(mask & ((1ull << i) - 1ull));
      if (first) {
        s_slot[i] = w + rank;
        s_row[w + rank] = p;
      }
      if (i < w) s_row[i] = k0 + i;
      if (i == 0) s_cnt = w + __popcll(mask);
    }
    __syncthreads();
    if (tid < w) {
      if (p < k0 + w) s_slot[tid] = p - k0;
      else if (!first) s_slot[tid] = s_slot[src_lane];
    }
    __syncthreads();
    // 2. replay the swap sequence on slots
    if (tid == 0) {
      const int cnt = s_cnt;
      for (int q = 0; q < cnt; ++q) s_src[q] = q;
      for (int q = 0; q < w; ++q) {
        const int c = s_slot[q];
        const int t = s_src[q];
        s_src[q] = s_src[c];
        s_src[c] = t;
      }
    }
    __syncthreads();
    const int cnt = s_cnt;
    if (col < ncols && !(col >= skip0 && col < skip1)) {
      T v[64];
#pragma unroll
      for (int q = 0; q < 64; ++q)
        if (q < cnt) v[q] = *A.at(b, s_row[s_src[q]], col);
#pragma unroll
      for (int q = 0; q < 64; ++q)
        if (q < cnt) *A.at(b, s_row[q], col) = v[q];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// LU panel in registers: rows k0..nrows-1 (R of them), columns k0..k0+w-1
// (w <= W) of one matrix per NT-thread workgroup; thread t holds panel rows
// t + NT s (s < RP) across all W columns.
//
// The column loop is bound by instruction issue: at the interface's batch of
// two matrices one workgroup per CU does all of the work, and in the leaf
// batches the column loops of all workgroups share the SIMDs.  So few
// wavefronts hold several rows each, and every step is written to issue as
// few instructions per wavefront as it can:
//  * rows never move between registers.  Each thread keeps its rows'
//    logical positions (where the interchanges so far have sent them) and
//    whether each is a pivot row already; rows at or past npiv (and padding
//    rows) are flagged once as never chosen, and they are never interchanged;
//  * per column every thread folds its candidate rows into one 64-bit key
//    (largest |a|, ties to the smaller logical position, as LAPACK's i?amax),
//    the wavefront reduces the keys with DPP, and the one lane holding the
//    wavefront's winner stores that row to LDS (a branch that one lane
//    takes, not a blend over all rows);
//  * after one barrier every wavefront merges the wavefront keys, reads the
//    pivot row, relabels the two interchanged positions, and scales and
//    updates its remaining rows in registers.
// Candidate buffers alternate between two slots, so writes for column j+1
// never meet reads of column j.  A column without candidates (all NaN) gets
// a NaN pivot, so everything below it becomes NaN.  The first zero pivot is
// kept in a register and reported once per panel.  The panel moves between
// HBM and registers through a per-wavefront LDS tile (row-contiguous global
// accesses); on the way out each row goes to its final logical position.
// ---------------------------------------------------------------------------
template <typename T, int W, int RP, int NT>
__global__ __launch_bounds__(NT) void k_getf2_reg(int batch, int nrows, int k0, int w, int npiv, View<T> A,
                                                  int* ipiv, int64_t pbs, int* info, bool vld, bool vst) {
  constexpr int NW = NT / 64;
  constexpr int TW = (sizeof(T) == 8 && W > 16) ? 16 : W;  // staging tile width
  static_assert(W % TW == 0 && NT % 64 == 0 && NT <= 1024, "panel shape");
  __shared__ T cand[2][NW][W];
  __shared__ unsigned long long red_k[2][NW];
  __shared__ T stage[NW][64][TW + 1];
  __shared__ int stage_pos[NW][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int R = nrows - k0, Rp = npiv - k0;
  const int rm1 = R - 1, wm1 = w - 1;
  for (int b = blockIdx.x; b < batch; b += gridDim.x) {
    T x[RP][W];
    // logical position of each row: the pivot rows of columns 0..j-1 hold
    // 0..j-1, the candidates of column j sit in j..Rp-1, padding rows at R..
    int lp[RP];
#pragma unroll
    for (int s = 0; s < RP; ++s) lp[s] = tid + NT * s;
    constexpr int VN = 16 / sizeof(T);
    typedef T vt __attribute__((ext_vector_type(VN)));
    if (vld) {
      // 16-byte aligned rows with W columns in bounds: every row of the
      // panel moves as W / VN vector loads, all of them in flight at once
#pragma unroll
      for (int s = 0; s < RP; ++s) {
        const int r = tid + NT * s;
        const vt* src = reinterpret_cast<const vt*>(A.at(b, k0 + (r < rm1 ? r : rm1), k0));
#pragma unroll
        for (int q = 0; q < W / VN; ++q) {
          const vt v = src[q];
#pragma unroll
          for (int e = 0; e < VN; ++e) x[s][q * VN + e] = ((lp[s] < R) & (q * VN + e < w)) ? v[e] : T(0);
        }
      }
    } else {
      // coalesced load: a wavefront moves its 64 rows of slot s in TW-wide
      // column tiles, 64 / TW rows per load instruction
      static_for<0, RP>([&](auto sc) __attribute__((always_inline)) {
        constexpr int sl = decltype(sc)::value;
        const int r0 = NT * sl + 64 * wave;
        static_for<0, W / TW>([&](auto tc) __attribute__((always_inline)) {
          constexpr int c0 = decltype(tc)::value * TW;
  #pragma unroll
          for (int q = 0; q < TW; ++q) {
            const int e = 64 * q + lane;
            const int rr = r0 + e / TW, cc = c0 + e % TW;
            stage[wave][e / TW][e % TW] = *A.at(b, k0 + (rr < rm1 ? rr : rm1), k0 + (cc < wm1 ? cc : wm1));
          }
          __syncthreads();
  #pragma unroll
          for (int c = 0; c < TW; ++c) {
            const T v = stage[wave][lane][c];
            x[sl][c0 + c] = ((lp[sl] < R) & (c0 + c < w)) ? v : T(0);
          }
          __syncthreads();
        });
      });
    }
    int zc = 0;    // panel column (1-based) of the first exactly zero pivot
    int pivl = 0;  // thread j: the 1-based pivot row of column j
    // one column.  Columns past w (a narrower last panel) hold zeros: their
    // pivot is their own row, nothing is updated, and they record no zero
    // pivot and no ipiv entry, so panels up to 16 wide (and FP64 ones) run
    // them as straight-line code with no branch but the owner's candidate
    // row write; wider FP32 panels skip them (the straight line of 32
    // columns would not fit the registers).
    constexpr bool kStraight = sizeof(T) == 8 || W <= 16;
    auto column = [&](auto jc) __attribute__((always_inline)) {
      constexpr int j = decltype(jc)::value;
      if (kStraight || j < w) {
        const int par = j & 1;
        unsigned long long best = 0;
#pragma unroll
        for (int s = 0; s < RP; ++s) {
          const T v = x[s][j];
          const bool ok = (lp[s] >= j) & (lp[s] < Rp) & (v == v);
          best = umax64(best, ok ? pivot_key(v, lp[s]) : 0ull);
        }
        best = wave_max_u64(best);
        if (NW > 1 && lane == 0) red_k[par][wave] = best;
        const int bi = pivot_key_index(best, T(0));
        const bool any = best != 0ull;
        // the wavefront's best candidate row (logical position bi >= j) to LDS
#pragma unroll
        for (int s = 0; s < RP; ++s) {
          if (any & (lp[s] == bi)) {
#pragma unroll
            for (int c = 0; c < W; ++c) cand[par][wave][c] = x[s][c];
          }
        }
        __syncthreads();
        bool valid;
        int p, wv;
        if constexpr (NW > 1) {
          // all wavefront keys first (one round of LDS reads), then a tree max
          unsigned long long mk[NW];
          int mw[NW];
#pragma unroll
          for (int q = 0; q < NW; ++q) {
            mk[q] = red_k[par][q];
            mw[q] = q;
          }
#pragma unroll
          for (int st = 1; st < NW; st *= 2)
#pragma unroll
            for (int q = 0; q + st < NW; q += 2 * st) {
              const bool take = mk[q + st] > mk[q];
              mk[q] = take ? mk[q + st] : mk[q];
              mw[q] = take ? mw[q + st] : mw[q];
            }
          valid = mk[0] != 0ull;
          p = valid ? pivot_key_index(mk[0], T(0)) : j;
          wv = mw[0];
        } else {
          valid = any;
          p = valid ? bi : j;
          wv = 0;
        }
        pivl = (tid == j) ? k0 + p + 1 : pivl;
        T pr[W];
#pragma unroll
        for (int c = 0; c < W; ++c) pr[c] = cand[par][wv][c];
        // interchange j <-> p: only the two logical positions change (pivot
        // rows sit below j, padding and never-chosen rows at or past Rp)
#pragma unroll
        for (int s = 0; s < RP; ++s) {
          const int o = lp[s];
          lp[s] = (o == p) ? j : ((o == j) ? p : o);
        }
        const T pv = valid ? pr[j] : T(NAN);
        const bool nz = pv != T(0);
        zc = (!nz & (zc == 0) & (j < w)) ? j + 1 : zc;
        const T dv = nz ? pv : T(1);
        // FP64: one reciprocal per column (as LAPACK's getf2 scales by
        // 1/pivot); FP32: true division, which refinement depends on
        const T rv = sizeof(T) == 8 ? T(1) / dv : T(0);
#pragma unroll
        for (int s = 0; s < RP; ++s) {
          const bool m = (lp[s] > j) & (lp[s] < R);
          const T l = sizeof(T) == 8 ? x[s][j] * rv : x[s][j] / dv;
          const T lm = m ? l : T(0);
          x[s][j] = m ? l : x[s][j];
#pragma unroll
          for (int c = j + 1; c < W; ++c) x[s][c] -= lm * pr[c];
        }
      }
    };
    static_for<0, W>(column);
    if (tid < w) ipiv[(int64_t)b * pbs + k0 + tid] = pivl;
    if ((zc != 0) & (tid == 0) & (info != nullptr))
      if (info[b] == 0) info[b] = k0 + zc;
    if (vst) {
      // each row straight to its final logical position
#pragma unroll
      for (int s = 0; s < RP; ++s) {
        const int r = tid + NT * s;
        if (r < R) {
          T* dst = A.at(b, k0 + lp[s], k0);
          if (w == W) {
#pragma unroll
            for (int q = 0; q < W / VN; ++q) {
              vt v;
#pragma unroll
              for (int e = 0; e < VN; ++e) v[e] = x[s][q * VN + e];
              reinterpret_cast<vt*>(dst)[q] = v;
            }
          } else {
#pragma unroll
            for (int c = 0; c < W; ++c)
              if (c < w) dst[c] = x[s][c];
          }
        }
      }
    } else {
      // coalesced store through the same tiles, each row to its final position
      static_for<0, RP>([&](auto sc) __attribute__((always_inline)) {
        constexpr int sl = decltype(sc)::value;
        const int r0 = NT * sl + 64 * wave;
        static_for<0, W / TW>([&](auto tc) __attribute__((always_inline)) {
          constexpr int c0 = decltype(tc)::value * TW;
          __syncthreads();
  #pragma unroll
          for (int c = 0; c < TW; ++c) stage[wave][lane][c] = x[sl][c0 + c];
          stage_pos[wave][lane] = lp[sl];
          __syncthreads();
  #pragma unroll
          for (int q = 0; q < TW; ++q) {
            const int e = 64 * q + lane;
            const int rr = r0 + e / TW, cc = c0 + e % TW;
            if (rr < R && cc < w) *A.at(b, k0 + stage_pos[wave][e / TW], k0 + cc) = stage[wave][e / TW][e % TW];
          }
        });
      });
    }
    __syncthreads();
  }
}

// Row interchanges of one panel (rows k0..k0+w-1, 1-based pivots
// ipiv[k0..k0+w-1] >= row, w <= W <= 32) on every column outside the panel,
// and U12 = L11^{-1} A12 on the columns right of it; one thread per column
// (one wavefront per workgroup).  The rows a panel touches are its own w
// rows plus the distinct pivot rows below it: each gets a slot, the swap
// sequence is replayed on slot ids (every lane tracks where its slot ends
// up), and each column then moves its touched entries with independent
// loads and stores, solving the unit-lower L11 (LDS) on the block rows in
// registers.
template <typename T, int W, int NT>
__global__ __launch_bounds__(NT) void k_swap_trsm(int batch, int ncols, int k0, int w, View<T> A, const int* ipiv,
                                                  int64_t pbs) {
  // the w block rows and up to w pivot rows get one slot per lane of
  // wavefront 0; every thread owns one column
  static_assert(W <= 32, "2 W slots must fit the 64 lanes");
  static_assert(NT % 64 == 0 && NT >= W * W / 32, "swap-trsm shape");
  __shared__ T L[W][W + 1];
  __shared__ int s_piv[W];
  __shared__ int s_slot[W];
  __shared__ int s_row[2 * W];
  __shared__ int s_src[2 * W];
  __shared__ int s_cnt;
  const int tid = threadIdx.x;
  const int col = blockIdx.x * NT + tid;
  const int cl = col < ncols ? col : ncols - 1;
  const int wm1 = w - 1;
  const bool right = col >= k0 + w;
  for (int b = blockIdx.y; b < batch; b += gridDim.y) {
    if (tid < W) s_piv[tid] = (tid < w) ? ipiv[(int64_t)b * pbs + k0 + tid] - 1 : -1;
    if constexpr (NT == 64) {
      // thread r < W loads row r of the block
      if (tid < W) {
        const int r = tid;
#pragma unroll
        for (int c = 0; c < W; ++c) {
          const T v = *A.at(b, k0 + (r < wm1 ? r : wm1), k0 + (c < wm1 ? c : wm1));
          L[r][c] = ((r < w) & (c < r)) ? v : T(0);
        }
      }
    } else {
      for (int e = tid; e < W * W; e += NT) {
        const int r = e / W, c = e % W;
        const T v = *A.at(b, k0 + (r < wm1 ? r : wm1), k0 + (c < wm1 ? c : wm1));
        L[r][c] = ((r < w) & (c < r)) ? v : T(0);
      }
    }
    __syncthreads();
    const int i = tid;
    int p = -1, src_lane = i;
    bool first = false;
    if (tid < 64) {
      p = (i < w) ? s_piv[i] : -1;
      first = (i < w) && (p >= k0 + w);
      for (int q = 0; q < W; ++q) {
        const bool hit = first & (q < i) & (s_piv[q] == p);
        src_lane = (hit && src_lane == i) ? q : src_lane;
      }
      first = first & (src_lane == i);
      const unsigned long long mask = __ballot(first);
      const int rank = __popcll(mask & ((1ull << i) - 1ull));
      if (first) {
        s_slot[i] = w + rank;
        s_row[w + rank] = p;
      }
      if (i < w) s_row[i] = k0 + i;
      if (i == 0) s_cnt = w + __popcll(mask);
    }
    __syncthreads();
    if (i < w) {
      if (p < k0 + w) s_slot[i] = p - k0;
      else if (!first) s_slot[i] = s_slot[src_lane];
    }
    __syncthreads();
    const int cnt = s_cnt;
    if (i < cnt) {
      int pos = i;
      for (int q = 0; q < w; ++q) {
        const int c = s_slot[q];
        pos = (pos == q) ? c : ((pos == c) ? q : pos);
      }
      s_src[pos] = i;
    }
    __syncthreads();
    T u[W], d[W];
#pragma unroll
    for (int q = 0; q < W; ++q) u[q] = *A.at(b, s_row[s_src[q < wm1 ? q : wm1]], cl);
#pragma unroll
    for (int t = 0; t < W; ++t) {
      const int q = w + t < cnt - 1 ? w + t : cnt - 1;
      d[t] = *A.at(b, s_row[s_src[q]], cl);
    }
    if (right) {
      // large batches: one column of L at a time (the compiler would
      // otherwise load the whole block ahead into registers and halve the
      // occupancy); the small-batch launches are latency-bound and keep the
      // loads ahead
#pragma unroll
      for (int a = 0; a < W; ++a) {
#pragma unroll
        for (int r = a + 1; r < W; ++r) u[r] -= L[r][a] * u[a];
        if constexpr (NT > 64) asm volatile("" ::: "memory");
      }
    }
    if (col < ncols && (col < k0 || right)) {
#pragma unroll
      for (int q = 0; q < W; ++q)
        if (q < w) *A.at(b, s_row[q], col) = u[q];
#pragma unroll
      for (int t = 0; t < W; ++t)
        if (w + t < cnt) *A.at(b, s_row[w + t], col) = d[t];
    }
    __syncthreads();
  }
}

static const size_t kPanelLdsBudget = 128 * 1024;

// Register panels, RP rows per thread in NT-thread workgroups (up to 2048
// rows).  A panel: its first column, width and kernel shape.
struct Panel {
  int k0, w, W, RP, NT;
};

static inline bool al16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// rows of A (columns from k0) as 16-byte vectors: aligned, row-major
template <typename T>
static bool rows_aligned(const View<T>& A, int batch, int k0) {
  return A.cs == 1 && al16(A.at(0, 0, k0)) && ((A.rs * (int64_t)sizeof(T)) & 15) == 0 &&
         (batch == 1 || ((A.bs * (int64_t)sizeof(T)) & 15) == 0);
}

template <typename T, int W, int RP, int NT>
static void panel_kernel(int batch, int nrows, int ncols, int k0, int w, int npiv, View<T> A, int* ipiv, int64_t pbs,
                         int* info, hipStream_t s) {
  // rows of the panel as 16-byte vectors: aligned rows, W columns in
  // bounds (columns past w are read but masked)
  const bool vst = rows_aligned(A, batch, k0), vld = vst && k0 + W <= ncols;
  hipLaunchKernelGGL((k_getf2_reg<T, W, RP, NT>), dim3(grid_batch(batch)), dim3(NT), 0, s, batch, nrows, k0, w, npiv,
                     A, ipiv, pbs, info, vld, vst);
}

// the panel kernel of shape (W, RP, NT)
template <typename T>
static hipError_t launch_panel(const Panel& P, int batch, int nrows, int ncols, int npiv, View<T> A, int* ipiv,
                               int64_t pbs, int* info, hipStream_t s) {
#define SLAB_PANEL(W_, RP_, NT_)                                                                                   \
  if (P.W == W_ && P.RP == RP_ && P.NT == NT_) {                                                                  \
    panel_kernel<T, W_, RP_, NT_>(batch, nrows, ncols, P.k0, P.w, npiv, A, ipiv, pbs, info, s);                \
    return hipGetLastError();                                                                                     \
  }
  if constexpr (sizeof(T) == 4) {
    SLAB_PANEL(32, 4, 64)
    SLAB_PANEL(32, 4, 128)
    SLAB_PANEL(32, 4, 256)
  } else {
    SLAB_PANEL(32, 2, 128)
    SLAB_PANEL(32, 2, 256)
    SLAB_PANEL(16, 4, 256)
    SLAB_PANEL(32, 1, 256)
    SLAB_PANEL(32, 1, 512)
    SLAB_PANEL(16, 2, 512)
  }
  SLAB_PANEL(8, 4, 512)
#undef SLAB_PANEL
  return hipErrorInvalidValue;
}

template <typename T>
static hipError_t launch_swap_trsm(int W, int batch, int ncols, int k0, int w, View<T> A, const int* ipiv,
                                   int64_t pbs, hipStream_t s) {
  // large batches (the leaf LUs): 256 columns per workgroup share the
  // block and the slot bookkeeping; small ones (the interface): 64, for
  // more workgroups per matrix
  if (batch > 64 && W == 32) {
    const dim3 g(ceil_div(ncols, 256), grid_batch(batch));
    hipLaunchKernelGGL((k_swap_trsm<T, 32, 256>), g, dim3(256), 0, s, batch, ncols, k0, w, A, ipiv, pbs);
    return hipGetLastError();
  }
  const dim3 g(ceil_div(ncols, 64), grid_batch(batch));
  if (W == 32)
    hipLaunchKernelGGL((k_swap_trsm<T, 32, 64>), g, dim3(64), 0, s, batch, ncols, k0, w, A, ipiv, pbs);
  else if (W == 16)
    hipLaunchKernelGGL((k_swap_trsm<T, 16, 64>), g, dim3(64), 0, s, batch, ncols, k0, w, A, ipiv, pbs);
  else
    hipLaunchKernelGGL((k_swap_trsm<T, 8, 64>), g, dim3(64), 0, s, batch, ncols, k0, w, A, ipiv, pbs);
  return hipGetLastError();
}

// The panels of an LU of nrows <= 2048 rows.  The panel shape follows the
// rows still below the diagonal: once they fit the next smaller shape, its
// panels take over (an FP64 1000-row LU runs 31 16-column panels, then 16
// 32-column panels).  FP32 and small FP64 batches use few wavefronts with
// four (two) rows per thread, which shortens each column's chain; large
// FP64 batches keep lighter wavefronts (more matrices per CU).
template <typename T>
static int plan_panels(int batch, int nrows, int nelim, Panel* out) {
  const bool few = batch <= 64;
  int n = 0, k0 = 0;
  while (k0 < nelim) {
    const int R = nrows - k0;
    int W, RP, NT, cap;
    if (R <= 256) {
      cap = 0;
      if (sizeof(T) == 4) W = 32, RP = 4, NT = 64;
      else if (few) W = 32, RP = 2, NT = 128;
      else W = 32, RP = 1, NT = 256;
    } else if (R <= 512) {
      cap = 256;
      if (sizeof(T) == 4) W = 32, RP = 4, NT = 128;
      else if (few) W = 32, RP = 2, NT = 256;
      else W = 32, RP = 1, NT = 512;
    } else if (R <= 1024) {
      cap = 512;
      if (sizeof(T) == 4) W = 32, RP = 4, NT = 256;
      else if (few) W = 16, RP = 4, NT = 256;
      else W = 16, RP = 2, NT = 512;
    } else {
      cap = 1024, W = 8, RP = 4, NT = 512;
    }
    // this shape's panels until the remaining rows fit `cap`
    const int need = R - cap;
    int k1 = cap == 0 ? nelim : k0 + ((need + W - 1) / W) * W;
    k1 = k1 < nelim ? k1 : nelim;
    for (; k0 < k1; k0 += W) out[n++] = Panel{k0, (k1 - k0) < W ? (k1 - k0) : W, W, RP, NT};
  }
  return n;
}

// right-looking: panel, interchanges + U12 on the other columns, trailing
// product, one after the other on one stream
template <typename T>
static hipError_t getrf_panels(int batch, int nrows, int ncols, int npiv, View<T> A, int* ipiv, int64_t pbs,
                               int* info, const Panel* P, int np, hipStream_t s) {
  hipError_t e;
  for (int i = 0; i < np; ++i) {
    const int k0 = P[i].k0, w = P[i].w;
    if ((e = launch_panel<T>(P[i], batch, nrows, ncols, npiv, A, ipiv, pbs, info, s)) != hipSuccess) return e;
    if ((e = launch_swap_trsm<T>(P[i].W, batch, ncols, k0, w, A, ipiv, pbs, s)) != hipSuccess) return e;
    const int mr = nrows - k0 - w, nr = ncols - k0 - w;
    if (mr > 0 && nr > 0) {
      e = gemm<T>(batch, mr, nr, w, T(-1), A.sub(k0 + w, k0), A.sub(k0, k0 + w), T(1), A.sub(k0 + w, k0 + w), s);
      if (e != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

template <typename T>
hipError_t getrf(int batch, int nrows, int ncols, int nelim, int npiv, View<T> A, int* ipiv, int64_t pbs, int* info,
                 hipStream_t s) {
  if (batch <= 0 || nelim <= 0) return hipSuccess;
  if (nelim > npiv || npiv > nrows || nelim > ncols) return hipErrorInvalidValue;
  if (nrows <= 2048) {
    Panel P[2048];
    const int np = plan_panels<T>(batch, nrows, nelim, P);
    return getrf_panels<T>(batch, nrows, ncols, npiv, A, ipiv, pbs, info, P, np, s);
  }
  // taller panels: LDS-resident panel, generic interchanges
  int nb = 32;
  while (nb > 4 && (size_t)nrows * nb * sizeof(T) > kPanelLdsBudget) nb >>= 1;
  if ((size_t)nrows * nb * sizeof(T) > kPanelLdsBudget) return hipErrorInvalidValue;
  static bool attr_set[2] = {false, false};
  const int ti = sizeof(T) == 8 ? 0 : 1;
  hipError_t e;
  if (!attr_set[ti]) {
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_getf2_panel<T>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPanelLdsBudget);
    if (e != hipSuccess) return e;
    attr_set[ti] = true;
  }
  for (int k0 = 0; k0 < nelim; k0 += nb) {
    const int w = (nelim - k0) < nb ? (nelim - k0) : nb;
    const size_t lds = (size_t)(nrows - k0) * w * sizeof(T);
    hipLaunchKernelGGL(k_getf2_panel<T>, dim3(grid_batch(batch)), dim3(256), lds, s, batch, nrows, k0, w, npiv, A,
                       ipiv, pbs, info);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // interchanges on every column outside the panel (the panel kernel
    // already swapped its own columns), then U12 = L11^{-1} A12
    dim3 g2(ceil_div(ncols, 256), grid_batch(batch));
    hipLaunchKernelGGL(k_laswp_panel<T>, g2, dim3(256), 0, s, batch, ncols, k0, w, A, ipiv, pbs, k0, k0 + w);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int mr = nrows - k0 - w, nr = ncols - k0 - w;
    if (nr > 0) {
      dim3 g3(ceil_div(nr, 64), grid_batch(batch));
      hipLaunchKernelGGL(k_trsm_diag<T>, g3, dim3(64), 0, s, batch, nr, 0, w, 1, 1, A.sub(k0, k0), A.sub(k0, k0 + w));
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (mr > 0 && nr > 0) {
      e = gemm<T>(batch, mr, nr, w, T(-1), A.sub(k0 + w, k0), A.sub(k0, k0 + w), T(1), A.sub(k0 + w, k0 + w), s);
      if (e != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

// Sum of v over the 256-thread workgroup, returned to every thread (red:
// 4 entries of LDS; the caller's next write to red follows a barrier).
template <typename T>
__device__ inline T block_sum256(T v, T* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const T t = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return t;
}

// Orthonormal columns spanning the columns of each m x k matrix A (m >= k,
// k <= 256), by Householder QR: one workgroup per matrix; per column the
// norm below the diagonal (workgroup sum), the reflector stored below the
// diagonal of A (LAPACK's v(0) = 1 scaling, tau in LDS), then thread c
// applies it to column c > j.  Q = H_0 ... H_{k-1} [I_k; 0] is then
// accumulated backwards into Q, one thread per column.  A is overwritten.
// A rank-deficient A still gets k orthonormal columns (a zero column leaves
// its reflector the identity, so e_j enters Q).
template <typename T>
__global__ __launch_bounds__(256) void k_house_orth(int batch, int m, int k, View<T> A, View<T> Q) {
  __shared__ T tau_s[256];
  __shared__ T red[4];
  const int tid = threadIdx.x;
  for (int b = blockIdx.x; b < batch; b += gridDim.x) {
    for (int j = 0; j < k; ++j) {
      T s = T(0);
      for (int i = j + 1 + tid; i < m; i += 256) {
        const T v = *A.at(b, i, j);
        s += v * v;
      }
      const T sigma = block_sum256(s, red);
      const T alpha = *A.at(b, j, j);
      T tau = T(0), beta = alpha, scale = T(1);
      if (sigma > T(0)) {
        beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
        tau = (beta - alpha) / beta;
        scale = T(1) / (alpha - beta);
      }
      for (int i = j + 1 + tid; i < m; i += 256) *A.at(b, i, j) *= scale;
      __syncthreads();
      if (tid == 0) {
        *A.at(b, j, j) = beta;
        tau_s[j] = tau;
      }
      const int c = tid;
      if (c > j && c < k && tau != T(0)) {
        T w = *A.at(b, j, c);
#pragma unroll 8
        for (int i = j + 1; i < m; ++i) w += *A.at(b, i, j) * *A.at(b, i, c);
        w *= tau;
        *A.at(b, j, c) -= w;
#pragma unroll 8
        for (int i = j + 1; i < m; ++i) *A.at(b, i, c) -= *A.at(b, i, j) * w;
      }
      __syncthreads();
    }
    for (int e = tid; e < m * k; e += 256) {
      const int r = e / k, c = e - r * k;
      *Q.at(b, r, c) = (r == c) ? T(1) : T(0);
    }
    __syncthreads();
    for (int j = k - 1; j >= 0; --j) {
      const T tau = tau_s[j];
      const int c = tid;
      // columns below j are e_c, which H_j leaves alone
      if (c >= j && c < k && tau != T(0)) {
        T w = *Q.at(b, j, c);
#pragma unroll 8
        for (int i = j + 1; i < m; ++i) w += *A.at(b, i, j) * *Q.at(b, i, c);
        w *= tau;
        *Q.at(b, j, c) -= w;
#pragma unroll 8
        for (int i = j + 1; i < m; ++i) *Q.at(b, i, c) -= *A.at(b, i, j) * w;
      }
      __syncthreads();
    }
  }
}

template <typename T>
hipError_t house_orth(int batch, int m, int k, View<T> A, View<T> Q, hipStream_t s) {
  if (batch <= 0 || k <= 0) return hipSuccess;
  if (m < k || k > 256) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_house_orth<T>, dim3(grid_batch(batch)), dim3(256), 0, s, batch, m, k, A, Q);
  return hipGetLastError();
}

template hipError_t house_orth<float>(int, int, int, View<float>, View<float>, hipStream_t);
template hipError_t house_orth<double>(int, int, int, View<double>, View<double>, hipStream_t);

template <typename T>
__global__ __launch_bounds__(256) void k_copy(int batch, int rows, int cols, View<T> src, View<T> dst) {
  const int64_t n = (int64_t)rows * cols;
  for (int b = blockIdx.y; b < batch; b += gridDim.y)
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
      const int r = (int)(e / cols), c = (int)(e - (int64_t)r * cols);
      *dst.at(b, r, c) = *src.at(b, r, c);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_identity(int batch, int n, View<T> dst) {
  const int64_t nn = (int64_t)n * n;
  for (int b = blockIdx.y; b < batch; b += gridDim.y)
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < nn; e += (int64_t)gridDim.x * 256) {
      const int r = (int)(e / n), c = (int)(e - (int64_t)r * n);
      *dst.at(b, r, c) = (r == c) ? T(1) : T(0);
    }
}

template <typename T>
hipError_t copy_batched(int batch, int rows, int cols, View<T> src, View<T> dst, hipStream_t s) {
  if (batch <= 0 || rows <= 0 || cols <= 0) return hipSuccess;
  const int64_t n = (int64_t)rows * cols;
  const int gx = (int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
  hipLaunchKernelGGL(k_copy<T>, dim3(gx, grid_batch(batch)), dim3(256), 0, s, batch, rows, cols, src, dst);
  return hipGetLastError();
}

template <typename T>
hipError_t set_identity(int batch, int n, View<T> dst, hipStream_t s) {
  if (batch <= 0 || n <= 0) return hipSuccess;
  const int64_t nn = (int64_t)n * n;
  const int gx = (int)((nn + 255) / 256 < 4096 ? (nn + 255) / 256 : 4096);
  hipLaunchKernelGGL(k_identity<T>, dim3(gx, grid_batch(batch)), dim3(256), 0, s, batch, n, dst);
  return hipGetLastError();
}

template <typename T>
hipError_t getrs(int batch, int n, int nrhs, bool trans, View<T> LU, const int* ipiv, int64_t pbs, View<T> B,
                 hipStream_t s) {
  hipError_t e;
  if (!trans && nrhs <= 8 && n <= 2048 && sizeof(T) * ((size_t)n * nrhs + 64 * 65) + sizeof(int) * n <= 160 * 1024)
    return getrs_small<T>(batch, n, nrhs, LU, ipiv, pbs, B, s);
  if (!trans) {
    if ((e = laswp<T>(batch, nrhs, 0, n, true, B, ipiv, pbs, s)) != hipSuccess) return e;
    if ((e = trsm_left<T>(batch, n, nrhs, true, true, LU, B, s)) != hipSuccess) return e;
    return trsm_left<T>(batch, n, nrhs, false, false, LU, B, s);
  }
  if ((e = trsm_left<T>(batch, n, nrhs, true, false, LU.t(), B, s)) != hipSuccess) return e;
  if ((e = trsm_left<T>(batch, n, nrhs, false, true, LU.t(), B, s)) != hipSuccess) return e;
  return laswp<T>(batch, nrhs, 0, n, false, B, ipiv, pbs, s);
}

// ---------------------------------------------------------------------------
// Inverse from getrf factors: A^{-1} = U^{-1} L^{-1} P.
//
// Both triangles are inverted as lower-triangular matrices (U through its
// transpose), in a workspace padded to np = 32 * 2^l >= n with the identity:
//   W[z]  z <  nb: LU of matrix z (its strict lower part is L, unit diagonal)
//         z >= nb: (LU of matrix z - nb)^T (its lower part is U^T)
//   X[z]  the inverse of that lower-triangular matrix.
// The 32 x 32 diagonal blocks are inverted by substitution in one launch;
// then recursive doubling joins neighbouring blocks, for s = 32, 64, ...:
//   X21 = -X22 (W21 X11)
// for every pair of s-blocks of every matrix in two batched GEMMs, so the
// whole inversion is 2 log2(np / 32) + 2 launches, without a sequential
// block loop.  Finally Y = X[nb + b]^T X[b] = U^{-1} L^{-1} and the column
// permutation A^{-1}[:, j] = Y[:, iperm[j]] (iperm[v] = the row where the
// interchanges send row v).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_getri_prep(int n, int np, int nb, View<T> LU, T* W) {
  __shared__ T t[32][33];
  const int b = blockIdx.z, r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t nn = (int64_t)np * np;
  T* W0 = W + b * nn;
  T* W1 = W + (nb + b) * nn;
  const int nm1 = n - 1;
#pragma unroll
  for (int i = ty; i < 32; i += 8) {
    const int r = r0 + i, c = c0 + tx;
    const T v = *LU.at(b, r < nm1 ? r : nm1, c < nm1 ? c : nm1);
    const T w = ((r < n) & (c < n)) ? v : T(0);
    t[i][tx] = w;
    W0[(int64_t)r * np + c] = w;
  }
  __syncthreads();
#pragma unroll
  for (int i = ty; i < 32; i += 8) W1[(int64_t)(c0 + i) * np + r0 + tx] = t[tx][i];
}

template <typename T>
__global__ __launch_bounds__(256) void k_getri_diag(int n, int np, int nb, const T* W, T* X) {
  __shared__ T Ls[32][33];
  const int z = blockIdx.y, d0 = blockIdx.x * 32, tid = threadIdx.x;
  const bool unit = z < nb;
  const int64_t nn = (int64_t)np * np;
  const T* A = W + z * nn;
  T* Xz = X + z * nn;
  for (int e = tid; e < 32 * 32; e += 256) {
    const int r = e >> 5, c = e & 31;
    const T v = A[(int64_t)(d0 + r) * np + d0 + c];
    Ls[r][c] = ((d0 + r < n) & (d0 + c < n)) ? v : ((r == c) ? T(1) : T(0));
  }
  __syncthreads();
  if (tid < 32) {
    // column c of the block inverse: forward substitution on e_c
    const int c = tid;
    T x[32];
#pragma unroll
    for (int i = 0; i < 32; ++i) {
      T v = (i == c) ? T(1) : T(0);
#pragma unroll
      for (int j = 0; j < i; ++j) v -= Ls[i][j] * x[j];
      x[i] = unit ? v : v / Ls[i][i];
    }
#pragma unroll
    for (int i = 0; i < 32; ++i) Xz[(int64_t)(d0 + i) * np + d0 + c] = x[i];
  }
  // the strip right of the diagonal block is zero in a lower-triangular inverse
  for (int r = 0; r < 32; ++r)
    for (int c = d0 + 32 + tid; c < np; c += 256) Xz[(int64_t)(d0 + r) * np + c] = T(0);
}

__global__ __launch_bounds__(1024) void k_piv_to_iperm(int n, const int* ipiv, int64_t pbs, int* iperm, int64_t qbs) {
  extern __shared__ int P[];
  const int b = blockIdx.x;
  for (int r = threadIdx.x; r < n; r += 1024) P[r] = ipiv[b * pbs + r] - 1;
  __syncthreads();
  for (int v = threadIdx.x; v < n; v += 1024) {
    int pos = v;
#pragma unroll 8
    for (int k = 0; k < n; ++k) {
      const int pk = P[k];
      pos = (pos == k) ? pk : ((pos == pk) ? k : pos);
    }
    iperm[b * qbs + v] = pos;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_getri_gather(int n, int np, const T* Y, const int* iperm, View<T> Xo) {
  const int b = blockIdx.y, i = blockIdx.x;
  const T* y = Y + (int64_t)b * np * np + (int64_t)i * np;
  const int* q = iperm + (int64_t)b * n;
  for (int j = threadIdx.x; j < n; j += 256) *Xo.at(b, i, j) = y[q[j]];
}

size_t getri_ws_bytes(int batch, int n, size_t elem) {
  int np = 32;
  while (np < n) np *= 2;
  const size_t nn = (size_t)np * np;
  // W and X (2 * batch each), the pair products (2 * batch quarter matrices), iperm
  return elem * (4 * (size_t)batch * nn + 2 * (size_t)batch * (nn / 4)) + sizeof(int) * (size_t)batch * n + 256;
}

template <typename T>
__global__ __launch_bounds__(256) void k_getri_split(int n, int np, int nb, const T* X, const int* iperm, View<T> Li,
                                                     View<T> Ui) {
  __shared__ T t[32][33];
  const int b = blockIdx.z, r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t nn = (int64_t)np * np;
  const T* XL = X + b * nn;
  const T* XU = X + (nb + b) * nn;
  const int* q = iperm + (int64_t)b * n;
  // Li = L^{-1} P: column j of it is column iperm[j] of L^{-1}
#pragma unroll
  for (int i = ty; i < 32; i += 8) {
    const int r = r0 + i, c = c0 + tx;
    if (r < n && c < n) *Li.at(b, r, c) = XL[(int64_t)r * np + q[c]];
  }
  // Ui = U^{-1} = (X[nb + b])^T through an LDS tile
#pragma unroll
  for (int i = ty; i < 32; i += 8) t[i][tx] = XU[(int64_t)(c0 + i) * np + r0 + tx];
  __syncthreads();
#pragma unroll
  for (int i = ty; i < 32; i += 8) {
    const int r = r0 + i, c = c0 + tx;
    if (r < n && c < n) *Ui.at(b, r, c) = t[tx][i];
  }
}

// The triangular inverses of getri in the workspace: X[z] = L^{-1} of matrix
// z (z < batch), X[batch + z] = (U^T)^{-1}; iperm of every matrix.  Returns
// the padded size np.
template <typename T>
static hipError_t getri_core(int batch, int n, View<T> LU, const int* ipiv, int64_t pbs, void* ws, hipStream_t s,
                             int* np_out) {
  int np = 32;
  while (np < n) np *= 2;
  *np_out = np;
  const int64_t nn = (int64_t)np * np;
  T* W = reinterpret_cast<T*>(ws);
  T* X = W + 2 * batch * nn;
  T* Tm = X + 2 * batch * nn;
  int* iperm = reinterpret_cast<int*>(Tm + 2 * batch * (nn / 4));
  const int z = 2 * batch, nt = np / 32;
  hipError_t e;
  hipLaunchKernelGGL(k_getri_prep<T>, dim3(nt, nt, batch), dim3(256), 0, s, n, np, batch, LU, W);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_getri_diag<T>, dim3(nt, z), dim3(256), 0, s, n, np, batch, (const T*)W, X);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_piv_to_iperm, dim3(batch), dim3(1024), sizeof(int) * n, s, n, ipiv, pbs, iperm,
                     (int64_t)n);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  for (int sb = 32; sb < np; sb *= 2) {
    // pairs of s-blocks: pair j has its first block at (2 j s, 2 j s)
    const int npairs = np / (2 * sb);
    const int64_t ps = 2LL * sb * (np + 1), ts = (int64_t)sb * sb;
    const View<T> W21{W + (int64_t)sb * np, np, 1, ps};
    const View<T> X11{X, np, 1, ps};
    const View<T> X21{X + (int64_t)sb * np, np, 1, ps};
    const View<T> X22{X + (int64_t)sb * np + sb, np, 1, ps};
    const View<T> Tv{Tm, sb, 1, ts};
    e = gemm_outer<T>(z, nn, nn, nn / 4, npairs, sb, sb, sb, T(1), W21, X11, T(0), Tv, s);
    if (e != hipSuccess) return e;
    e = gemm_outer<T>(z, nn, nn / 4, nn, npairs, sb, sb, sb, T(-1), X22, Tv, T(0), X21, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename T>
hipError_t getri(int batch, int n, View<T> LU, const int* ipiv, int64_t pbs, View<T> Xo, void* ws, hipStream_t s) {
  if (batch <= 0 || n <= 0) return hipSuccess;
  if (batch > 32767 || n > 16384) return hipErrorInvalidValue;
  int np;
  hipError_t e = getri_core<T>(batch, n, LU, ipiv, pbs, ws, s, &np);
  if (e != hipSuccess) return e;
  const int64_t nn = (int64_t)np * np;
  T* W = reinterpret_cast<T*>(ws);
  T* X = W + 2 * batch * nn;
  const int* iperm = reinterpret_cast<const int*>(X + 2 * batch * nn + 2 * batch * (nn / 4));
  // Y = U^{-1} L^{-1} into W[0..batch) (the copies of LU are no longer read)
  e = gemm<T>(batch, n, n, n, T(1), View<T>{X + batch * nn, 1, np, nn}, View<T>{X, np, 1, nn}, T(0),
              View<T>{W, np, 1, nn}, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_getri_gather<T>, dim3(n, batch), dim3(256), 0, s, n, np, (const T*)W, iperm, Xo);
  return hipGetLastError();
}

template <typename T>
hipError_t getri_factors(int batch, int n, View<T> LU, const int* ipiv, int64_t pbs, View<T> Li, View<T> Ui,
                         void* ws, hipStream_t s) {
  if (batch <= 0 || n <= 0) return hipSuccess;
  if (batch > 32767 || n > 16384) return hipErrorInvalidValue;
  int np;
  hipError_t e = getri_core<T>(batch, n, LU, ipiv, pbs, ws, s, &np);
  if (e != hipSuccess) return e;
  const int64_t nn = (int64_t)np * np;
  const T* X = reinterpret_cast<const T*>(ws) + 2 * batch * nn;
  const int* iperm = reinterpret_cast<const int*>(X + 2 * batch * nn + 2 * batch * (nn / 4));
  const int nt = (n + 31) / 32;
  hipLaunchKernelGGL(k_getri_split<T>, dim3(nt, nt, batch), dim3(256), 0, s, n, np, batch, X, iperm, Li, Ui);
  return hipGetLastError();
}

#define SLAB_INSTANTIATE(T)                                                                                        \
  template hipError_t gemm<T>(int, int, int, int, T, View<T>, View<T>, T, View<T>, hipStream_t);                  \
  template hipError_t trsm_left<T>(int, int, int, bool, bool, View<T>, View<T>, hipStream_t);                     \
  template hipError_t laswp<T>(int, int, int, int, bool, View<T>, const int*, int64_t, hipStream_t);              \
  template hipError_t getrf<T>(int, int, int, int, int, View<T>, int*, int64_t, int*, hipStream_t);               \
  template hipError_t getrs<T>(int, int, int, bool, View<T>, const int*, int64_t, View<T>, hipStream_t);         \
  template hipError_t copy_batched<T>(int, int, int, View<T>, View<T>, hipStream_t);                              \
  template hipError_t set_identity<T>(int, int, View<T>, hipStream_t);                                            \
  template hipError_t getri<T>(int, int, View<T>, const int*, int64_t, View<T>, void*, hipStream_t);             \
  template hipError_t getri_factors<T>(int, int, View<T>, const int*, int64_t, View<T>, View<T>, void*,          \
                                       hipStream_t);
SLAB_INSTANTIATE(double)
SLAB_INSTANTIATE(float)

}  // namespace slab
