// Batched getrs for a few right-hand sides: the leaf interior solves with Acc
// (torch.linalg.solve in src_disc/hps_parallel_leaf_ops.py:86-100), here with
// the factors kept from the leaf factorization.
#include "lds_dev.h"

namespace slab {

namespace {

// Batched LU solve for a few right-hand sides with the factors of getrf
// (row interchanges, unit lower, upper), one workgroup per matrix: the
// right-hand side lives in LDS (gathered through the composed permutation
// as it is loaded), each 64-row diagonal block is staged in LDS and solved
// by substitution with lane = row (one wavefront per column), and the rows
// below/above are updated by wavefronts reading rows of LU coalesced, eight
// rows in flight per wavefront.
template <typename T>
__global__ __launch_bounds__(256) void k_getrs_small(int batch, int n, int nrhs, View<T> LU, const int* ipiv,
                                                     int64_t pbs, View<T> B) {
  extern __shared__ unsigned char smem_raw[];
  T* X = reinterpret_cast<T*>(smem_raw);  // n x nrhs
  T* Dg = X + n * nrhs;                   // 64 x 65
  int* P = reinterpret_cast<int*>(Dg + 64 * 65);
  int* Q = P + n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int b = blockIdx.x; b < batch; b += gridDim.x) {
    for (int r = tid; r < n; r += 256) P[r] = ipiv[(int64_t)b * pbs + r] - 1;
    __syncthreads();
    piv_to_perm(P, n, Q);
    for (int e = tid; e < n * nrhs; e += 256) {
      const int r = e / nrhs, c = e - r * nrhs;
      X[e] = *B.at(b, Q[r], c);
    }
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
      const bool lower = (pass == 0);
      const int nblk = (n + 63) / 64;
      for (int bi = 0; bi < nblk; ++bi) {
        const int blk = lower ? bi : nblk - 1 - bi;
        const int d0 = blk * 64;
        const int tb = (n - d0) < 64 ? (n - d0) : 64;
        for (int e0 = tid; e0 < 64 * 64; e0 += 256 * 8) {
          T v[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int e = e0 + 256 * u, r = e >> 6, c = e & 63;
            const T w = *LU.at(b, d0 + (r < tb ? r : tb - 1), d0 + (c < tb ? c : tb - 1));
            v[u] = ((r < tb) & (c < tb)) ? w : T(0);
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int e = e0 + 256 * u;
            Dg[(e >> 6) * 65 + (e & 63)] = v[u];
          }
        }
        __syncthreads();
        const bool act = lane < tb;
        const int rl = act ? lane : 0;
        for (int c = wave; c < nrhs; c += 4) {
          T x = X[(d0 + rl) * nrhs + c];
          if (lower) {
#pragma unroll 8
            for (int j = 0; j < tb; ++j) {
              const T a = Dg[rl * 65 + j];
              const T xj = rdlane(x, j);
              x = (lane > j) ? x - a * xj : x;
            }
          } else {
            const T rinv = T(1) / Dg[rl * 65 + rl];
#pragma unroll 8
            for (int j = tb - 1; j >= 0; --j) {
              const T a = Dg[rl * 65 + j];
              const T xj = rdlane(x, j) * rdlane(rinv, j);
              x = (lane == j) ? xj : ((lane < j) ? x - a * xj : x);
            }
          }
          if (act) X[(d0 + lane) * nrhs + c] = x;
        }
        __syncthreads();
        const int r0 = lower ? d0 + tb : 0;
        const int r1 = lower ? n : d0;
        for (int rb = r0 + 8 * wave; rb < r1; rb += 32) {
          T l[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int r = rb + u;
            const T v = *LU.at(b, r < r1 ? r : r1 - 1, d0 + (act ? lane : 0));
            l[u] = ((r < r1) & act) ? v : T(0);
          }
          for (int c = 0; c < nrhs; ++c) {
            const T xv = X[(d0 + rl) * nrhs + c];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              T v = l[u] * xv;
#pragma unroll
              for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
              if (lane == 0 && rb + u < r1) X[(rb + u) * nrhs + c] -= v;
            }
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

// Row segment row[0..63] of a factor (entries past tb read as zero) with
// 16-byte loads when the rows are aligned.
template <typename T>
__device__ inline void load_seg(const T* row, bool vec, int tb, T (&a)[64]) {
  constexpr int V = 16 / sizeof(T);
  if (vec && tb == 64) {
    typedef T vec_t __attribute__((ext_vector_type(V)));
    const vec_t* rv = reinterpret_cast<const vec_t*>(row);
#pragma unroll
    for (int q = 0; q < 64 / V; ++q) {
      const vec_t v = rv[q];
#pragma unroll
      for (int u = 0; u < V; ++u) a[q * V + u] = v[u];
    }
  } else {
    const int tm1 = tb - 1;
#pragma unroll
    for (int j = 0; j < 64; ++j) {
      const T v = row[j < tm1 ? j : tm1];
      a[j] = j < tb ? v : T(0);
    }
  }
}

// Batched LU solve with the composed row permutation (perm[r]: the row of B
// that P B puts at row r) for up to NC right-hand sides, one wavefront per
// matrix (four per workgroup) and no workgroup barriers.  Per 64-row
// diagonal block each lane holds its row of the block in registers and the
// block is solved by substitution with lane = row and readlane broadcasts;
// then each lane takes rows below/above the block, streaming their 64-entry
// segments of LU with independent 16-byte loads against the block's solved
// values (LDS broadcasts).  With one wavefront per matrix, thousands of
// matrices are in flight and their substitution and memory phases overlap.
template <typename T, int NC>
__global__ __launch_bounds__(256) void k_getrs_wave(int batch, int n, int nrhs, View<T> LU, const int* perm,
                                                    int64_t qbs, View<T> B) {
  extern __shared__ unsigned char smem_raw[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T* X = reinterpret_cast<T*>(smem_raw) + (size_t)wave * n * NC;  // n x NC
  const int b = blockIdx.x * (int)(blockDim.x >> 6) + wave;
  if (b >= batch) return;
  constexpr int V = 16 / sizeof(T);
  const T* base = LU.at(b, 0, 0);
  const bool vec = LU.cs == 1 && (LU.rs % V) == 0 && (reinterpret_cast<uintptr_t>(base) % 16) == 0;
  const int* pq = perm + (int64_t)b * qbs;
  const int cm1 = nrhs - 1;
  for (int r = lane; r < n; r += 64) {
    const int src = pq[r];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const T v = *B.at(b, src, c < cm1 ? c : cm1);
      X[r * NC + c] = c < nrhs ? v : T(0);
    }
  }
  __builtin_amdgcn_wave_barrier();
  const int nblk = (n + 63) >> 6;
  for (int q = 0; q < 2 * nblk; ++q) {
    const bool lower = q < nblk;
    const int d0 = (lower ? q : 2 * nblk - 1 - q) * 64;
    const int tb = (n - d0) < 64 ? (n - d0) : 64;
    const bool act = lane < tb;
    const int rl = d0 + (act ? lane : tb - 1);
    T a[64];
    load_seg<T>(LU.at(b, rl, d0), vec, tb, a);
    T x[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) x[c] = X[rl * NC + c];
    if (lower) {
      static_for<0, 64>([&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        if (j < tb) {
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const T xj = rdlane(x[c], j);
            x[c] = (lane > j) ? x[c] - a[j] * xj : x[c];
          }
        }
      });
    } else {
      const T rinv = T(1) / *LU.at(b, rl, rl);
      static_for<0, 64>([&](auto jc) __attribute__((always_inline)) {
        constexpr int j = 63 - decltype(jc)::value;
        if (j < tb) {
          const T rj = rdlane(rinv, j);
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const T xj = rdlane(x[c], j) * rj;
            x[c] = (lane == j) ? xj : ((lane < j) ? x[c] - a[j] * xj : x[c]);
          }
        }
      });
    }
    if (act) {
#pragma unroll
      for (int c = 0; c < NC; ++c) X[(d0 + lane) * NC + c] = x[c];
    }
    __builtin_amdgcn_wave_barrier();
    // rows outside the diagonal block: X[r] -= LU[r, d0:d0+tb] . X[d0:d0+tb].
    // Sixteen lanes share a row (four consecutive columns each), so a load
    // instruction reads four row segments contiguously; the lane's four
    // solution entries stay in registers and each row's sum is a DPP
    // reduction within its 16-lane row.  Sixteen (FP32) or eight (FP64) row
    // groups per iteration keep as many loads per lane in flight.
    const int r0 = lower ? d0 + tb : 0;
    const int r1 = lower ? n : d0;
    const int ro = lane >> 4, cc = (lane & 15) * 4;
    T xs[4][NC];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const T v = X[(d0 + (cc + u < tb ? cc + u : tb - 1)) * NC + c];
        xs[u][c] = cc + u < tb ? v : T(0);
      }
    const bool vseg = vec && (tb % 4) == 0;
    const int ccl = cc < tb ? cc : 0;  // column chunk inside the block (masked by xs)
    constexpr int U = sizeof(T) == 4 ? 16 : 8;
    const int rm1 = r1 - 1;
    for (int rb = r0; rb < r1; rb += 4 * U) {
      T l[U][4];
#pragma unroll
      for (int g = 0; g < U; ++g) {
        const int r = rb + 4 * g + ro;
        const T* rowp = LU.at(b, r < rm1 ? r : rm1, d0);
        if (vseg) {
          typedef T v4_t __attribute__((ext_vector_type(4)));
          const v4_t v = *reinterpret_cast<const v4_t*>(rowp + ccl);
#pragma unroll
          for (int u = 0; u < 4; ++u) l[g][u] = v[u];
        } else {
#pragma unroll
          for (int u = 0; u < 4; ++u) l[g][u] = rowp[cc + u < tb ? cc + u : tb - 1];
        }
      }
#pragma unroll
      for (int g = 0; g < U; ++g) {
        const int r = rb + 4 * g + ro;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          T acc = l[g][0] * xs[0][c];
#pragma unroll
          for (int u = 1; u < 4; ++u) acc += l[g][u] * xs[u][c];
          acc += dpp<0xB1>(acc);   // quad_perm [1,0,3,2]
          acc += dpp<0x4E>(acc);   // quad_perm [2,3,0,1]
          acc += dpp<0x141>(acc);  // row_half_mirror
          acc += dpp<0x140>(acc);  // row_mirror
          if ((lane & 15) == 0 && r < r1) X[r * NC + c] -= acc;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  for (int r = lane; r < n; r += 64)
#pragma unroll
    for (int c = 0; c < NC; ++c)
      if (c < nrhs) *B.at(b, r, c) = X[r * NC + c];
}

// perm[r] = the row that the interchanges ipiv (1-based, as getrf) bring to
// row r, one workgroup per matrix.
__global__ __launch_bounds__(256) void k_piv_to_perm(int batch, int n, const int* ipiv, int64_t pbs, int* perm,
                                                     int64_t qbs) {
  extern __shared__ unsigned char smem_raw[];
  int* P = reinterpret_cast<int*>(smem_raw);
  int* Q = P + n;
  for (int b = blockIdx.x; b < batch; b += gridDim.x) {
    for (int r = threadIdx.x; r < n; r += 256) P[r] = ipiv[(int64_t)b * pbs + r] - 1;
    __syncthreads();
    piv_to_perm(P, n, Q);
    for (int r = threadIdx.x; r < n; r += 256) perm[(int64_t)b * qbs + r] = Q[r];
    __syncthreads();
  }
}

}  // namespace

template <typename T>
hipError_t getrs_small(int batch, int n, int nrhs, View<T> LU, const int* ipiv, int64_t pbs, View<T> B,
                       hipStream_t s) {
  if (nrhs > 8) return hipErrorInvalidValue;
  const size_t lds = sizeof(T) * ((size_t)n * nrhs + 64 * 65) + 2 * sizeof(int) * n;
  if (lds > kLdsMax) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_getrs_small<T>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_getrs_small<T>, dim3(batch < 65535 ? batch : 65535), dim3(256), lds, s, batch, n, nrhs, LU,
                     ipiv, pbs, B);
  return hipGetLastError();
}

template <typename T>
hipError_t getrs_perm(int batch, int n, int nrhs, View<T> LU, const int* perm, int64_t qbs, View<T> B,
                      hipStream_t s) {
  if (nrhs > 8 || nrhs < 1 || n < 1) return hipErrorInvalidValue;
  const int nc = nrhs == 1 ? 1 : (nrhs <= 2 ? 2 : (nrhs <= 4 ? 4 : 8));
  // one wavefront (one matrix) per workgroup: the dispatcher spreads the
  // matrices over the CUs one at a time instead of four at a time
  const size_t lds = (size_t)n * nc * sizeof(T);
  if (lds > kLdsMax) return hipErrorInvalidValue;
  auto launch = [&](auto kern) -> hipError_t {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3(batch), dim3(64), lds, s, batch, n, nrhs, LU, perm, qbs, B);
    return hipGetLastError();
  };
  if (nc == 1) return launch(&k_getrs_wave<T, 1>);
  if (nc == 2) return launch(&k_getrs_wave<T, 2>);
  if (nc == 4) return launch(&k_getrs_wave<T, 4>);
  return launch(&k_getrs_wave<T, 8>);
}

hipError_t piv_to_perm_batched(int batch, int n, const int* ipiv, int64_t pbs, int* perm, int64_t qbs,
                               hipStream_t s) {
  if (batch <= 0 || n <= 0) return hipSuccess;
  const size_t lds = 2 * sizeof(int) * (size_t)n;
  if (lds > kLdsMax) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_piv_to_perm),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_piv_to_perm, dim3(batch < 65535 ? batch : 65535), dim3(256), lds, s, batch, n, ipiv, pbs,
                     perm, qbs);
  return hipGetLastError();
}

template hipError_t getrs_perm<double>(int, int, int, View<double>, const int*, int64_t, View<double>,
                                       hipStream_t);
template hipError_t getrs_perm<float>(int, int, int, View<float>, const int*, int64_t, View<float>, hipStream_t);
template hipError_t getrs_small<double>(int, int, int, View<double>, const int*, int64_t, View<double>,
                                        hipStream_t);
template hipError_t getrs_small<float>(int, int, int, View<float>, const int*, int64_t, View<float>, hipStream_t);

}  // namespace slab
