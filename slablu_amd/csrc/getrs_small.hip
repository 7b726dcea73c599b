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
            v[u] = (r < tb && c < tb) ? w : T(0);
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
            l[u] = (r < r1 && act) ? v : T(0);
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

template hipError_t getrs_small<double>(int, int, int, View<double>, const int*, int64_t, View<double>,
                                        hipStream_t);
template hipError_t getrs_small<float>(int, int, int, View<float>, const int*, int64_t, View<float>, hipStream_t);

}  // namespace slab
