// Batched tall-skinny products of the FP64 Krylov refinement: for every
// right-hand-side column c at once, the k dot products of a vector against
// its basis (Gram-Schmidt projections, norms, Gram matrices) and the update
// of a vector by a combination of its basis.  The vectors are long (one
// entry per unknown of the un-condensed system: 1.1 M at n = 1000) and few
// (k <= 32), so both are HBM-bound streams: each reads the k basis vectors
// once, coalesced, with deterministic (fixed-order) reductions.
#include "slab_internal.h"

namespace slab {

namespace {

constexpr int kMdotMaxK = 32;
constexpr int kChunk = 256 * 8;  // vector entries per workgroup

// part[c][chunk][i] = sum over the chunk's t of V[c][i][t] w[c][t]
__global__ __launch_bounds__(256) void k_mdot_part(int k, int64_t n, const double* V, int64_t vcs, int64_t vks,
                                                   const double* w, int64_t wcs, double* part, int nchunk) {
  __shared__ double red[4][kMdotMaxK];
  const int c = blockIdx.y, ch = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double* Vc = V + (int64_t)c * vcs;
  const double* wc = w + (int64_t)c * wcs;
  double acc[kMdotMaxK];
#pragma unroll
  for (int i = 0; i < kMdotMaxK; ++i) acc[i] = 0.0;
  const int64_t t0 = (int64_t)ch * kChunk;
#pragma unroll 2
  for (int u = 0; u < 8; ++u) {
    const int64_t t = t0 + u * 256 + tid;
    if (t < n) {
      const double wv = wc[t];
#pragma unroll
      for (int i = 0; i < kMdotMaxK; ++i)
        if (i < k) acc[i] += Vc[(int64_t)i * vks + t] * wv;
    }
  }
#pragma unroll
  for (int i = 0; i < kMdotMaxK; ++i) {
    if (i < k) {
      double v = acc[i];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if (lane == 0) red[wave][i] = v;
    }
  }
  __syncthreads();
  if (tid < k) part[((int64_t)c * nchunk + ch) * k + tid] = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
}

// out[c][i] = sum over chunks of part[c][chunk][i], chunks in order
__global__ __launch_bounds__(256) void k_mdot_sum(int k, const double* part, int nchunk, double* out, int64_t ocs) {
  __shared__ double red[256];
  const int c = blockIdx.y, i = blockIdx.x, tid = threadIdx.x;
  double v = 0.0;
  for (int ch = tid; ch < nchunk; ch += 256) v += part[((int64_t)c * nchunk + ch) * k + i];
  red[tid] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  if (tid == 0) out[(int64_t)c * ocs + i] = red[0];
}

// w[c][t] += alpha sum_i V[c][i][t] coef[c][i]
__global__ __launch_bounds__(256) void k_maxpy(int k, int64_t n, double alpha, const double* V, int64_t vcs,
                                               int64_t vks, const double* coef, int64_t ccs, double* w,
                                               int64_t wcs) {
  __shared__ double cf[kMdotMaxK];
  const int c = blockIdx.y, tid = threadIdx.x;
  if (tid < k) cf[tid] = alpha * coef[(int64_t)c * ccs + tid];
  __syncthreads();
  const double* Vc = V + (int64_t)c * vcs;
  double* wc = w + (int64_t)c * wcs;
  const int64_t t0 = (int64_t)blockIdx.x * kChunk;
#pragma unroll 2
  for (int u = 0; u < 8; ++u) {
    const int64_t t = t0 + u * 256 + tid;
    if (t < n) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < kMdotMaxK; ++i)
        if (i < k) s += Vc[(int64_t)i * vks + t] * cf[i];
      wc[t] += s;
    }
  }
}

}  // namespace

size_t mdot_ws_elems(int nc, int k, int64_t n) {
  const int64_t nchunk = (n + kChunk - 1) / kChunk;
  return (size_t)nc * (size_t)nchunk * (size_t)(k > 0 ? k : 1);
}

hipError_t mdot(int nc, int k, int64_t n, const double* V, int64_t vcs, int64_t vks, const double* w, int64_t wcs,
                double* out, int64_t ocs, double* ws, hipStream_t s) {
  if (nc <= 0 || k <= 0) return hipSuccess;
  if (k > kMdotMaxK || nc > 65535) return hipErrorInvalidValue;
  const int nchunk = (int)((n + kChunk - 1) / kChunk);
  if (nchunk <= 0) {
    hipLaunchKernelGGL(k_mdot_sum, dim3(k, nc), dim3(256), 0, s, k, ws, 0, out, ocs);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_mdot_part, dim3(nchunk, nc), dim3(256), 0, s, k, n, V, vcs, vks, w, wcs, ws, nchunk);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mdot_sum, dim3(k, nc), dim3(256), 0, s, k, (const double*)ws, nchunk, out, ocs);
  return hipGetLastError();
}

hipError_t maxpy(int nc, int k, int64_t n, double alpha, const double* V, int64_t vcs, int64_t vks,
                 const double* coef, int64_t ccs, double* w, int64_t wcs, hipStream_t s) {
  if (nc <= 0 || k <= 0 || n <= 0) return hipSuccess;
  if (k > kMdotMaxK || nc > 65535) return hipErrorInvalidValue;
  const int nchunk = (int)((n + kChunk - 1) / kChunk);
  hipLaunchKernelGGL(k_maxpy, dim3(nchunk, nc), dim3(256), 0, s, k, n, alpha, V, vcs, vks, coef, ccs, w, wcs);
  return hipGetLastError();
}

}  // namespace slab
