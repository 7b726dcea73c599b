// Wavefront (64-lane) primitives for gfx950: lane broadcasts and an argmax
// all-reduce built on DPP row permutations plus four readlanes, so a pivot
// search costs a few VALU operations instead of LDS round trips.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>

namespace slab {

// Broadcast lane j's value of v to the wavefront (j uniform).
__device__ inline int rdlane(int v, int j) { return __builtin_amdgcn_readlane(v, j); }
__device__ inline float rdlane(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ inline double rdlane(double v, int j) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, j);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), j);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// v from the lane selected by the DPP control CTRL (all lanes valid sources)
template <int CTRL>
__device__ inline int dpp(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ inline float dpp(float v) {
  return __int_as_float(dpp<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ inline double dpp(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)dpp<CTRL>((int)(unsigned)u);
  const unsigned hi = (unsigned)dpp<CTRL>((int)(unsigned)(u >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// (v, i) <- the larger value, ties to the smaller index (LAPACK i?amax order)
template <typename T>
__device__ inline void argmax_merge(T& v, int& i, T ov, int oi) {
  const bool take = (ov > v) | ((ov == v) & (oi < i));
  v = take ? ov : v;
  i = take ? oi : i;
}

// All-reduce argmax over the wavefront: lanes pair up by quad permutations
// (xor 1, xor 2), then half-row and row mirrors (which pair the two quads of
// each 8 and the two 8s of each 16), then the four 16-lane rows are merged
// from their first lanes.  Every lane returns the same (v, i).
template <typename T>
__device__ inline void wave_argmax(T& v, int& i) {
  argmax_merge(v, i, dpp<0xB1>(v), dpp<0xB1>(i));    // quad_perm [1,0,3,2]
  argmax_merge(v, i, dpp<0x4E>(v), dpp<0x4E>(i));    // quad_perm [2,3,0,1]
  argmax_merge(v, i, dpp<0x141>(v), dpp<0x141>(i));  // row_half_mirror
  argmax_merge(v, i, dpp<0x140>(v), dpp<0x140>(i));  // row_mirror
  T v0 = rdlane(v, 0);
  int i0 = rdlane(i, 0);
  argmax_merge(v0, i0, rdlane(v, 16), rdlane(i, 16));
  argmax_merge(v0, i0, rdlane(v, 32), rdlane(i, 32));
  argmax_merge(v0, i0, rdlane(v, 48), rdlane(i, 48));
  v = v0;
  i = i0;
}

// Pivot search keys: one unsigned 64-bit number per candidate whose
// unsigned order is (|a| ascending, then row index descending), so a plain
// maximum picks the largest magnitude and, among equal magnitudes, the
// smallest index (LAPACK i?amax order).  FP32: |a| bits above the inverted
// 32-bit index (exact).  FP64: |a| bits with the low 12 mantissa bits
// replaced by the inverted index (indices below 2048; magnitudes that agree
// to 2^-40 relative count as equal).  0 means no candidate; the sign of a
// is ignored, a must not be NaN.
__device__ inline unsigned long long pivot_key(float a, int idx) {
  return ((unsigned long long)(__float_as_uint(a) & 0x7FFFFFFFu) << 32) |
         (unsigned long long)(0xFFFFFFFFu - (unsigned)idx);
}
__device__ inline unsigned long long pivot_key(double a, int idx) {
  return ((unsigned long long)__double_as_longlong(a) & 0x7FFFFFFFFFFFF000ull) | (unsigned long long)(0xFFF - idx);
}
__device__ inline int pivot_key_index(unsigned long long k, float) { return (int)(0xFFFFFFFFu - (unsigned)k); }
__device__ inline int pivot_key_index(unsigned long long k, double) { return 0xFFF - (int)(k & 0xFFFull); }

template <int CTRL, int ROWMASK = 0xF>
__device__ inline unsigned long long dpp_u64(unsigned long long v) {
  // lanes outside ROWMASK keep their own value
  const int lo = (int)(unsigned)v, hi = (int)(unsigned)(v >> 32);
  const unsigned rlo = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWMASK, 0xF, false);
  const unsigned rhi = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWMASK, 0xF, false);
  return ((unsigned long long)rhi << 32) | rlo;
}

__device__ inline unsigned long long umax64(unsigned long long a, unsigned long long b) { return a > b ? a : b; }

// All-reduce maximum of a 64-bit key over the wavefront with VALU-only
// steps: quad permutations and row mirrors reduce each 16-lane row, then
// row_bcast:15 / row_bcast:31 fold the rows into lane 63, which is read
// once.
__device__ inline unsigned long long wave_max_u64(unsigned long long k) {
  k = umax64(k, dpp_u64<0xB1>(k));        // quad_perm [1,0,3,2]
  k = umax64(k, dpp_u64<0x4E>(k));        // quad_perm [2,3,0,1]
  k = umax64(k, dpp_u64<0x141>(k));       // row_half_mirror
  k = umax64(k, dpp_u64<0x140>(k));       // row_mirror
  k = umax64(k, dpp_u64<0x142, 0xA>(k));  // row_bcast:15 into rows 1 and 3
  k = umax64(k, dpp_u64<0x143, 0xC>(k));  // row_bcast:31 into rows 2 and 3
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)k, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(k >> 32), 63);
  return ((unsigned long long)hi << 32) | lo;
}

// Bitwise blend: (a & ~m) | (b & m) on the bit patterns, m all ones or all
// zeros.  Picking one of several register-resident values this way keeps
// every index compile-time (a select chain would be folded into a load
// through a selected address, which forces the array into scratch).
__device__ inline float bit_blend(float a, float b, bool take) {
  const unsigned m = take ? 0xFFFFFFFFu : 0u;
  return __uint_as_float((__float_as_uint(a) & ~m) | (__float_as_uint(b) & m));
}
__device__ inline double bit_blend(double a, double b, bool take) {
  const unsigned long long m = take ? ~0ull : 0ull;
  const unsigned long long ua = (unsigned long long)__double_as_longlong(a);
  const unsigned long long ub = (unsigned long long)__double_as_longlong(b);
  return __longlong_as_double((long long)((ua & ~m) | (ub & m)));
}

// f(std::integral_constant<int, J>) for J = B..E-1: loop bodies whose index
// must be a compile-time constant (register-resident arrays), independent of
// the unroller's size limits.
template <int B, int E, typename F>
__device__ __attribute__((always_inline)) inline void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// MFMA wrappers: 16x16x4 tiles.  Lane l holds A[l%16][l/16] and
// B[l/16][l%16]; accumulator element q of lane l is D[row(l, q)][l%16] with
//   FP32 (v_mfma_f32_16x16x4f32): row = 4*(l/16) + q
//   FP64 (v_mfma_f64_16x16x4f64): row = 4*q + l/16
// (both measured on gfx950 with single-lane probes).
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  __device__ static inline acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int q) { return 4 * q + (lane >> 4); }
};
template <>
struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  __device__ static inline acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int q) { return 4 * (lane >> 4) + q; }
};

}  // namespace slab
