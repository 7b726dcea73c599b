// Leaf operators of the HPS discretization on gfx950.
//
// k_leaf_assemble forms, for every leaf at once, the rows of the spectral
// collocation operator that the static condensation needs, directly from
// the 1-D Chebyshev matrices and the per-node PDE coefficients (the dense
// p^2 x p^2 operator is never materialised):
//   M = [ A_cc | A_cx ]           (n_i x (n_i + n_x))       border = 0
//   M = [ A_cc | A_cx ; N_c | N_x ]  ((n_i+n_x) x (n_i+n_x))  border = 1
// with n_i = (p-2)^2 interior nodes and n_x = 4(p-2) face nodes.
//
// Entry formula (node id i*p + j, i the x index, j the y index), summed in
// the order c11, c22, c12, c1, c2, c as in
// src_disc/hps_parallel_leaf_ops.py:10-37:
//   A[(i1,i2),(j1,j2)] = -c11 Dsq[i1,j1] d(i2,j2) - c22 d(i1,j1) Dsq[i2,j2]
//                        - 2 c12 D[i1,j1] D[i2,j2] - c1 D[i1,j1] d(i2,j2)
//                        - c2 d(i1,j1) D[i2,j2] + c d(row,col)
// where D is the scaled Chebyshev matrix on the descending grid
// (src_disc/hps_subdomain_disc.py:136-152) and coefficients are taken at
// the row node.  The arithmetic is done in FP64 and rounded once to T.
#include "slab_internal.h"

namespace slab {

__device__ inline void leaf_col_node(int col, int ni, int f, int p, int& j1, int& j2) {
  if (col < ni) {
    j1 = 1 + col / f;
    j2 = 1 + col % f;
  } else {
    const int q = col - ni, face = q / f, t = q % f;
    // [l, r, d, u] faces, each ascending in the tangential coordinate
    // (src_disc/hps_subdomain_disc.py:105-121)
    if (face == 0) { j1 = 0; j2 = 1 + t; }
    else if (face == 1) { j1 = p - 1; j2 = 1 + t; }
    else if (face == 2) { j1 = 1 + t; j2 = 0; }
    else { j1 = 1 + t; j2 = p - 1; }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_leaf_assemble(int nboxes, int p, int nrows, int rows_per_block,
                                                       const double* __restrict__ D, const double* __restrict__ Dsq,
                                                       const double* c11, const double* c22, const double* c12,
                                                       const double* c1, const double* c2, const double* c,
                                                       const double* __restrict__ Nxtot, T* M, int64_t ldm,
                                                       int64_t mbs) {
  // separate multiply and add roundings, as the reference's
  // `Aloc += f_vals * D` accumulation does (no fused multiply-add)
#pragma clang fp contract(off)
  const int f = p - 2, ni = f * f, nx = 4 * f, ncols = ni + nx;
  const int b = blockIdx.x;
  if (b >= nboxes) return;
  const int r0 = blockIdx.y * rows_per_block;
  const int p2 = p * p;
  T* Mb = M + (int64_t)b * mbs;
  for (int rr = 0; rr < rows_per_block; ++rr) {
    const int r = r0 + rr;
    if (r >= nrows) break;
    if (r < ni) {
      const int i1 = 1 + r / f, i2 = 1 + r % f;
      const int node = i1 * p + i2;
      const int64_t ci = (int64_t)b * p2 + node;
      const double v11 = -c11[ci];
      const double v22 = -c22[ci];
      const double v12 = c12 ? -2.0 * c12[ci] : 0.0;
      const double v1 = c1 ? c1[ci] : 0.0;
      const double v2 = c2 ? c2[ci] : 0.0;
      const double v0 = c ? c[ci] : 0.0;
      for (int col = threadIdx.x; col < ncols; col += blockDim.x) {
        int j1, j2;
        leaf_col_node(col, ni, f, p, j1, j2);
        double val = 0.0;
        val += v11 * (i2 == j2 ? Dsq[i1 * p + j1] : 0.0);
        val += v22 * (i1 == j1 ? Dsq[i2 * p + j2] : 0.0);
        if (c12) val += v12 * (D[i1 * p + j1] * D[i2 * p + j2]);
        if (c1) val += v1 * (-(i2 == j2 ? D[i1 * p + j1] : 0.0));
        if (c2) val += v2 * (-(i1 == j1 ? D[i2 * p + j2] : 0.0));
        if (c) val += v0 * ((i1 == j1 && i2 == j2) ? 1.0 : 0.0);
        Mb[(int64_t)r * ldm + col] = (T)val;
      }
    } else {
      const double* nrow = Nxtot + (int64_t)(r - ni) * ncols;
      for (int col = threadIdx.x; col < ncols; col += blockDim.x) Mb[(int64_t)r * ldm + col] = (T)nrow[col];
    }
  }
}

template <typename T>
hipError_t leaf_assemble(int nboxes, int p, int border, const double* D, const double* Dsq, const double* c11,
                         const double* c22, const double* c12, const double* c1, const double* c2, const double* c,
                         const double* Nxtot, T* M, int64_t ldm, int64_t mbs, hipStream_t s) {
  const int f = p - 2, ni = f * f, nx = 4 * f;
  const int nrows = ni + (border ? nx : 0);
  const int rpb = 8;
  dim3 grid(nboxes, (nrows + rpb - 1) / rpb);
  hipLaunchKernelGGL(k_leaf_assemble<T>, grid, dim3(256), 0, s, nboxes, p, nrows, rpb, D, Dsq, c11, c22, c12, c1,
                     c2, c, Nxtot, M, ldm, mbs);
  return hipGetLastError();
}

template hipError_t leaf_assemble<double>(int, int, int, const double*, const double*, const double*, const double*,
                                          const double*, const double*, const double*, const double*, const double*,
                                          double*, int64_t, int64_t, hipStream_t);
template hipError_t leaf_assemble<float>(int, int, int, const double*, const double*, const double*, const double*,
                                         const double*, const double*, const double*, const double*, const double*,
                                         float*, int64_t, int64_t, hipStream_t);

}  // namespace slab
