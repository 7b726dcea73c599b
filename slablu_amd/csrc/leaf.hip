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

// xnodes (optional): node ids (i*p + j) of the columns past the interior
// ones, nxc of them, e.g. the 4(p-1) corner-inclusive exterior nodes in
// Jxreorder order; NULL: the 4(p-2) face nodes in [l, r, d, u] order.
template <typename T>
__global__ __launch_bounds__(256) void k_leaf_assemble(int nboxes, int p, int row_begin, int nrows, int rows_per_block,
                                                       const double* __restrict__ D, const double* __restrict__ Dsq,
                                                       const double* c11, const double* c22, const double* c12,
                                                       const double* c1, const double* c2, const double* c,
                                                       const double* __restrict__ Nxtot, T* M, int64_t ldm,
                                                       int64_t mbs, const int* __restrict__ xnodes, int nxc) {
  // separate multiply and add roundings, as the reference's
  // `Aloc += f_vals * D` accumulation does (no fused multiply-add)
#pragma clang fp contract(off)
  const int f = p - 2, ni = f * f, nx = xnodes ? nxc : 4 * f, ncols = ni + nx;
  const int b = blockIdx.x;
  if (b >= nboxes) return;
  const int r0 = row_begin + blockIdx.y * rows_per_block;
  const int p2 = p * p;
  T* Mb = M + (int64_t)b * mbs;
  // the nodes of this thread's columns, once for all rows of the block
  constexpr int CPT = 4;
  int jj1[CPT], jj2[CPT];
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int col = threadIdx.x + 256 * k;
    const int cc = col < ncols ? col : ncols - 1;
    if (xnodes && cc >= ni) {
      const int node = xnodes[cc - ni];
      jj1[k] = node / p;
      jj2[k] = node % p;
    } else {
      leaf_col_node(cc, ni, f, p, jj1[k], jj2[k]);
    }
  }
  const bool tabled = ncols <= 256 * CPT && blockDim.x == 256 && rows_per_block <= 64;
  // the coefficients of the block's rows, loaded once (one round trip
  // instead of one per row)
  __shared__ double cf[5][64];
  if (tabled && c12 == nullptr) {
    const int t = threadIdx.x;
    if (t < rows_per_block) {
      const int r = r0 + t;
      const int rc = r < ni ? r : ni - 1;
      const int64_t ci = (int64_t)b * p2 + (1 + rc / f) * p + 1 + rc % f;
      cf[0][t] = -c11[ci];
      cf[1][t] = -c22[ci];
      cf[2][t] = c1 ? c1[ci] : 0.0;
      cf[3][t] = c2 ? c2[ci] : 0.0;
      cf[4][t] = c ? c[ci] : 0.0;
    }
    __syncthreads();
  }
  for (int rr = 0; rr < rows_per_block; ++rr) {
    const int r = r0 + rr;
    if (r >= nrows) break;
    if (r < ni && tabled && c12 == nullptr) {
      // without the cross term a row couples only to its node's grid lines
      // (j1 == i1 or j2 == i2); every other entry is an exact zero, written
      // without evaluating the terms (same value as the full sum: +0)
      const int i1 = 1 + r / f, i2 = 1 + r % f;
      const double v11 = cf[0][rr], v22 = cf[1][rr], v1 = cf[2][rr], v2 = cf[3][rr], v0 = cf[4][rr];
#pragma unroll
      for (int k = 0; k < CPT; ++k) {
        const int col = threadIdx.x + 256 * k;
        if (col < ncols) {
          const int j1 = jj1[k], j2 = jj2[k];
          double val = 0.0;
          if (i1 == j1 || i2 == j2) {
            val += v11 * (i2 == j2 ? Dsq[i1 * p + j1] : 0.0);
            val += v22 * (i1 == j1 ? Dsq[i2 * p + j2] : 0.0);
            if (c1) val += v1 * (-(i2 == j2 ? D[i1 * p + j1] : 0.0));
            if (c2) val += v2 * (-(i1 == j1 ? D[i2 * p + j2] : 0.0));
            if (c) val += v0 * ((i1 == j1 && i2 == j2) ? 1.0 : 0.0);
          }
          Mb[(int64_t)r * ldm + col] = (T)val;
        }
      }
      continue;
    }
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
        if (xnodes && col >= ni) {
          const int node = xnodes[col - ni];
          j1 = node / p;
          j2 = node % p;
        } else {
          leaf_col_node(col, ni, f, p, j1, j2);
        }
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

constexpr int kAsmMaxP2 = 1024;  // p^2 of the LDS copies of D and D^2 (p <= 32)

// The interior rows without a cross term (c12 == NULL), as 16-byte vector
// stores: each thread owns VN consecutive columns of a row, ncols / VN
// threads a row, 256 / (ncols / VN) rows at a time; the 1-D Chebyshev
// tables D and D^2 are read from LDS copies.  The entries are those
// of k_leaf_assemble (same terms in the same order, FP64, one rounding to
// T).  Requires ncols % VN == 0, ncols / VN <= 256 and 16-byte aligned rows.
template <typename T>
__global__ __launch_bounds__(256) void k_leaf_assemble_vec(int nboxes, int p, int rows_per_block,
                                                           const double* __restrict__ D, const double* __restrict__ Dsq,
                                                           const double* c11, const double* c22, const double* c1,
                                                           const double* c2, const double* c, T* M, int64_t ldm,
                                                           int64_t mbs, const int* __restrict__ xnodes, int nxc) {
#pragma clang fp contract(off)
  constexpr int VN = 16 / sizeof(T);
  typedef T vt __attribute__((ext_vector_type(VN)));
  const int f = p - 2, ni = f * f, nx = xnodes ? nxc : 4 * f, ncols = ni + nx;
  const int tpr = ncols / VN, rpi = 256 / tpr;
  const int b = blockIdx.x;
  const int r0 = blockIdx.y * rows_per_block;
  const int p2 = p * p;
  const int t = threadIdx.x;
  const int q = t % tpr, sub = t / tpr;
  __shared__ double cf[5][64];
  __shared__ double sD[kAsmMaxP2], sDsq[kAsmMaxP2];
  if (t < rows_per_block) {
    const int r = r0 + t;
    const int rc = r < ni ? r : ni - 1;
    const int64_t ci = (int64_t)b * p2 + (1 + rc / f) * p + 1 + rc % f;
    cf[0][t] = -c11[ci];
    cf[1][t] = -c22[ci];
    cf[2][t] = c1 ? c1[ci] : 0.0;
    cf[3][t] = c2 ? c2[ci] : 0.0;
    cf[4][t] = c ? c[ci] : 0.0;
  }
  for (int e = t; e < p2; e += 256) {
    sD[e] = D[e];
    sDsq[e] = Dsq[e];
  }
  __syncthreads();
  if (sub >= rpi || b >= nboxes) return;
  int jj1[VN], jj2[VN];
#pragma unroll
  for (int e = 0; e < VN; ++e) {
    const int col = q * VN + e;
    if (xnodes && col >= ni) {
      const int node = xnodes[col - ni];
      jj1[e] = node / p;
      jj2[e] = node % p;
    } else {
      leaf_col_node(col, ni, f, p, jj1[e], jj2[e]);
    }
  }
  T* Mb = M + (int64_t)b * mbs;
  for (int rr = sub; rr < rows_per_block; rr += rpi) {
    const int r = r0 + rr;
    if (r >= ni) break;
    const int i1 = 1 + r / f, i2 = 1 + r % f;
    const double v11 = cf[0][rr], v22 = cf[1][rr], v1 = cf[2][rr], v2 = cf[3][rr], v0 = cf[4][rr];
    vt out;
#pragma unroll
    for (int e = 0; e < VN; ++e) {
      const int j1 = jj1[e], j2 = jj2[e];
      double val = 0.0;
      if (i1 == j1 || i2 == j2) {
        val += v11 * (i2 == j2 ? sDsq[i1 * p + j1] : 0.0);
        val += v22 * (i1 == j1 ? sDsq[i2 * p + j2] : 0.0);
        if (c1) val += v1 * (-(i2 == j2 ? sD[i1 * p + j1] : 0.0));
        if (c2) val += v2 * (-(i1 == j1 ? sD[i2 * p + j2] : 0.0));
        if (c) val += v0 * ((i1 == j1 && i2 == j2) ? 1.0 : 0.0);
      }
      out[e] = (T)val;
    }
    reinterpret_cast<vt*>(Mb + (int64_t)r * ldm)[q] = out;
  }
}

template <typename T>
hipError_t leaf_assemble(int nboxes, int p, int border, const double* D, const double* Dsq, const double* c11,
                         const double* c22, const double* c12, const double* c1, const double* c2, const double* c,
                         const double* Nxtot, T* M, int64_t ldm, int64_t mbs, hipStream_t s, const int* xnodes,
                         int nxc) {
  const int f = p - 2, ni = f * f, nx = 4 * f;
  const int nrows = ni + (border ? nx : 0);
  const int ncols = ni + (xnodes ? nxc : nx);
  const int rpb = 32;
  constexpr int VN = 16 / sizeof(T);
  const bool vec = c12 == nullptr && p * p <= kAsmMaxP2 && ncols % VN == 0 && ncols / VN <= 256 &&
                   (reinterpret_cast<uintptr_t>(M) & 15) == 0 && ((ldm * (int64_t)sizeof(T)) & 15) == 0 &&
                   (nboxes == 1 || ((mbs * (int64_t)sizeof(T)) & 15) == 0);
  int row_begin = 0;
  if (vec) {
    dim3 gv(nboxes, (ni + rpb - 1) / rpb);
    hipLaunchKernelGGL(k_leaf_assemble_vec<T>, gv, dim3(256), 0, s, nboxes, p, rpb, D, Dsq, c11, c22, c1, c2, c, M,
                       ldm, mbs, xnodes, nxc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nrows == ni) return e;
    row_begin = ni;
  }
  dim3 grid(nboxes, (nrows - row_begin + rpb - 1) / rpb);
  hipLaunchKernelGGL(k_leaf_assemble<T>, grid, dim3(256), 0, s, nboxes, p, row_begin, nrows, rpb, D, Dsq, c11, c22,
                     c12, c1, c2, c, Nxtot, M, ldm, mbs, xnodes, nxc);
  return hipGetLastError();
}

template hipError_t leaf_assemble<double>(int, int, int, const double*, const double*, const double*, const double*,
                                          const double*, const double*, const double*, const double*, const double*,
                                          double*, int64_t, int64_t, hipStream_t, const int*, int);
template hipError_t leaf_assemble<float>(int, int, int, const double*, const double*, const double*, const double*,
                                         const double*, const double*, const double*, const double*, const double*,
                                         float*, int64_t, int64_t, hipStream_t, const int*, int);

// FP64 residual of the un-condensed collocation system, matrix-free, one
// workgroup per leaf.  The leaf's p^2 node values are assembled in LDS from
// the interior values uc (at the Jc nodes) and the face values ub through
// the face-to-node interpolation Intmap (rows at the Jxreo nodes, as
// src_disc/hps_parallel_leaf_ops.py:84 assembles them); then
//   rI = -(A_loc u) at the interior nodes (same operator as k_leaf_assemble)
//   fl = N_x u, the outward normal derivatives at the 4(p-2) face nodes,
// with N_x in compressed rows of p entries (nx_idx, nx_val).  uu, when not
// NULL, receives all node values.  Element (b, row, r) of every array
// argument sits at base + b*bs + row*ld + r.
__global__ __launch_bounds__(256) void k_leaf_residual(
    int nboxes, int p, int nrhs, const double* __restrict__ uc, int64_t uc_ld, int64_t uc_bs,
    const double* __restrict__ ub, int64_t ub_ld, int64_t ub_bs, const double* __restrict__ Intmap,
    const int* __restrict__ Jc, const int* __restrict__ Jxreo, const double* __restrict__ D,
    const double* __restrict__ Dsq, const double* __restrict__ c11, const double* __restrict__ c22,
    const double* __restrict__ c12, const double* __restrict__ c1, const double* __restrict__ c2,
    const double* __restrict__ c, const int* __restrict__ nx_idx, const double* __restrict__ nx_val,
    double* __restrict__ rI, int64_t rI_ld, int64_t rI_bs, double* __restrict__ fl, int64_t fl_ld, int64_t fl_bs,
    double* __restrict__ uu, int64_t uu_ld, int64_t uu_bs) {
  extern __shared__ double smem[];
  const int f = p - 2, ni = f * f, nx = 4 * f, p2 = p * p, nxn = p2 - ni;
  double* s_D = smem;             // p*p
  double* s_Dsq = s_D + p2;       // p*p
  double* s_u = s_Dsq + p2;       // p*p node values
  double* s_ux = s_u + p2;        // p*p x-derivatives (c12 term)
  double* s_ub = s_ux + p2;       // nx face values
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  for (int t = tid; t < p2; t += blockDim.x) {
    s_D[t] = D[t];
    s_Dsq[t] = Dsq[t];
  }
  const int64_t cb = (int64_t)b * p2;
  for (int r = 0; r < nrhs; ++r) {
    __syncthreads();
    for (int t = tid; t < nx; t += blockDim.x) s_ub[t] = ub[(int64_t)b * ub_bs + (int64_t)t * ub_ld + r];
    __syncthreads();
    for (int t = tid; t < p2; t += blockDim.x) {
      int node;
      double v;
      if (t < ni) {
        node = Jc[t];
        v = uc[(int64_t)b * uc_bs + (int64_t)t * uc_ld + r];
      } else {
        const int q = t - ni;
        node = Jxreo[q];
        const double* row = Intmap + (int64_t)q * nx;
        double a0 = 0.0, a1 = 0.0;
        int k = 0;
        for (; k + 1 < nx; k += 2) {
          a0 += row[k] * s_ub[k];
          a1 += row[k + 1] * s_ub[k + 1];
        }
        if (k < nx) a0 += row[k] * s_ub[k];
        v = a0 + a1;
      }
      s_u[node] = v;
      if (uu) uu[(int64_t)b * uu_bs + (int64_t)node * uu_ld + r] = v;
    }
    __syncthreads();
    if (c12) {
      for (int t = tid; t < p2; t += blockDim.x) {
        const int i1 = t / p, i2 = t % p;
        double a = 0.0;
        for (int j = 0; j < p; ++j) a += s_D[i1 * p + j] * s_u[j * p + i2];
        s_ux[t] = a;
      }
      __syncthreads();
    }
    (void)nxn;
    for (int t = tid; t < ni + nx; t += blockDim.x) {
      if (t < ni) {
        const int i1 = 1 + t / f, i2 = 1 + t % f;
        const int node = i1 * p + i2;
        double axx = 0.0, ayy = 0.0, ax = 0.0, ay = 0.0;
        for (int j = 0; j < p; ++j) {
          const double ux = s_u[j * p + i2], uy = s_u[i1 * p + j];
          axx += s_Dsq[i1 * p + j] * ux;
          ayy += s_Dsq[i2 * p + j] * uy;
          ax += s_D[i1 * p + j] * ux;
          ay += s_D[i2 * p + j] * uy;
        }
        double res = -c11[cb + node] * axx;
        res -= c22[cb + node] * ayy;
        if (c12) {
          double axy = 0.0;
          for (int j = 0; j < p; ++j) axy += s_D[i2 * p + j] * s_ux[i1 * p + j];
          res -= 2.0 * c12[cb + node] * axy;
        }
        if (c1) res -= c1[cb + node] * ax;
        if (c2) res -= c2[cb + node] * ay;
        if (c) res += c[cb + node] * s_u[node];
        rI[(int64_t)b * rI_bs + (int64_t)t * rI_ld + r] = -res;
      } else {
        const int q = t - ni;
        const int* ix = nx_idx + q * p;
        const double* vx = nx_val + q * p;
        double a0 = 0.0, a1 = 0.0;
        int k = 0;
        for (; k + 1 < p; k += 2) {
          a0 += vx[k] * s_u[ix[k]];
          a1 += vx[k + 1] * s_u[ix[k + 1]];
        }
        if (k < p) a0 += vx[k] * s_u[ix[k]];
        fl[(int64_t)b * fl_bs + (int64_t)q * fl_ld + r] = a0 + a1;
      }
    }
  }
}

hipError_t leaf_residual(int nboxes, int p, int nrhs, const double* uc, int64_t uc_ld, int64_t uc_bs,
                         const double* ub, int64_t ub_ld, int64_t ub_bs, const double* Intmap, const int* Jc,
                         const int* Jxreo, const double* D, const double* Dsq, const double* c11, const double* c22,
                         const double* c12, const double* c1, const double* c2, const double* c, const int* nx_idx,
                         const double* nx_val, double* rI, int64_t rI_ld, int64_t rI_bs, double* fl, int64_t fl_ld,
                         int64_t fl_bs, double* uu, int64_t uu_ld, int64_t uu_bs, hipStream_t s) {
  if (nboxes == 0 || nrhs == 0) return hipSuccess;
  const int f = p - 2;
  const size_t lds = sizeof(double) * (4 * (size_t)p * p + 4 * f);
  hipLaunchKernelGGL(k_leaf_residual, dim3(nboxes), dim3(256), lds, s, nboxes, p, nrhs, uc, uc_ld, uc_bs, ub, ub_ld,
                     ub_bs, Intmap, Jc, Jxreo, D, Dsq, c11, c22, c12, c1, c2, c, nx_idx, nx_val, rI, rI_ld, rI_bs, fl,
                     fl_ld, fl_bs, uu, uu_ld, uu_bs);
  return hipGetLastError();
}

}  // namespace slab
