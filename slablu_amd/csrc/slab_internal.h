// Internal declarations shared by the HIP translation units of libslabhip.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace slab {

// A strided matrix view: element (i, j) of batch b lives at
//   base + b * bs + i * rs + j * cs.
// Transposition is a swap of (rs, cs); every kernel takes views, so op(A)
// never needs a copy.
template <typename T>
struct View {
  T* p;
  int64_t rs, cs, bs;
  __host__ __device__ T* at(int64_t b, int64_t i, int64_t j) const { return p + b * bs + i * rs + j * cs; }
  __host__ __device__ View<T> sub(int64_t i, int64_t j) const { return View<T>{p + i * rs + j * cs, rs, cs, bs}; }
  __host__ __device__ View<T> t() const { return View<T>{p, cs, rs, bs}; }
};

template <typename T>
inline View<T> view(void* p, int64_t rs, int64_t cs, int64_t bs) {
  return View<T>{reinterpret_cast<T*>(p), rs, cs, bs};
}

// C(b) = alpha * A(b) @ B(b) + beta * C(b);  A m x k, B k x n, C m x n.
template <typename T>
hipError_t gemm(int batch, int m, int n, int k, T alpha, View<T> A, View<T> B, T beta, View<T> C,
                hipStream_t s);

// B(b) <- op(A(b))^{-1} B(b) for a triangular m x m A; lower/upper refer to
// the view as given (pass A.t() for a transposed solve).
template <typename T>
hipError_t trsm_left(int batch, int m, int n, bool lower, bool unit, View<T> A, View<T> B, hipStream_t s);

// Row interchanges rows k1..k2-1 of an (nrows x ncols) view with the 1-based
// pivots ipiv[k] (forward: k ascending; backward: descending).
template <typename T>
hipError_t laswp(int batch, int ncols, int k1, int k2, bool forward, View<T> A, const int* ipiv, int64_t pbs,
                 hipStream_t s);

// Right-looking blocked LU with partial pivoting eliminating the first
// `nelim` columns of an nrows x ncols row-major matrix; pivot candidates are
// restricted to rows < npiv.  ipiv is 1-based (LAPACK).  info[b] receives the
// 1-based index of the first exactly-zero pivot (0 if none) when non-null.
template <typename T>
hipError_t getrf(int batch, int nrows, int ncols, int nelim, int npiv, View<T> A, int* ipiv, int64_t pbs,
                 int* info, hipStream_t s);

// Solve op(A) X = B with the factors of getrf (square n x n).
template <typename T>
hipError_t getrs(int batch, int n, int nrhs, bool trans, View<T> LU, const int* ipiv, int64_t pbs, View<T> B,
                 hipStream_t s);

}  // namespace slab

namespace slab {
// Launch probe (probe.hip): probe_on(kernel) is cheap; probe_begin returns
// a slot (or -1: off, full, or the stream is under capture) and records the
// start event, probe_end records the end event and the launch's algorithmic
// flops and bytes.
bool probe_on(int kernel);
int probe_begin(int kernel, hipStream_t s);
void probe_end(int slot, hipStream_t s, double flops, double bytes);
}  // namespace slab

namespace slab {
// Batched Krylov products (FP64): out[c*ocs + i] = V_c[i] . w_c for i < k
// (V_c[i] at V + c*vcs + i*vks, n entries each; ws holds mdot_ws_elems
// partial sums), and w_c += alpha sum_i coef[c*ccs + i] V_c[i].
size_t mdot_ws_elems(int nc, int k, int64_t n);
hipError_t mdot(int nc, int k, int64_t n, const double* V, int64_t vcs, int64_t vks, const double* w, int64_t wcs,
                double* out, int64_t ocs, double* ws, hipStream_t s);
hipError_t maxpy(int nc, int k, int64_t n, double alpha, const double* V, int64_t vcs, int64_t vks,
                 const double* coef, int64_t ccs, double* w, int64_t wcs, hipStream_t s);
hipError_t leaf_residual(int nboxes, int p, int nrhs, const double* uc, int64_t uc_ld, int64_t uc_bs,
                         const double* ub, int64_t ub_ld, int64_t ub_bs, const double* Intmap, const int* Jc,
                         const int* Jxreo, const double* D, const double* Dsq, const double* c11, const double* c22,
                         const double* c12, const double* c1, const double* c2, const double* c, const int* nx_idx,
                         const double* nx_val, double* rI, int64_t rI_ld, int64_t rI_bs, double* fl, int64_t fl_ld,
                         int64_t fl_bs, double* uu, int64_t uu_ld, int64_t uu_bs, hipStream_t s);
template <typename T>
hipError_t leaf_assemble(int nboxes, int p, int border, const double* D, const double* Dsq, const double* c11,
                         const double* c22, const double* c12, const double* c1, const double* c2, const double* c,
                         const double* Nxtot, T* M, int64_t ldm, int64_t mbs, hipStream_t s,
                         const int* xnodes = nullptr, int nxc = 0);
}  // namespace slab

namespace slab {
// dst(b) = src(b) for rows x cols views
template <typename T>
hipError_t copy_batched(int batch, int rows, int cols, View<T> src, View<T> dst, hipStream_t s);
// dst(b) = I (n x n)
template <typename T>
hipError_t set_identity(int batch, int n, View<T> dst, hipStream_t s);
// X(b) = A(b)^{-1} from the factors of getrf (n x n), out of place; ws holds
// orthonormal columns of each m x k block (Householder, A destroyed)
template <typename T>
hipError_t house_orth(int batch, int m, int k, View<T> A, View<T> Q, hipStream_t s);

// getri_ws_bytes(batch, n, sizeof(T)) bytes
size_t getri_ws_bytes(int batch, int n, size_t elem);
template <typename T>
hipError_t getri(int batch, int n, View<T> LU, const int* ipiv, int64_t pbs, View<T> X, void* ws, hipStream_t s);
// The two triangular inverses instead of their product: Li = L^{-1} P and
// Ui = U^{-1} (Ui may be LU itself; same workspace as getri)
template <typename T>
hipError_t getri_factors(int batch, int n, View<T> LU, const int* ipiv, int64_t pbs, View<T> Li, View<T> Ui,
                         void* ws, hipStream_t s);
// getrs for nrhs <= 8, non-transposed, one workgroup per matrix
// LU solve with the composed pivot permutation (perm from piv_to_perm_batched),
// nrhs <= 8, one wavefront per matrix
template <typename T>
hipError_t getrs_perm(int batch, int n, int nrhs, View<T> LU, const int* perm, int64_t qbs, View<T> B,
                      hipStream_t s);
hipError_t piv_to_perm_batched(int batch, int n, const int* ipiv, int64_t pbs, int* perm, int64_t qbs,
                               hipStream_t s);
template <typename T>
hipError_t getrs_small(int batch, int n, int nrhs, View<T> LU, const int* ipiv, int64_t pbs, View<T> B,
                       hipStream_t s);
// autotuned right-hand-side tile width of the chain sweeps for (dtype 0 =
// FP64 / 1 = FP32, block size, inverse factors); 0 = built-in rule.  set
// returns the previous width or -1 for bad arguments.
int sweep_tile_override(int dtype, int bs, bool inv);
int set_sweep_tile(int dtype, int bs, int inv, int nr);
// LU-form block-tridiagonal factorization / sweeps, one workgroup per chain
template <typename T>
bool btri_lu_fused_ok(int bs);
template <typename T>
hipError_t btri_lu_factor_fused(int nb, int m, int bs, T* diag, int64_t dbs, int64_t dis, T* sub, int64_t lbs,
                                int64_t lis, T* sup, int64_t ubs, int64_t uis, int* ipiv, int64_t pbs, int64_t pis,
                                int* info, hipStream_t s);
template <typename T>
hipError_t btri_lu_solve_fused(int nb, int m, int bs, T* diag, int64_t dbs, int64_t dis, T* sub, int64_t lbs,
                               int64_t lis, T* sup, int64_t ubs, int64_t uis, const int* ipiv, int64_t pbs,
                               int64_t pis, T* f, int64_t fbs, int64_t ldf, int nrhs, hipStream_t s,
                               bool trans = false);
// twisted variants: blocks 0..m/2-1 top-down, m-1..m/2+1 bottom-up, then the
// middle block m/2 (same storage; the factors serve the twisted sweeps only)
template <typename T>
hipError_t btri_lu_factor_twisted(int nb, int m, int bs, T* diag, int64_t dbs, int64_t dis, T* sub, int64_t lbs,
                                  int64_t lis, T* sup, int64_t ubs, int64_t uis, int* ipiv, int64_t pbs,
                                  int64_t pis, int* info, hipStream_t s);
// linv (optional): L^{-1} P of every diagonal block in diag's layout, diag
// then holding U^{-1} (getri_factors): every sweep stage is a block product.
// trans: solve with the transposed matrix.
template <typename T>
hipError_t btri_lu_solve_twisted(int nb, int m, int bs, T* diag, int64_t dbs, int64_t dis, T* sub, int64_t lbs,
                                 int64_t lis, T* sup, int64_t ubs, int64_t uis, const int* ipiv, int64_t pbs,
                                 int64_t pis, T* f, int64_t fbs, int64_t ldf, int nrhs, hipStream_t s,
                                 const T* linv = nullptr, bool trans = false);
// block-Thomas factorization / sweeps with explicit diagonal inverses
template <typename T>
bool btri_inv_fused_ok(int bs);
template <typename T>
hipError_t btri_inv_factor(int nb, int m, int bs, T* diag, int64_t dbs, int64_t dis, T* sub, int64_t lbs,
                           int64_t lis, T* sup, int64_t ubs, int64_t uis, int* info, T* ws, int* iws,
                           hipStream_t s);
template <typename T>
hipError_t btri_inv_solve(int nb, int m, int bs, T* diag, int64_t dbs, int64_t dis, T* sub, int64_t lbs,
                          int64_t lis, T* sup, int64_t ubs, int64_t uis, T* f, int64_t fbs, int64_t ldf, int nrhs,
                          T* ws, hipStream_t s);
}  // namespace slab
