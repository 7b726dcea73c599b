/* libslabhip: C ABI of the MI355X (gfx950) SlabLU factor+solve kernels.
 *
 * Conventions
 *   - dtype: SLAB_F64 (0) or SLAB_F32 (1); every pointer is device memory
 *     owned by the caller (the library never frees caller memory).
 *   - Matrices are strided views: element (i, j) of batch b is at
 *       ptr + b*bs + i*rs + j*cs   (strides in elements).
 *     Transposed operands are passed by swapping rs and cs.
 *   - Pivots are int32, 1-based (LAPACK getrf convention).
 *   - Every call is asynchronous on `stream` (a hipStream_t; NULL = the
 *     default stream).  Return value: 0 on success, < 0 for an invalid
 *     argument (-1) or a HIP launch error (-(1000 + hipError_t)).  Exactly
 *     zero pivots are reported through the optional `info` device array
 *     (1-based row of the first zero pivot, 0 if none), as LAPACK does.
 *
 * The Python host layer (slablu_amd/) keeps the reference's API and calls
 * these entry points through ctypes; each declaration names the reference
 * code it replaces (paths relative to the annayesy/slabLU checkout).
 */
#ifndef SLAB_HIP_H
#define SLAB_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SLAB_F64 = 0, SLAB_F32 = 1 };

/* Library version (major*10000 + minor*100 + patch). */
int slab_version(void);

/* Batched C = alpha*A@B + beta*C (A m x k, B k x n).
 * Replaces the batched matmuls of src_disc/hps_parallel_leaf_ops.py:58 and
 * src_solver/block_tridiag_solver_dense.py:169,193,200. */
int slab_gemm(int dtype, int batch, int m, int n, int k, double alpha,
              const void* A, int64_t a_rs, int64_t a_cs, int64_t a_bs,
              const void* B, int64_t b_rs, int64_t b_cs, int64_t b_bs,
              double beta, void* C, int64_t c_rs, int64_t c_cs, int64_t c_bs, void* stream);

/* Batched left triangular solve B <- A^{-1} B, A m x m lower/upper (as
 * viewed), unit or non-unit diagonal.  Replaces torch.linalg.solve_triangular
 * in src_solver/block_tridiag_solver_dense.py:111-140. */
int slab_trsm(int dtype, int batch, int m, int n, int lower, int unit,
              const void* A, int64_t a_rs, int64_t a_cs, int64_t a_bs,
              void* B, int64_t b_rs, int64_t b_cs, int64_t b_bs, void* stream);

/* Batched row interchanges rows k1..k2-1 (forward or backward order). */
int slab_laswp(int dtype, int batch, int ncols, int k1, int k2, int forward,
               void* A, int64_t a_rs, int64_t a_cs, int64_t a_bs,
               const int* ipiv, int64_t p_bs, void* stream);

/* Batched LU with partial pivoting eliminating the first nelim columns of
 * an nrows x ncols row-major matrix (lda, batch stride a_bs); pivot rows are
 * restricted to rows < npiv.  nrows = ncols = nelim = npiv is a plain getrf
 * (torch.linalg.lu_factor, src_solver/block_tridiag_solver_dense.py:157,169);
 * nelim < ncols additionally forms L^{-1} P A12, and nrows > npiv carries
 * bordered rows whose trailing block becomes the Schur complement. */
int slab_getrf(int dtype, int batch, int nrows, int ncols, int nelim, int npiv,
               void* A, int64_t lda, int64_t a_bs, int* ipiv, int64_t p_bs, int* info, void* stream);

/* Batched solve with getrf factors: op(A) X = B (n x n, nrhs columns).
 * Replaces torch.linalg.lu_solve (block_tridiag_solver_structured.py:58). */
int slab_getrs(int dtype, int batch, int n, int nrhs, int trans,
               const void* LU, int64_t lda, int64_t lu_bs, const int* ipiv, int64_t p_bs,
               void* B, int64_t ldb, int64_t b_bs, void* stream);

/* Batched orthonormal basis: Q (m x k, row-major) with orthonormal columns
 * spanning the columns of each A (m x k, m >= k, k <= 256; A is destroyed)
 * by Householder QR, k columns even when A is rank-deficient.  Replaces the
 * torch.linalg.qr calls of the HBS reconstruction
 * (src_solver/hbs_reconstruction.py:454-470): the node bases U, V. */
int slab_orth(int dtype, int batch, int m, int k, void* A, int64_t lda, int64_t a_bs, void* Q, int64_t ldq,
              int64_t q_bs, void* stream);

/* Batched inverse X = A^{-1} (n x n, out of place) from the factors of
 * slab_getrf: both triangles inverted by recursive doubling of their
 * diagonal blocks, then X = U^{-1} L^{-1} P.  Replaces torch.linalg.lu_factor
 * + lu_solve of each interface Schur complement against its coupling block
 * (block_tridiag_solver_structured.py:46-65): with the inverse kept, every
 * later solve is a product.  ws: device workspace of slab_getri_ws bytes. */
size_t slab_getri_ws(int dtype, int batch, int n);
int slab_getri(int dtype, int batch, int n, const void* LU, int64_t lda, int64_t lu_bs, const int* ipiv,
               int64_t p_bs, void* X, int64_t ldx, int64_t x_bs, void* ws, void* stream);

/* slab_getri with no allocation and no workspace argument, for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops); it has no counterpart in the reference.  The workspace is a
 * file-scope device array of SLAB_GETRI_NOALLOC_MAX_BYTES (40 MiB) that each
 * device holds for the life of the process: problems whose slab_getri_ws
 * exceeds it are rejected with -2 before any launch (e.g. FP64: one 1000 x
 * 1000 matrix fits, two do not).  Every call on one device shares that
 * device's array, so calls that overlap on one device (concurrent streams,
 * threads) race; calls in order on one stream do not.  Otherwise as
 * slab_getri (same results, bit for bit). */
#define SLAB_GETRI_NOALLOC_MAX_BYTES (40u * 1024u * 1024u)
int slab_getri_noalloc(int dtype, int batch, int n, const void* LU, int64_t lda, int64_t lu_bs, const int* ipiv,
                       int64_t p_bs, void* X, int64_t ldx, int64_t x_bs, void* stream);

/* perm[r] (0-based) = the row of B that the interchanges ipiv (1-based, as
 * slab_getrf writes them) bring to row r; batched, perm row stride q_bs.
 * Lets repeated solves with one factorization skip replaying the swaps. */
int slab_piv_to_perm(int batch, int n, const int* ipiv, int64_t p_bs, int* perm, int64_t q_bs, void* stream);

/* slab_getrs (non-transposed, 1 <= nrhs <= 8) with the permutation of
 * slab_piv_to_perm, one wavefront per matrix: the leaf interior solves
 * A_cc^{-1} f (torch.linalg.solve in src_disc/hps_parallel_leaf_ops.py:86-100)
 * with the factors kept from the leaf condensation. */
int slab_getrs_perm(int dtype, int batch, int n, int nrhs, const void* LU, int64_t lda, int64_t lu_bs,
                    const int* perm, int64_t q_bs, void* B, int64_t ldb, int64_t b_bs, void* stream);

/* Block-tridiagonal LU without inter-block pivoting over b independent
 * chains of m blocks of size bs (row-major bs x bs blocks).  Overwrites
 * diag_i with the packed LU of D_i - Lfac_i Ufac_i, sub_i with
 * Lfac = sub_i U_i^{-1} and sup_i with Ufac = L_i^{-1} P_i^T sup_i.
 * Block i of chain k is at ptr + k*bstride + i*istride.
 * Replaces build_sweep (src_solver/block_tridiag_solver_dense.py:152-177)
 * and the interface Schur recursion
 * (src_solver/block_tridiag_solver_structured.py:34-76), which is the same
 * factorization with b = 1. */
int slab_btri_factor(int dtype, int b, int m, int bs,
                     void* diag, int64_t d_bstride, int64_t d_istride,
                     void* sub, int64_t l_bstride, int64_t l_istride,
                     void* sup, int64_t u_bstride, int64_t u_istride,
                     int* ipiv, int64_t p_bstride, int64_t p_istride, int* info, void* stream);

/* Forward/backward block sweeps with the factors of slab_btri_factor, in
 * place on f (b chains of m*bs rows, nrhs columns, row stride ldf, chain
 * stride f_bstride); transpose solves A^T x = f.
 * Replaces lusolve_sweep / lusolve_sweep_transpose
 * (src_solver/block_tridiag_solver_dense.py:179-222) and
 * HBS_BlockTridiagonal solve (block_tridiag_solver_structured.py:157-187). */
int slab_btri_solve(int dtype, int b, int m, int bs,
                    const void* diag, int64_t d_bstride, int64_t d_istride,
                    const void* sub, int64_t l_bstride, int64_t l_istride,
                    const void* sup, int64_t u_bstride, int64_t u_istride,
                    const int* ipiv, int64_t p_bstride, int64_t p_istride,
                    void* f, int64_t f_bstride, int64_t ldf, int nrhs, int transpose, void* stream);

/* Twisted variant of slab_btri_factor (same arguments and storage): the
 * blocks 0..k-1 (k = m/2) are eliminated top-down and m-1..k+1 bottom-up,
 * both chains of every system concurrently, then the middle block k.  The
 * factors are only meaningful to slab_btri_solve_twisted.  The same
 * Gaussian elimination as build_sweep
 * (src_solver/block_tridiag_solver_dense.py:152-177) in a different block
 * order, with half the dependent steps.  Requires
 * slab_btri_fused(dtype, bs) == 1; returns -1 otherwise. */
int slab_btri_factor_twisted(int dtype, int b, int m, int bs,
                             void* diag, int64_t d_bstride, int64_t d_istride,
                             void* sub, int64_t l_bstride, int64_t l_istride,
                             void* sup, int64_t u_bstride, int64_t u_istride,
                             int* ipiv, int64_t p_bstride, int64_t p_istride, int* info, void* stream);

/* Sweeps with the factors of slab_btri_factor_twisted, in place on f (as
 * slab_btri_solve); transpose solves A^T x = f.  Replaces lusolve_sweep and
 * lusolve_sweep_transpose (src_solver/block_tridiag_solver_dense.py:179-222)
 * for twisted factors. */
int slab_btri_solve_twisted(int dtype, int b, int m, int bs,
                            const void* diag, int64_t d_bstride, int64_t d_istride,
                            const void* sub, int64_t l_bstride, int64_t l_istride,
                            const void* sup, int64_t u_bstride, int64_t u_istride,
                            const int* ipiv, int64_t p_bstride, int64_t p_istride,
                            void* f, int64_t f_bstride, int64_t ldf, int nrhs, int transpose, void* stream);

/* Inverse triangular factors of every diagonal block after
 * slab_btri_factor (or _twisted): diag's packed LU is replaced by U^{-1}
 * and linv (same layout) receives L^{-1} P, so the sweeps of
 * slab_btri_solve_twisted_inv are block products on the matrix cores
 * instead of triangular substitutions.  The Gaussian elimination and its
 * pivots are those of block_tridiag_solver_dense.py:152-177; only the way
 * each diagonal block is applied differs.  Blocks must be contiguous
 * (d_istride = bs*bs, d_bstride = m*bs*bs, p_istride = bs, p_bstride =
 * m*bs); ws: slab_getri_ws(dtype, b*m, bs) bytes. */
int slab_btri_invert_diag(int dtype, int b, int m, int bs, void* diag, int64_t d_bstride, int64_t d_istride,
                          const int* ipiv, int64_t p_bstride, int64_t p_istride, void* linv, void* ws,
                          void* stream);

/* slab_btri_solve_twisted with the factors of slab_btri_invert_diag
 * (lusolve_sweep / lusolve_sweep_transpose,
 * src_solver/block_tridiag_solver_dense.py:179-222, with the permutation
 * folded into L^{-1} P). */
int slab_btri_solve_twisted_inv(int dtype, int b, int m, int bs,
                                const void* diag, int64_t d_bstride, int64_t d_istride, const void* linv,
                                const void* sub, int64_t l_bstride, int64_t l_istride,
                                const void* sup, int64_t u_bstride, int64_t u_istride,
                                void* f, int64_t f_bstride, int64_t ldf, int nrhs, int transpose, void* stream);

/* 1 when blocks of size bs run in the fused one-workgroup-per-chain LU-form
 * kernels (the diagonal block, a right-hand-side chunk and staging fit in
 * LDS), 0 otherwise, -1 for a bad dtype. */
int slab_btri_fused(int dtype, int bs);

/* Block-tridiagonal factorization in block-Thomas form with explicit
 * diagonal inverses (same chains and storage as slab_btri_factor):
 *   S_0 = D_0, S_i = D_i - sub_{i-1} G_{i-1}, G_i = S_i^{-1} sup_i;
 * diag_i <- S_i^{-1}, sup_i <- G_i, sub_i unchanged.  This is the Schur
 * recursion of src_solver/block_tridiag_solver_structured.py:34-76 with
 * the inverse kept instead of the LU, so that every sweep step is a matrix
 * product.  Blocks up to slab_btri_inv_fused(dtype, bs) == 1 run as one
 * workgroup per chain (Gauss-Jordan inverse in LDS); otherwise ws must hold
 * 2*b*bs*bs elements and iws b*bs ints. */
int slab_btri_inv_fused(int dtype, int bs);
int slab_btri_inv_factor(int dtype, int b, int m, int bs,
                         void* diag, int64_t d_bstride, int64_t d_istride,
                         void* sub, int64_t l_bstride, int64_t l_istride,
                         void* sup, int64_t u_bstride, int64_t u_istride,
                         int* info, void* ws, int* iws, void* stream);

/* Sweeps with the factors of slab_btri_inv_factor, in place on f (as
 * slab_btri_solve, non-transposed).  Replaces lusolve_sweep
 * (src_solver/block_tridiag_solver_dense.py:179-202) and
 * HBS_BlockTridiagonal solve (block_tridiag_solver_structured.py:157-187).
 * For non-fused block sizes ws must hold b*bs*nrhs elements. */
int slab_btri_inv_solve(int dtype, int b, int m, int bs,
                        const void* diag, int64_t d_bstride, int64_t d_istride,
                        const void* sub, int64_t l_bstride, int64_t l_istride,
                        const void* sup, int64_t u_bstride, int64_t u_istride,
                        void* f, int64_t f_bstride, int64_t ldf, int nrhs, void* ws, void* stream);

/* Leaf collocation rows for static condensation, all leaves in one launch:
 * M = [A_cc | A_cx] (border = 0) or [A_cc | A_cx ; N_c | N_x] (border = 1)
 * per leaf, from the scaled 1-D Chebyshev matrix D (p x p), Dsq = D@D and
 * per-node coefficients (nboxes*p*p FP64 each; NULL = term absent; c11 and
 * c22 required).  Nxtot (4(p-2) x ((p-2)^2 + 4(p-2)), FP64) is copied into
 * the border rows.  Replaces get_Aloc_2d / Aloc_acc and the Acc/Acx
 * gathers of src_disc/hps_parallel_leaf_ops.py:10-58. */
int slab_leaf_assemble(int dtype, int nboxes, int p, int border, const double* D, const double* Dsq,
                       const double* c11, const double* c22, const double* c12,
                       const double* c1, const double* c2, const double* c,
                       const double* Nxtot, void* M, int64_t ldm, int64_t m_bs, void* stream);

/* Interior rows [A_cc | A_c,X] of every leaf where the columns past the
 * (p-2)^2 interior ones are the nxc leaf nodes xnodes (node id i*p + j),
 * e.g. the 4(p-1) corner-inclusive exterior nodes in Jxreorder order, which
 * the condensation needs when the cross term c12 couples interior and
 * corner nodes.  Coefficients and accumulation as slab_leaf_assemble.
 * Replaces the c12 branch of form_DtNs, Aloc[Jc][:, Jxreorder]
 * (src_disc/hps_parallel_leaf_ops.py:59-65). */
int slab_leaf_assemble_nodes(int dtype, int nboxes, int p, int nxc, const int* xnodes, const double* D,
                             const double* Dsq, const double* c11, const double* c22, const double* c12,
                             const double* c1, const double* c2, const double* c, void* M, int64_t ldm, int64_t m_bs,
                             void* stream);

/* FP64 residual of the un-condensed leaf collocation system for every leaf
 * in one launch (matrix-free): node values from interior values uc
 * (nboxes, (p-2)^2, nrhs) at the Jc nodes and per-leaf face values ub
 * (nboxes, 4(p-2), nrhs) through Intmap ((p^2-(p-2)^2) x 4(p-2), rows at the
 * Jxreo nodes); outputs rI = -(A_loc u) at the interior nodes and
 * fl = N_x u at the face nodes (N_x as nx_idx/nx_val, p entries per row);
 * uu (nboxes, p^2, nrhs) receives all node values when not NULL.
 * Coefficients as in slab_leaf_assemble.  Element (b, row, r) of every array
 * sits at base + b*bs + row*ld + r.  Replaces the leaf residual
 * Aloc @ u - f and the flux N_x u of src_disc/hps_parallel_leaf_ops.py:73-92
 * and hps_multidomain_disc.py:227-243 (resloc), applied to a given iterate. */
int slab_leaf_residual(int nboxes, int p, int nrhs, const double* uc, int64_t uc_ld, int64_t uc_bs,
                       const double* ub, int64_t ub_ld, int64_t ub_bs, const double* Intmap, const int* Jc,
                       const int* Jxreo, const double* D, const double* Dsq, const double* c11, const double* c22,
                       const double* c12, const double* c1, const double* c2, const double* c, const int* nx_idx,
                       const double* nx_val, double* rI, int64_t rI_ld, int64_t rI_bs, double* fl, int64_t fl_ld,
                       int64_t fl_bs, double* uu, int64_t uu_ld, int64_t uu_bs, void* stream);

/* Right-hand-side tile width (nr in {1, 2, 4, 8, 16, 32}; 0 restores the
 * built-in rule) of the one-workgroup-per-chain sweeps for blocks of size
 * bs in dtype, LU form (inv_factors = 0) or inverse-factor form (1): the
 * knob the autotuner (slablu_amd/autotune.py) sets after timing candidate
 * widths on the actual chains.  Process-wide; returns the previous width,
 * -1 for bad arguments.  No counterpart in the reference (its sweeps are
 * torch library calls, block_tridiag_solver_dense.py:179-222). */
int slab_set_sweep_tile(int dtype, int bs, int inv_factors, int nr);

/* Batched Krylov products of the FP64 refinement, all right-hand-side
 * columns in one launch: out[c*out_cs + i] = V_c[i] . w_c for i < k <= 32,
 * where vector i of column c starts at V + c*v_cs + i*v_ks and w_c at
 * w + c*w_cs (n entries each); ws: slab_mdot_ws(nc, k, n) bytes of device
 * workspace for the fixed-order partial sums.  No counterpart in the
 * reference (its solve is direct); used by the FGMRES refinement of the
 * mixed-precision form (Gram-Schmidt projections, norms, Gram matrices). */
size_t slab_mdot_ws(int nc, int k, int64_t n);
int slab_mdot(int nc, int k, int64_t n, const double* V, int64_t v_cs, int64_t v_ks, const double* w, int64_t w_cs,
              double* out, int64_t out_cs, void* ws, void* stream);

/* w_c += alpha * sum_{i<k} coef[c*coef_cs + i] V_c[i] for every column c
 * (layouts as slab_mdot); the Gram-Schmidt update and the FGMRES solution
 * update of the refinement. */
int slab_maxpy(int nc, int k, int64_t n, double alpha, const double* V, int64_t v_cs, int64_t v_ks,
               const double* coef, int64_t coef_cs, double* w, int64_t w_cs, void* stream);

/* Launch probe: HIP events around every launch of one kernel, on the stream
 * it is launched on, with each launch's algorithmic flops (2 m n k per
 * product) and bytes (A, B, C read once, C written once), so that a caller
 * times one kernel inside a larger run without a profiler.  Launches into a
 * stream under graph capture are not probed.  slab_probe_start creates
 * 2 * max_launches events (-3 when a probe is running); slab_probe_stop
 * waits for the recorded events and returns the probed launches, the
 * launches past max_launches (not timed), the summed event intervals (ms)
 * and the summed flops and bytes.  One probe per process at a time.  No
 * counterpart in the reference (measurement of this build's kernels). */
enum {
  SLAB_PROBE_GEMM_F32 = 1, /* k_gemm<float, int, true, true>: FP32 MFMA GEMM, 16-byte rows of A, B and C */
  SLAB_PROBE_GEMM_F64 = 2  /* k_gemm<double, int, false, false>: FP64 MFMA GEMM, strided operands */
};
int slab_probe_start(int kernel, int max_launches);
int slab_probe_stop(int* launches, int* dropped, double* total_ms, double* flops, double* bytes);

#ifdef __cplusplus
}
#endif
#endif /* SLAB_HIP_H */
