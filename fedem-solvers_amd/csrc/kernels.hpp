// Host-callable launchers of the gfx950 kernels (kernels.hip).
#pragma once
#include <string>
#include "fgpu.hpp"
#include "kern_common.hpp"

namespace fgpu {

void launch_hex8(hipStream_t st, int nhex, const int* elist, const int* eptr, const int* enodes,
                 const double* xyz, const double* eprop, const int64_t* eoff, double* Kp, double* Mp,
                 int* ebad);
// ANDES4 shells (elem_shell.hip): P^T K P stiffness and FQS35 lumped mass
void launch_andes4(hipStream_t st, int nel, const int* elist, const int* eptr, const int* enodes, const double* xyz,
                   const double* eprop, const int64_t* eoff, double* Kp, double* Mp, int* ebad);
// 2-node beams (elem_beam.hip): BEAM31 stiffness and BEAM35 consistent mass
void launch_beam(hipStream_t st, int nbeam, const int* elist, const int* eptr, const int* enodes, const double* xyz,
                 const double* eprop, const int64_t* eoff, double* Kp, double* Mp, int* ebad);
// ccoef: per-contribution constraint coefficients, or nullptr (all 1)
void launch_assemble(hipStream_t st, int64_t nnz, const int* cptr, const int* cidx, const double* ccoef,
                     const double* Kp, const double* Mp, double* Kv, double* Mv, const int64_t* fmap, double* pool);
// Device construction of the matrix structure (structure.hip): neighbour
// lists sorted in permuted order with the DOF offset of each neighbour in
// the node's rows; CSR columns; the element-matrix contributions of every
// CSR entry (pass 0 counts into cnt, pass 1 appends at cptr + cnt; cidx =
// packed element pool position, ccoef the constraint coefficient or
// nullptr); their fixed order; the front map of the lower entries; an
// exclusive prefix sum (out has n + 1 entries, tsum scan_tiles(n) + 1).
// bad: bit 0 a contribution outside the CSR pattern, bit 1 an entry outside
// its front.
void launch_nbr_sort(hipStream_t st, int nnod, const int* nbptr, int* nbl, const int* nfirst, const int* neqd,
                     int* nbo);
void launch_csr_cols(hipStream_t st, int neq, const int* rnode, const int* rowptr, const int* nbptr, const int* nbl,
                     const int* nbo, const int* nfirst, const int* neqd, int* colind);
void launch_contrib(hipStream_t st, int pass, int nel, const int* etype, const int* eptr, const int* enodes,
                    const int64_t* eoff, const int* madof, const int* msc, const int* nndof, const int* ceq_ptr,
                    const int* ceq_node, const int* ceq_dof, const double* ceq_coef, const int* drank,
                    const int* nfirst, const int* rowptr, const int* nbptr, const int* nbl, const int* nbo, int* cnt,
                    const int* cptr, int* cidx, double* ccoef, int* bad);
void launch_seg_sort(hipStream_t st, int nnz, const int* cptr, int* cidx, double* ccoef);
void launch_fmap(hipStream_t st, int neq, const int* rowptr, const int* colind, const int* colsn, const DevSym& S,
                 int64_t* fmap, int* bad);
int scan_tiles(int n);
void launch_exclusive_scan(hipStream_t st, int n, const int* in, int* out, int* tsum);
void launch_fill(hipStream_t st, double* p, int64_t n, double v);
// no-allocation reduced loads (noalloc.hip): this GPU's scratch array
// (address and size), the permuted-row gather of column-major load vectors
// and the scatter of the reduced rows into Fred
int noalloc_scratch(double** ptr, size_t* bytes, std::string& err);
void launch_perm_rows(hipStream_t st, int neq, int nrhs, const int* p2e, const double* F, double* Fp);
void launch_scatter_reduced(hipStream_t st, int nd, int n2, int ngen, int nrhs, const int* pos2, const double* r2,
                            const double* r1, double* Fred);
// singular-pivot support of the factorisation: the original (or, with rs,
// equilibrated) diagonal of the first n permuted equations and its largest
// magnitude (amax must start at 0); the pivot statistics
// out[0..5] = negative, true-zero, reduced-to-zero pivots, first column of
// each (out must start at {0, 0, 0, INT_MAX, INT_MAX, INT_MAX})
void launch_diag_extract(hipStream_t st, int n, const int* rowptr, const int* colind, const double* val,
                         const double* rs, double* adiag, double* amax);
void launch_pivot_stats(hipStream_t st, int n, const double* dpiv, const int* sing, int* out);
// Pivot-block inverse tiles whose sum X_ij = -X_ii sum_k L_ik X_kj has more
// than INV_CHUNK terms are summed in INV_CHUNK-tile chunks in parallel.
constexpr int INV_CHUNK = 4;

// Factorisation and solve launchers for both factor storage types: double
// (FP64 path) and float (mixed precision: FP32 fronts and FP32 MFMA updates,
// diagonal blocks factored in FP64; solve vectors stay FP64).
#define FGPU_FACTOR_DECLS(T)                                                                              \
  void launch_extend_add(hipStream_t st, const FTask* t, int n, const DevSym& S, const int* xrange,        \
                         T* pool);                                                                         \
  void launch_diag(hipStream_t st, const FTask* t, int n, const DevSym& S, T* pool, const int64_t* ioff,  \
                   T* ipool, int* piv);                                                                   \
  void launch_diag_trsm(hipStream_t st, const FTask* t, int n, const DevSym& S, T* pool,                  \
                        const int64_t* ioff, T* ipool, int* piv);                                          \
  void launch_trsm(hipStream_t st, const FTask* t, int n, const DevSym& S, T* pool, const int64_t* ioff,  \
                   const T* ipool);                                                                       \
  void launch_update(hipStream_t st, const FTask* t, int n, const DevSym& S, T* pool);                    \
  void launch_inv_off(hipStream_t st, const FTask* t, int n, const DevSym& S, const T* pool,              \
                      const int64_t* ioff, T* ipool);                                                     \
  void launch_inv_part(hipStream_t st, const FTask* t, int n, const DevSym& S, const T* pool,             \
                       const int64_t* ioff, const T* ipool, T* ipart);                                    \
  void launch_inv_fin(hipStream_t st, const FTask* t, int n, const DevSym& S, const int64_t* ioff,        \
                      T* ipool, const T* ipart);                                                          \
  /* solve GEMM tasks, mode 0 FW_INV, 1 FW_UPD, 2 BW_UPD, 3 BW_INV */                                     \
  void launch_mv(hipStream_t st, const SolveTask* t, int n, const DevSym& S, const T* pool,               \
                 const int64_t* ioff, const T* ipool, double* W, double* V, double* X, double* part,      \
                 int nrhs, int mode);                                                                     \
  void launch_fw_out(hipStream_t st, const SolveTask* t, int n, const DevSym& S, const T* pool,           \
                     const double* W, const double* V, double* Y, int nrhs, int scale, int ndof1);
FGPU_FACTOR_DECLS(double)
FGPU_FACTOR_DECLS(float)
#undef FGPU_FACTOR_DECLS
void launch_mv_red(hipStream_t st, const SolveTask* t, int n, const DevSym& S, double* W, double* V, double* X,
                   const double* part, int nrhs, int mode);
void launch_fw_init(hipStream_t st, const SolveTask* t, int n, const DevSym& S, const int* gptr, const int* gsrc,
                    const int* ugrp, const int* ugroups, const double* part, const double* R, double* W, int nrhs,
                    int ndof1);
void launch_bw_init(hipStream_t st, const SolveTask* t, int n, const DevSym& S, const double* Y, double* W,
                    int nrhs);
void launch_spmm(hipStream_t st, int n, int nrhs, const int* rowptr, const int* colind, const double* val,
                 const double* X, int ldx, double* Z, int ldz,
                 bool skip_zeros = false);
int xty_chunks(int n);
// C = X^T Y (X n x na, Y n x nb, row-major; C na x nb row-major); acc
// (optional) = C, or acc + C with acc_add
void launch_xty(hipStream_t st, int n, int na, int nb, const double* X, int ldx, const double* Y, int ldy,
                double* part, double* C, double* acc = nullptr, bool acc_add = false);
// Y = alpha Y + beta X S (X n x na, S na x nb, row-major); up to 64 columns
// on FP64 MFMA, 16 columns of S per launch.  For nb <= 16, Y may be X itself
// (same leading dimension): each row tile reads its rows before it writes
// them.
void launch_gemm_small(hipStream_t st, int n, int na, int nb, const double* X, int ldx, const double* Sm,
                       double* Y, int ldy, double alpha, double beta);
// the same for two pairs (X1 -> Y1, X2 -> Y2) with one S, in one launch
void launch_gemm_small2(hipStream_t st, int n, int na, int nb, const double* X1, int ldx1, double* Y1, int ldy1,
                        const double* X2, int ldx2, double* Y2, int ldy2, const double* Sm, double alpha,
                        double beta);
// sign convention of the modes Phi (nrows x nev row-major): largest-magnitude
// internal component (rows p1[0..n1-1], equation order) made positive;
// sign: nev * (1 + 3 * MODE_SIGN_GROUPS) doubles of scratch
constexpr int MODE_SIGN_GROUPS = 256;
void launch_mode_signs(hipStream_t st, int n1, const int* p1, double* Phi, int64_t nrows, int nev, double* sign);
void launch_identity_rows(hipStream_t st, double* X, int64_t nrows, int nrhs, int ndof1);
// Lanczos: G (n x n, n <= 64) = R^T R, R and R^-1 row-major; *bad = 1 if
// G is not positive definite
void launch_chol_inv(hipStream_t st, int n, const double* G, double* R, double* Ri, int* bad);
// mixed precision: s = 1/sqrt(diag K) per perm row; equilibrated lower
// entries of K into the FP32 fronts; X = diag(rs) X diag(cs) (cs may be
// null); Y = a X + b Y; partial sums of squares of x and y over
// SUMSQ_BLOCKS blocks (2 doubles each, summed on the host in block order)
constexpr int SUMSQ_BLOCKS = 512;
void launch_diag_scale(hipStream_t st, int n, const int* rowptr, const int* colind, const double* val, double* s);
void launch_front_scatter32(hipStream_t st, int n, const int* rowptr, const int* colind, const double* val,
                            const double* s, const int64_t* fmap, float* pool);
void launch_scale_rc(hipStream_t st, int64_t nrows, int nrhs, const double* rs, const double* cs, double* X);
void launch_recip(hipStream_t st, int n, const double* x, double* y);
void launch_axpby(hipStream_t st, int64_t n, double a, const double* X, double b, double* Y);
void launch_sumsq2(hipStream_t st, int64_t n, const double* x, const double* y, double* part);
// Column dot products out[k * nr + c] = sum_r A_k[r][c] B_k[r][c] of npair
// (<= 3) pairs of row-major n x nr arrays, summed in a fixed order; part
// holds coldot_blocks(n) * npair * nr doubles.
constexpr int COLDOT_ROWS = 256;
int coldot_blocks(int64_t n);
void launch_coldots(hipStream_t st, int64_t n, int nr, int npair, const double* A0, const double* B0,
                    const double* A1, const double* B1, const double* A2, const double* B2, double* part,
                    double* out);
// Y = diag-scaled columns: Y[r][c] = a[c] X[r][c] + b[c] Y[r][c] (a, b device arrays; null: 1)
void launch_col_axpby(hipStream_t st, int64_t n, int nr, const double* a, const double* X, const double* b,
                      double* Y);
// element masses (x-translation mass of each element's mass blocks; nen
// nodes of d DOFs) of one element type list, and their fixed-order total
void launch_emass(hipStream_t st, int n, const int* elist, int nen, int d, const int64_t* eoff, const double* Mp,
                  double* emass);
void launch_sum_fixed(hipStream_t st, int n, const double* x, double* out);

}  // namespace fgpu
