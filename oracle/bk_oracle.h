/*
 * bk_oracle.h -- CPU restatement of the reference blocked Bunch-Kaufman
 * LDL^T factorization (test infrastructure only).
 *
 * THIS IS THE ORACLE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker.  The product path
 * (blockedbunchkaufmanldlt_amd) never links or calls anything here.
 *
 * Storage model.  The reference keeps the upper triangle in separately
 * allocated blocks (src/block_matrix.cc:209-231) whose views move when a
 * 2x2 pivot straddles a panel edge (src/block_ldlt_decomposition.cc:589-631).
 * Every element keeps its global (row, col) position through those moves,
 * so the restatement stores factors_ and workspace_ as two flat n x n
 * row-major arrays (only i <= j is meaningful) and tracks only the panel
 * boundaries [panel_start[P], panel_end[P]).
 *
 * Arithmetic.  Operation order and association follow the reference
 * exactly; the file is compiled with -ffp-contract=off and no -march, as
 * the reference is, so results are bitwise equal to the reference build.
 */
#ifndef BK_ORACLE_H_
#define BK_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Per-pivot decision record, indexed by the pivot's global column k. */
typedef struct {
  int kind;        /* 1: 1x1 (|a_kk| >= alpha*omega or omega == 0)
                      3: 1x1 after the candidate test (no swap)
                      -1: swapped 1x1, 2: 2x2 starting at k, 0: not a pivot */
  int r;           /* candidate column (argmax of row k), -1 if none */
  double omega;    /* max_{j>k} |W(k,j)| */
  double omega2;   /* largest |W(k,j)|, j>k, j != r (argmax gap) */
  double akk;      /* |W(k,k)| */
  double sigma;    /* max_{i>=k, i!=r} |W(i,r)| (0 if not computed) */
  double arr;      /* |W(r,r)| (0 if not computed) */
} bko_decision;

/* Blocked BK LDL^T, restating BlockLdltDecomposition::Factorize()
 * (src/block_ldlt_decomposition.cc:50-59) in kCompactOnly form.
 *   F : n*n row-major; upper triangle = A on entry, compact factors on exit
 *   W : n*n row-major scratch, must be zero on entry (workspace_ restated)
 *   perm, piv : n ints each (Permutation(), PivotTypes())
 *   panel_start, panel_end : ceil(n/bs) ints each (final panel geometry)
 *   trace : n records or NULL
 * Returns 0, or -1 on invalid arguments. */
int bko_dsytrf(int n, int bs, double* F, double* W, int* perm, int* piv,
               int* panel_start, int* panel_end, bko_decision* trace);

/* Same as bko_dsytrf with the reference's OpenMP regions enabled
 * (src/block_ldlt_decomposition.cc:125-128,146-149,642-645,679-682,713-718).
 * Bitwise identical to bko_dsytrf: every region writes disjoint elements
 * with serial inner sums.  Used as the CPU baseline. */
int bko_dsytrf_omp(int n, int bs, double* F, double* W, int* perm, int* piv,
                   int* panel_start, int* panel_end, bko_decision* trace);

/* Dense L (unit lower) and D (block diagonal) from compact factors,
 * restating ExtractRegularFactorsFromCompact
 * (src/block_ldlt_decomposition.cc:746-763,773-775).  L, D: n*n row-major. */
void bko_extract_ld(int n, const double* F, const int* piv, double* L,
                    double* D);

/* ||P L D L^T P^T - A||_F / max(||A||_F, 1): ReconstructOriginal
 * (src/block_ldlt_decomposition.cc:61-78) followed by RelativeError
 * (tests/block_ldlt_decomposition_test.cc:131-135).  A is n*n dense. */
double bko_relative_residual(int n, const double* A, const double* F,
                             const int* perm, const int* piv);

/* One entry of (L D L^T)(row, col) from compact factors, restating
 * ReconstructedPermutedEntry (tests/block_ldlt_large_stress_test.cc:174-194). */
double bko_reconstructed_entry(int n, const double* F, const int* piv,
                               int row, int col);

/* Sampled relative residual over the given (orig_row, orig_col) pairs,
 * restating SampledRelativeResidual
 * (tests/block_ldlt_large_stress_test.cc:220-244).  A is n*n dense. */
double bko_sampled_residual(int n, const double* A, const double* F,
                            const int* perm, const int* piv, int count,
                            const int* sample_rows, const int* sample_cols);

/* Solve A x = b (nrhs columns, b/x row-major n x nrhs) with the compact
 * factors: x = P L^-T D^-1 L^-1 P^T b.  2x2 blocks use LAPACK dsytrs's
 * scaled formula.  No reference counterpart (new in this build). */
void bko_dsytrs(int n, int nrhs, const double* F, const int* perm,
                const int* piv, double* b);

/* Standalone helpers restating src/kernels.cc (test-only in the reference). */
void bko_kernel_scale_1x1(int count, double pivot, const double* src,
                          int src_stride, double* dst, int dst_stride);
void bko_kernel_scale_2x2(int count, double a, double b, double c,
                          const double* s1, int s1_stride, const double* s2,
                          int s2_stride, double* d1, int d1_stride, double* d2,
                          int d2_stride);
void bko_kernel_update_trailing(int rows, int cols, int inner,
                                const double* left, int left_stride,
                                const double* work, int work_stride,
                                double* target, int target_stride);

#ifdef __cplusplus
}
#endif

#endif /* BK_ORACLE_H_ */
