/*
 * bk_oracle.c -- CPU restatement of the reference blocked Bunch-Kaufman
 * LDL^T (TEST INFRASTRUCTURE: the oracle, never the product).
 *
 * Every function cites the reference file:line it restates.  Storage is
 * flat (see bk_oracle.h); the loops below visit the same elements in the
 * same order with the same association as the reference, so with
 * -ffp-contract=off the results are bitwise equal to the reference build.
 *
 * Parity pins (tests/test_oracle.py), per exported function:
 *   bko_dsytrf        bitwise equal to the reference compiled from its own
 *                     sources (oracle/_ref): factors, perm, piv, workspace,
 *                     block geometry; golden vectors; the reference's 4x4
 *                     KAT; LAPACK dsytrf pivots; exact rational BK on tiny
 *                     matrices; the unpivoted closed form.
 *   bko_dsytrf_omp    bitwise equal to the reference's OpenMP build at the
 *                     extended stress sizes (n = 400..1000) and on the
 *                     golden stress case.
 *   bko_relative_residual   equal to the reference's RelativeError of
 *                     ReconstructOriginal() on the same cases.
 *   bko_sampled_residual, bko_reconstructed_entry   equal to the stress
 *                     test's SampledRelativeResidual (compiled from the
 *                     reference's test source) on the same factors, and to
 *                     the exact formula on planted errors.
 *   bko_extract_ld    through the full product L D L^T against the pinned
 *                     sampled residual; closed form on dominant matrices.
 *   bko_dsytrs        numpy's solve and the backward-error bound.
 *   bko_kernel_*      bitwise equal to the reference's kernels::* on
 *                     strided shapes, exact-zero pivots and empty sizes.
 */
#include "bk_oracle.h"

#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* src/block_ldlt_decomposition.cc:14 */
#define BKO_ALPHA 0.6403882032022076
/* src/block_ldlt_decomposition.cc:19 */
#define BKO_MIN_PARALLEL_BLOCKS 3

typedef struct {
  int n, bs, nb;
  double* F; /* factors_ */
  double* W; /* workspace_ */
  int* perm;
  int* piv;
  int* start; /* panel row0 */
  int* end;   /* panel row0 + rows */
  bko_decision* trace;
  int omp;
} bko_ctx;

#define AT(M, i, j) ((M)[(size_t)(i) * (size_t)c->n + (size_t)(j)])

/* src/block_ldlt_decomposition.cc:21-26 */
static double safe_denominator(double value) {
  if (fabs(value) <= DBL_EPSILON) return 0.0;
  return value;
}

static void swap_d(double* a, double* b) {
  double t = *a;
  *a = *b;
  *b = t;
}

static int parallel_gate(const bko_ctx* c, int panel) {
  return c->omp && (c->nb - panel > BKO_MIN_PARALLEL_BLOCKS);
}

/* CopyCurrentRowToWorkspace, src/block_ldlt_decomposition.cc:122-137:
 * W(k, j) = F(k, j) for j in [k, n). */
static void copy_current_row(bko_ctx* c, int panel, int k) {
  int n = c->n;
  int par = parallel_gate(c, panel);
#pragma omp parallel for schedule(static) if (par)
  for (int j = k; j < n; ++j) AT(c->W, k, j) = AT(c->F, k, j);
}

/* UpdateCurrentRowInWorkspace, src/block_ldlt_decomposition.cc:139-162:
 * W(k, j) -= sum_{p in [s, k)} F(p, k) * W(p, j), p ascending. */
static void update_current_row(bko_ctx* c, int panel, int s, int k) {
  if (k == s) return;
  int n = c->n;
  int par = parallel_gate(c, panel);
#pragma omp parallel for schedule(static) if (par)
  for (int j = k; j < n; ++j) {
    double update = 0.0;
    for (int p = s; p < k; ++p) update += AT(c->F, p, k) * AT(c->W, p, j);
    AT(c->W, k, j) -= update;
  }
}

/* FindCurrentRowMax, src/block_ldlt_decomposition.cc:164-182: first
 * argmax of |W(k, j)| over j > k with a strict '>' comparison. */
static void find_current_row_max(const bko_ctx* c, int k, double* omega,
                                 int* r) {
  double best = 0.0;
  int arg = 0;
  for (int j = k + 1; j < c->n; ++j) {
    double v = fabs(AT(c->W, k, j));
    if (v > best) {
      best = v;
      arg = j;
    }
  }
  *omega = best;
  *r = arg;
}

/* CopyCandidateToWorkspace, src/block_ldlt_decomposition.cc:184-247:
 * column r above the diagonal for rows [k, r), row r from r on. */
static void copy_candidate(bko_ctx* c, int k, int r) {
  for (int i = k; i < r; ++i) AT(c->W, i, r) = AT(c->F, i, r);
  for (int j = r; j < c->n; ++j) AT(c->W, r, j) = AT(c->F, r, j);
}

/* UpdateCandidateInWorkspace, src/block_ldlt_decomposition.cc:249-352. */
static void update_candidate(bko_ctx* c, int s, int k, int r) {
  if (k == s) return;
  for (int i = k; i < r; ++i) {
    double update = 0.0;
    for (int p = s; p < k; ++p) update += AT(c->F, p, r) * AT(c->W, p, i);
    AT(c->W, i, r) -= update;
  }
  for (int j = r; j < c->n; ++j) {
    double update = 0.0;
    for (int p = s; p < k; ++p) update += AT(c->F, p, r) * AT(c->W, p, j);
    AT(c->W, r, j) -= update;
  }
}

/* FindCandidateColumnMax + GetSymmetricGlobal,
 * src/block_ldlt_decomposition.cc:354-367,765-771 (std::max semantics). */
static double find_candidate_column_max(const bko_ctx* c, int k, int r) {
  double omega = 0.0;
  for (int i = k; i < c->n; ++i) {
    if (i == r) continue;
    double v = i < r ? fabs(AT(c->W, i, r)) : fabs(AT(c->W, r, i));
    omega = (omega < v) ? v : omega;
  }
  return omega;
}

/* SelectPivot, src/block_ldlt_decomposition.cc:369-390.
 * Returns 1 (1x1 at k), -1 (swap k<->r, 1x1) or 2 (swap k+1<->r, 2x2). */
static int select_pivot(const bko_ctx* c, int k, int r, double omega,
                        double* sigma_out, double* arr_out) {
  const double pivot_diagonal_abs = fabs(AT(c->W, k, k));
  const double candidate_column_max = find_candidate_column_max(c, k, r);
  const double candidate_diagonal_abs = fabs(AT(c->W, r, r));
  *sigma_out = candidate_column_max;
  *arr_out = candidate_diagonal_abs;
  if (pivot_diagonal_abs * candidate_column_max -
          BKO_ALPHA * omega * omega >=
      0.0)
    return 1;
  if (candidate_diagonal_abs - BKO_ALPHA * candidate_column_max >= 0.0)
    return -1;
  return 2;
}

/* SwapOneByOnePivotInMatrix, src/block_ldlt_decomposition.cc:414-498,
 * on the flat array M; the element pairs are disjoint so order is free. */
static void swap_1x1_in_matrix(bko_ctx* c, double* M, int s, int k, int r,
                               int swap_previous_rows) {
  int n = c->n;
  swap_d(&AT(M, k, k), &AT(M, r, r));
  for (int j = r + 1; j < n; ++j) swap_d(&AT(M, k, j), &AT(M, r, j));
  for (int j = k + 1; j < r; ++j) swap_d(&AT(M, k, j), &AT(M, j, r));
  if (swap_previous_rows)
    for (int i = 0; i < s; ++i) swap_d(&AT(M, i, k), &AT(M, i, r));
  for (int i = s; i < k; ++i) swap_d(&AT(M, i, k), &AT(M, i, r));
}

/* SwapTwoByTwoPivotInMatrix, src/block_ldlt_decomposition.cc:500-587. */
static void swap_2x2_in_matrix(bko_ctx* c, double* M, int s, int k, int r,
                               int swap_previous_rows) {
  int n = c->n;
  int next = k + 1;
  swap_d(&AT(M, next, next), &AT(M, r, r));
  swap_d(&AT(M, k, next), &AT(M, k, r));
  for (int j = r + 1; j < n; ++j) swap_d(&AT(M, next, j), &AT(M, r, j));
  for (int j = k + 2; j < r; ++j) swap_d(&AT(M, next, j), &AT(M, j, r));
  if (swap_previous_rows)
    for (int i = 0; i < s; ++i) swap_d(&AT(M, i, next), &AT(M, i, r));
  for (int i = s; i < k; ++i) swap_d(&AT(M, i, next), &AT(M, i, r));
}

/* SwapOneByOnePivot / SwapTwoByTwoPivot,
 * src/block_ldlt_decomposition.cc:392-412. */
static void swap_1x1(bko_ctx* c, int s, int k, int r) {
  swap_1x1_in_matrix(c, c->F, s, k, r, 1);
  swap_1x1_in_matrix(c, c->W, s, k, r, 0);
  int t = c->perm[k];
  c->perm[k] = c->perm[r];
  c->perm[r] = t;
}

static void swap_2x2(bko_ctx* c, int s, int k, int r) {
  swap_2x2_in_matrix(c, c->F, s, k, r, 1);
  swap_2x2_in_matrix(c, c->W, s, k, r, 0);
  int t = c->perm[k + 1];
  c->perm[k + 1] = c->perm[r];
  c->perm[r] = t;
}

/* ScaleOneByOnePivot, src/block_ldlt_decomposition.cc:633-656. */
static void scale_1x1(bko_ctx* c, int panel, int k) {
  int n = c->n;
  c->piv[k] = 1;
  const double diagonal = AT(c->W, k, k);
  const double denominator = safe_denominator(diagonal);
  int par = parallel_gate(c, panel);
#pragma omp parallel for schedule(static) if (par)
  for (int j = k + 1; j < n; ++j)
    AT(c->F, k, j) =
        denominator == 0.0 ? 0.0 : AT(c->W, k, j) / denominator;
  AT(c->F, k, k) = diagonal;
}

/* ScaleTwoByTwoPivot, src/block_ldlt_decomposition.cc:658-704.  The
 * reference also writes the lower (unused) slots of the diagonal block
 * when the 2x2 closes the panel (:686-689); those slots never feed an
 * upper-triangle value and are not restated. */
static void scale_2x2(bko_ctx* c, int panel, int k) {
  int n = c->n;
  int next = k + 1;
  c->piv[k] = 2;
  c->piv[next] = -2;
  const double a = AT(c->W, k, k);
  const double b = AT(c->W, k, next);
  const double d = AT(c->W, next, next);
  const double determinant = a * d - b * b;
  const double inverse_top_left = d / determinant;
  const double inverse_bottom_right = a / determinant;
  const double inverse_off_diagonal = -b / determinant;
  int par = parallel_gate(c, panel);
#pragma omp parallel for schedule(static) if (par)
  for (int j = k + 2; j < n; ++j) {
    const double w0 = AT(c->W, k, j);
    const double w1 = AT(c->W, next, j);
    AT(c->F, k, j) = inverse_top_left * w0 + inverse_off_diagonal * w1;
    AT(c->F, next, j) = inverse_off_diagonal * w0 + inverse_bottom_right * w1;
  }
  AT(c->F, k, k) = a;
  AT(c->F, k, next) = b;
  AT(c->F, next, next) = d;
}

/* ShiftBoundaryColumnToNextBlock, src/block_ldlt_decomposition.cc:589-631:
 * in flat storage no element moves; only the panel boundary does. */
static void shift_boundary(bko_ctx* c, int panel) {
  c->end[panel] -= 1;
  c->start[panel + 1] -= 1;
}

/* UpdateTrailingBlocks + kernels::UpdateTrailingBlock (non-MKL),
 * src/block_ldlt_decomposition.cc:706-744, src/kernels.cc:67-91:
 * F(i, j) -= sum_{p in panel} F(p, i) * W(p, j) for end <= i <= j < n.
 * (The reference also updates the lower half of diagonal blocks, which
 * never feeds an upper-triangle value.) */
static void update_trailing(bko_ctx* c, int panel) {
  int n = c->n;
  int s = c->start[panel], e = c->end[panel];
  if (e <= s) return;
  int par = c->omp && ((c->nb - panel - 1) * (c->nb - panel) / 2 >
                       BKO_MIN_PARALLEL_BLOCKS);
#pragma omp parallel if (par)
  {
    /* Per-row accumulator: each element still sums p = s..e-1 in order
     * into its own 'update' before the single subtraction, exactly as the
     * reference's i-j-k loop; only the loop nest is reordered. */
    double* upd = (double*)malloc(sizeof(double) * (size_t)n);
#pragma omp for schedule(dynamic, 8)
    for (int i = e; i < n; ++i) {
      for (int j = i; j < n; ++j) upd[j] = 0.0;
      for (int p = s; p < e; ++p) {
        const double l = AT(c->F, p, i);
        const double* wrow = &AT(c->W, p, 0);
        for (int j = i; j < n; ++j) upd[j] += l * wrow[j];
      }
      for (int j = i; j < n; ++j) AT(c->F, i, j) -= upd[j];
    }
    free(upd);
  }
}

static void record(bko_ctx* c, int k, int kind, int r, double omega,
                   double akk, double sigma, double arr) {
  if (!c->trace) return;
  bko_decision* d = &c->trace[k];
  d->kind = kind;
  d->r = (omega == 0.0) ? -1 : r;
  d->omega = omega;
  d->akk = akk;
  d->sigma = sigma;
  d->arr = arr;
  /* second largest |W(k, j)| over j > k, j != r (argmax gap) */
  double second = 0.0;
  for (int j = k + 1; j < c->n; ++j) {
    if (j == r) continue;
    double v = fabs(AT(c->W, k, j));
    if (v > second) second = v;
  }
  d->omega2 = second;
}

/* FactorPanel, src/block_ldlt_decomposition.cc:80-120. */
static void factor_panel(bko_ctx* c, int panel) {
  const int s = c->start[panel];
  const int panel_cols = c->end[panel] - s;
  int local_pivot = 0;
  while (local_pivot < panel_cols) {
    const int k = s + local_pivot;
    copy_current_row(c, panel, k);
    update_current_row(c, panel, s, k);
    double omega;
    int r;
    find_current_row_max(c, k, &omega, &r);
    int pivot_size = 1;
    const double diagonal = fabs(AT(c->W, k, k));
    int kind = 1;
    double sigma = 0.0, arr = 0.0;
    double akk = diagonal;
    double omega2 = 0.0;
    (void)omega2;
    if (!(diagonal >= BKO_ALPHA * omega || omega == 0.0)) {
      if (c->trace) record(c, k, 0, r, omega, akk, 0.0, 0.0);
      copy_candidate(c, k, r);
      update_candidate(c, s, k, r);
      const int choice = select_pivot(c, k, r, omega, &sigma, &arr);
      pivot_size = abs(choice);
      kind = choice == 1 ? 3 : choice;
      if (c->trace) {
        c->trace[k].kind = kind;
        c->trace[k].sigma = sigma;
        c->trace[k].arr = arr;
      }
      if (choice == -1) {
        swap_1x1(c, s, k, r);
      } else if (pivot_size == 2) {
        if (local_pivot == panel_cols - 1 && panel + 1 < c->nb) {
          shift_boundary(c, panel);
          break;
        }
        swap_2x2(c, s, k, r);
      }
    } else if (c->trace) {
      record(c, k, 1, r, omega, akk, 0.0, 0.0);
    }
    if (pivot_size == 1)
      scale_1x1(c, panel, k);
    else
      scale_2x2(c, panel, k);
    local_pivot += pivot_size;
  }
}

static int factorize(int n, int bs, double* F, double* W, int* perm,
                     int* piv, int* panel_start, int* panel_end,
                     bko_decision* trace, int omp) {
  if (n <= 0 || bs <= 0 || !F || !W || !perm || !piv || !panel_start ||
      !panel_end)
    return -1;
  bko_ctx ctx;
  bko_ctx* c = &ctx;
  c->n = n;
  c->bs = bs;
  c->nb = (n + bs - 1) / bs;
  c->F = F;
  c->W = W;
  c->perm = perm;
  c->piv = piv;
  c->start = panel_start;
  c->end = panel_end;
  c->trace = trace;
  c->omp = omp;
  for (int i = 0; i < n; ++i) {
    perm[i] = i; /* src/block_ldlt_decomposition.cc:45-47 */
    piv[i] = 0;  /* :43 */
  }
  if (trace) memset(trace, 0, sizeof(bko_decision) * (size_t)n);
  for (int P = 0; P < c->nb; ++P) { /* src/block_matrix.cc:215-222 */
    panel_start[P] = P * bs;
    panel_end[P] = (P + 1) * bs < n ? (P + 1) * bs : n;
  }
  for (int P = 0; P < c->nb; ++P) { /* Factorize, :50-59 */
    factor_panel(c, P);
    update_trailing(c, P);
  }
  return 0;
}

int bko_dsytrf(int n, int bs, double* F, double* W, int* perm, int* piv,
               int* panel_start, int* panel_end, bko_decision* trace) {
  return factorize(n, bs, F, W, perm, piv, panel_start, panel_end, trace, 0);
}

int bko_dsytrf_omp(int n, int bs, double* F, double* W, int* perm, int* piv,
                   int* panel_start, int* panel_end, bko_decision* trace) {
  return factorize(n, bs, F, W, perm, piv, panel_start, panel_end, trace, 1);
}

/* ---------------------------------------------------------------------- */
/* Oracle arithmetic on the results.                                       */

#undef AT
#define AT(M, i, j) ((M)[(size_t)(i) * (size_t)n + (size_t)(j)])

/* ExtractRegularFactorsFromCompact + IsTwoByTwoPivotCoupling,
 * src/block_ldlt_decomposition.cc:746-763,773-775. */
void bko_extract_ld(int n, const double* F, const int* piv, double* L,
                    double* D) {
  memset(L, 0, sizeof(double) * (size_t)n * (size_t)n);
  memset(D, 0, sizeof(double) * (size_t)n * (size_t)n);
  for (int row = 0; row < n; ++row) {
    AT(L, row, row) = 1.0;
    AT(D, row, row) = AT(F, row, row);
    for (int col = row + 1; col < n; ++col) {
      const double value = AT(F, row, col);
      if (col == row + 1 && piv[row] == 2) {
        AT(D, row, col) = value;
        AT(D, col, row) = value;
      } else {
        AT(L, col, row) = value;
      }
    }
  }
}

/* Multiply (non-MKL), src/regular_matrix.cc:92-116: out = left * right. */
static void dense_multiply(int n, const double* left, const double* right,
                           double* out) {
  memset(out, 0, sizeof(double) * (size_t)n * (size_t)n);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      const double lv = AT(left, i, k);
      for (int j = 0; j < n; ++j) AT(out, i, j) += lv * AT(right, k, j);
    }
}

/* FrobeniusNorm (non-MKL), src/regular_matrix.cc:75-90. */
static double frobenius(int n, const double* M) {
  double sum = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) sum += AT(M, i, j) * AT(M, i, j);
  return sqrt(sum);
}

double bko_relative_residual(int n, const double* A, const double* F,
                             const int* perm, const int* piv) {
  size_t nn = (size_t)n * (size_t)n;
  double* L = (double*)malloc(sizeof(double) * nn);
  double* D = (double*)malloc(sizeof(double) * nn);
  double* LD = (double*)malloc(sizeof(double) * nn);
  double* LT = (double*)malloc(sizeof(double) * nn);
  double* R = (double*)malloc(sizeof(double) * nn);
  bko_extract_ld(n, F, piv, L, D);
  dense_multiply(n, L, D, LD);
  for (int i = 0; i < n; ++i) /* Transpose, src/regular_matrix.cc:118-126 */
    for (int j = 0; j < n; ++j) AT(LT, j, i) = AT(L, i, j);
  dense_multiply(n, LD, LT, L); /* L now holds the permuted product */
  for (int i = 0; i < n; ++i)   /* ReconstructOriginal scatter, :71-76 */
    for (int j = 0; j < n; ++j) AT(R, perm[i], perm[j]) = AT(L, i, j);
  for (size_t t = 0; t < nn; ++t) /* Subtract, src/regular_matrix.cc:139-145 */
    LD[t] = R[t] - A[t];
  /* RelativeError, tests/block_ldlt_decomposition_test.cc:131-135 */
  double den = frobenius(n, A);
  if (den < 1.0) den = 1.0;
  double res = frobenius(n, LD) / den;
  free(L);
  free(D);
  free(LD);
  free(LT);
  free(R);
  return res;
}

/* CompactLValue, tests/block_ldlt_large_stress_test.cc:160-172. */
static double compact_l(int n, const double* F, const int* piv, int row,
                        int col) {
  if (row == col) return 1.0;
  if (row < col) return 0.0;
  if (piv[col] == 2 && row == col + 1) return 0.0;
  return AT(F, col, row);
}

/* ReconstructedPermutedEntry, tests/block_ldlt_large_stress_test.cc:174-194 */
double bko_reconstructed_entry(int n, const double* F, const int* piv,
                               int row, int col) {
  double value = 0.0;
  for (int p = 0; p < n; ++p) {
    const double rl = compact_l(n, F, piv, row, p);
    const double cl = compact_l(n, F, piv, col, p);
    value += rl * AT(F, p, p) * cl;
    if (piv[p] == 2) {
      const double off = AT(F, p, p + 1);
      const double rn = compact_l(n, F, piv, row, p + 1);
      const double cn = compact_l(n, F, piv, col, p + 1);
      value += off * (rl * cn + rn * cl);
    }
  }
  return value;
}

/* SampledRelativeResidual, tests/block_ldlt_large_stress_test.cc:196-244. */
double bko_sampled_residual(int n, const double* A, const double* F,
                            const int* perm, const int* piv, int count,
                            const int* sample_rows, const int* sample_cols) {
  int* inv = (int*)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; ++i) inv[perm[i]] = i;
  double rs = 0.0, os = 0.0;
  for (int t = 0; t < count; ++t) {
    const int orow = sample_rows[t], ocol = sample_cols[t];
    const double rec =
        bko_reconstructed_entry(n, F, piv, inv[orow], inv[ocol]);
    const double expected = AT(A, orow, ocol);
    const double diff = rec - expected;
    rs += diff * diff;
    os += expected * expected;
  }
  free(inv);
  double den = sqrt(os);
  if (den < 1.0) den = 1.0;
  return sqrt(rs) / den;
}

/* Solve with compact factors: (L D L^T)(i, j) = A(perm[i], perm[j]), so
 * A x = b  <=>  L D L^T y = c with c[i] = b[perm[i]], x[perm[i]] = y[i].
 * Forward/backward substitution walks the rows of U = L^T; 2x2 blocks use
 * LAPACK dsytrs's scaled formula. */
void bko_dsytrs(int n, int nrhs, const double* F, const int* perm,
                const int* piv, double* b) {
  double* y = (double*)malloc(sizeof(double) * (size_t)n);
  for (int q = 0; q < nrhs; ++q) {
    for (int i = 0; i < n; ++i) y[i] = b[(size_t)perm[i] * nrhs + q];
    for (int i = 0; i < n; ++i) { /* L z = c */
      const double yi = y[i];
      int j0 = (piv[i] == 2) ? i + 2 : i + 1;
      if (piv[i] == 2) {
        /* L(i+1, i) = 0 inside a 2x2 */
      }
      for (int j = j0; j < n; ++j) y[j] -= AT(F, i, j) * yi;
    }
    for (int i = 0; i < n;) { /* D w = z */
      if (piv[i] == 2) {
        const double akm1k = AT(F, i, i + 1);
        const double akm1 = AT(F, i, i) / akm1k;
        const double ak = AT(F, i + 1, i + 1) / akm1k;
        const double denom = akm1 * ak - 1.0;
        const double bkm1 = y[i] / akm1k;
        const double bk = y[i + 1] / akm1k;
        y[i] = (ak * bkm1 - bk) / denom;
        y[i + 1] = (akm1 * bk - bkm1) / denom;
        i += 2;
      } else {
        y[i] = y[i] / AT(F, i, i);
        i += 1;
      }
    }
    for (int i = n - 1; i >= 0; --i) { /* L^T v = w */
      int j0 = (piv[i] == 2) ? i + 2 : i + 1;
      double acc = y[i];
      for (int j = j0; j < n; ++j) acc -= AT(F, i, j) * y[j];
      y[i] = acc;
    }
    for (int i = 0; i < n; ++i) b[(size_t)perm[i] * nrhs + q] = y[i];
  }
  free(y);
}

#undef AT

/* kernels::ScaleOneByOnePivot (non-MKL), src/kernels.cc:11-31. */
void bko_kernel_scale_1x1(int count, double pivot, const double* src,
                          int src_stride, double* dst, int dst_stride) {
  if (count <= 0) return;
  if (fabs(pivot) == 0.0) {
    for (int i = 0; i < count; ++i) dst[(ptrdiff_t)i * dst_stride] = 0.0;
    return;
  }
  for (int i = 0; i < count; ++i)
    dst[(ptrdiff_t)i * dst_stride] = src[(ptrdiff_t)i * src_stride] / pivot;
}

/* kernels::ScaleTwoByTwoPivot, src/kernels.cc:33-65. */
void bko_kernel_scale_2x2(int count, double a, double b, double c,
                          const double* s1, int s1_stride, const double* s2,
                          int s2_stride, double* d1, int d1_stride, double* d2,
                          int d2_stride) {
  if (count <= 0) return;
  const double det = a * c - b * b;
  if (fabs(det) == 0.0) {
    for (int i = 0; i < count; ++i) {
      d1[(ptrdiff_t)i * d1_stride] = 0.0;
      d2[(ptrdiff_t)i * d2_stride] = 0.0;
    }
    return;
  }
  for (int i = 0; i < count; ++i) {
    const double v1 = s1[(ptrdiff_t)i * s1_stride];
    const double v2 = s2[(ptrdiff_t)i * s2_stride];
    d1[(ptrdiff_t)i * d1_stride] = (v1 * c - v2 * b) / det;
    d2[(ptrdiff_t)i * d2_stride] = (-v1 * b + v2 * a) / det;
  }
}

/* kernels::UpdateTrailingBlock (non-MKL), src/kernels.cc:67-91. */
void bko_kernel_update_trailing(int rows, int cols, int inner,
                                const double* left, int left_stride,
                                const double* work, int work_stride,
                                double* target, int target_stride) {
  if (rows <= 0 || cols <= 0 || inner <= 0) return;
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) {
      double update = 0.0;
      for (int k = 0; k < inner; ++k)
        update += left[(ptrdiff_t)k * left_stride + i] *
                  work[(ptrdiff_t)k * work_stride + j];
      target[(ptrdiff_t)i * target_stride + j] -= update;
    }
}
