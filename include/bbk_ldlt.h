/*
 * bbk_ldlt.h -- C ABI of the MI355X (gfx950) blocked Bunch-Kaufman LDL^T
 * library (libbbk_ldlt.so).
 *
 * Matrix layout.  Every matrix argument is row-major with leading
 * dimension lda >= n, and only its UPPER triangle (i <= j) is read and
 * written.  Row-major upper is byte-identical to LAPACK's column-major
 * lower ('L'), so a LAPACK caller passes its 'L' array unchanged.
 *
 * Output encoding (identical to the reference's compact factor storage,
 * include/block_ldlt_decomposition.h:25-30 and
 * src/block_ldlt_decomposition.cc:633-704,746-775):
 *   perm[i]  : (L D L^T)(i, j) = A(perm[i], perm[j])   -- Permutation()
 *   piv[i]   : 1 for a 1x1 pivot; 2, -2 at i, i+1 for a 2x2 -- PivotTypes()
 *   A(i, i)  : D(i, i);  A(i, i+1) = D(i, i+1) when piv[i] == 2
 *   A(i, j)  : L(j, i) for every other i < j (L unit lower, L(i+1,i) = 0
 *              inside a 2x2).
 * Pivot tests, 2x2 inversion and |d| <= DBL_EPSILON singular handling
 * follow src/block_ldlt_decomposition.cc:21-26,369-390 in binary64 for
 * every precision.
 *
 * Errors.  Functions return BBK_OK (0), a negative value -i when argument
 * i (1-based, handle = 1) is invalid (LAPACK convention), or a positive
 * BBK_ERR_* code.  The C++ layer (naive_block_ldlt) maps these to the
 * reference's exception types.
 *
 * Threading.  A handle owns one HIP stream and its workspace; calls on one
 * handle must not overlap.  All device-pointer functions are asynchronous
 * on the handle's stream unless stated otherwise.
 */
#ifndef BBK_LDLT_H_
#define BBK_LDLT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BBK_OK 0
#define BBK_ERR_HIP 1          /* a HIP runtime call failed */
#define BBK_ERR_ALLOC 2        /* device/host allocation failed */
#define BBK_ERR_UNSUPPORTED 3  /* size beyond what the kernels handle */
#define BBK_ERR_NOT_CONVERGED 4 /* refinement stopped above tolerance */

typedef struct bbk_handle* bbk_handle_t;

/* Create a handle on the current HIP device.  stream may be NULL (the
 * handle then creates its own default-priority non-blocking stream; the
 * caller orders it against its own streams, e.g. with events).  A given
 * stream must not be the null stream or another stream that synchronizes
 * with it: cooperative launches are refused there. */
int bbk_create(bbk_handle_t* handle, void* stream);
/* The stream the handle queues its work on. */
int bbk_get_stream(bbk_handle_t handle, void** stream);
int bbk_destroy(bbk_handle_t handle);
/* Block until all work queued on the handle's stream has finished. */
int bbk_synchronize(bbk_handle_t handle);
/* Text of the last error seen by this handle (never NULL). */
const char* bbk_last_error(bbk_handle_t handle);

/* Failure reporting (no counterpart in the reference, which reports only
 * argument errors, by exception).
 * The cooperative kernels (panel exchanges, triangular sweeps) wait for
 * records written by co-resident workgroups of the same launch.  The waits
 * of a launch are bounded by a deadline on the device clock; a wait that
 * runs past it marks an error word on the device and the launch runs to
 * its end with invalid values instead of hanging.  bbk_synchronize and
 * every synchronous entry point then return BBK_ERR_HIP (bbk_last_error
 * says which wait) and clear the word.
 * bbk_set_wait_limit: milliseconds per launch; 0 = the default (10 s, far
 * beyond any real wait); a negative value makes every wait fail at once
 * (for testing the error path). */
int bbk_set_wait_limit(bbk_handle_t handle, int ms);
/* Panel kernel variant of the handle's factorizations (no counterpart in
 * the reference, whose FactorPanel is serial): 0 (default) picks by order,
 * workgroups of 256 columns up to n = 65536 and of 512 columns up to
 * n = 131072 (binary64 panels then at most 64 wide); 256 or 512 forces that
 * variant (orders its grid cannot hold fail with BBK_ERR_UNSUPPORTED).
 * Both variants make the same pivot decisions and, at the same panel
 * width, bitwise the same factors. */
int bbk_set_panel_threads(bbk_handle_t handle, int threads);
/* After the last factorization queued on the handle (waits for the
 * stream): info = k + 1 for the first column k whose 1x1 pivot had
 * |d| <= DBL_EPSILON (its L column is set to zero, as the reference's
 * SafeDenominator does, src/block_ldlt_decomposition.cc:21-26) or whose 2x2
 * block had a zero determinant, 0 when there is none (LAPACK's dsytrf
 * info convention); nonfinite = 1 when D holds a NaN or an infinity (a
 * non-finite input reaches D).  Either pointer may be NULL.  Returns BBK_ERR_HIP as
 * bbk_synchronize does when a wait error is pending. */
int bbk_get_info(bbk_handle_t handle, int* info, int* nonfinite);

/* ---- plain factorization (device pointers, asynchronous) -------------
 * Replaces BlockLdltDecomposition::Factorize()
 * (include/block_ldlt_decomposition.h:22, src/block_ldlt_decomposition.cc:50-59)
 * for kCompactOnly storage.  nb is the panel width (0 = tuned default);
 * results are the same Bunch-Kaufman decisions for any nb (a 2x2 that
 * straddles a panel edge widens that panel by one column instead of
 * being deferred).  d_A: n x lda; d_perm, d_piv: n ints. */
int bbk_dsytrf(bbk_handle_t handle, int n, int nb, double* d_A, int64_t lda,
               int* d_perm, int* d_piv);
/* FP32 form: same algorithm on binary32 data with binary32 MFMA
 * accumulation; pivot tests and 2x2 inverses in binary64. */
int bbk_ssytrf(bbk_handle_t handle, int n, int nb, float* d_A, int64_t lda,
               int* d_perm, int* d_piv);

/* Host-buffer convenience (synchronous): copies A (row-major, lda, upper
 * triangle) to the device, factors it, and copies the compact factors,
 * perm and piv back into the same host arrays.  This is the entry the C++
 * drop-in layer uses. */
int bbk_dsytrf_host(bbk_handle_t handle, int n, int nb, double* A,
                    int64_t lda, int* perm, int* piv);

/* ---- out-of-core (non-GPU-resident) factorization ----------------------
 * The reference's Factorize() (src/block_ldlt_decomposition.cc:50-59) for
 * a matrix that stays in host memory; the reference cites this setting
 * (README.md:24-25) but is in-core only (README.md:155-157).  A: row-major,
 * lda, upper triangle, inside ONE pinned allocation from bbk_host_alloc
 * that covers (n - 1) lda + n elements (other memory returns -4); it is
 * factored in place, perm and piv are host arrays of n ints.  device_bytes
 * bounds the HBM the call allocates for the matrix's working copies (0:
 * the free HBM less 1 GiB); the handle's own scratch, O(nb n), is extra.
 * While the trailing matrix is larger than that, the panel runs on the
 * host copy across PCIe and the trailing update streams column slabs
 * through HBM (copies in and out on two streams of the call's own,
 * overlapped with the update); once it fits, it moves to HBM for the
 * remaining panels.  The interchanges of the rows factored on the host
 * copy are applied at the end, one gather per row.  The factors, perm and
 * piv are bitwise those of bbk_dsytrf; the strictly lower triangle is left
 * as it was.  Synchronous.  Returns BBK_OK, -i for an invalid argument i,
 * BBK_ERR_UNSUPPORTED (order beyond the panel grid, or device_bytes below
 * (nb + 1) ldd + 128 n elements, ldd = n rounded up to a multiple of 64
 * (plus 64 when that is a multiple of 1024), each part rounded up to 256
 * bytes), BBK_ERR_ALLOC or BBK_ERR_HIP. */
int bbk_dsytrf_ooc(bbk_handle_t handle, int n, int nb, double* A, int64_t lda, int* perm,
                   int* piv, size_t device_bytes);
/* Pinned host memory for bbk_dsytrf_ooc (hipHostMalloc: mapped into every
 * GPU's address space); bbk_host_free releases it. */
int bbk_host_alloc(void** p, size_t bytes);
int bbk_host_free(void* p);

/* ---- factorization that allocates nothing -----------------------------
 * No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): bbk_dsytrf's factorization (same
 * kernels, same results) with its scratch in a fixed device array of
 * 16.1 MiB that every GPU the library runs on holds until the process
 * ends; calls no hipMalloc/hipFree and needs no handle.  n <= 16384
 * (larger orders return BBK_ERR_UNSUPPORTED).  stream: a non-null stream
 * that does not synchronize with the null stream (cooperative launches);
 * the call is asynchronous on it.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU race and must be ordered by the
 * caller.  Returns BBK_OK, -i for an invalid argument i (1-based: n = 1),
 * BBK_ERR_UNSUPPORTED or BBK_ERR_HIP. */
int bbk_dsytrf_noalloc(int n, int nb, double* d_A, int64_t lda, int* d_perm,
                       int* d_piv, void* stream);

/* ---- solves (new; no reference counterpart) -------------------------
 * Solve A X = B with compact factors from bbk_?sytrf.  B: n x nrhs,
 * row-major with leading dimension ldb >= nrhs; overwritten by X. */
int bbk_dsytrs(bbk_handle_t handle, int n, int nrhs, const double* d_F,
               int64_t lda, const int* d_perm, const int* d_piv, double* d_B,
               int64_t ldb);
int bbk_ssytrs(bbk_handle_t handle, int n, int nrhs, const float* d_F,
               int64_t lda, const int* d_perm, const int* d_piv, float* d_B,
               int64_t ldb);

/* Normwise backward error per right-hand side, computed in binary64:
 * eta_q = ||b_q - A x_q||_inf / (||A||_inf ||x_q||_inf + ||b_q||_inf).
 * A: symmetric, upper triangle of d_A.  Writes nrhs doubles to d_eta. */
int bbk_dsyberr(bbk_handle_t handle, int n, int nrhs, const double* d_A,
                int64_t lda, const double* d_B, int64_t ldb,
                const double* d_X, int64_t ldx, double* d_eta);

/* Mixed-precision solve: factor a binary32 copy of A, solve in binary32,
 * refine in binary64 (residual b - A x and the correction accumulation)
 * until max_q eta_q <= tol or max_iter corrections.  d_A is not modified.
 * Synchronous.  iters/berr (host pointers, may be NULL) receive the number
 * of corrections applied and the final max backward error.  Returns
 * BBK_ERR_NOT_CONVERGED if tol was not reached (X still holds the best
 * iterate). */
int bbk_dsysv_mixed(bbk_handle_t handle, int n, int nrhs, const double* d_A,
                    int64_t lda, const double* d_B, int64_t ldb, double* d_X,
                    int64_t ldx, double tol, int max_iter, int* iters,
                    double* berr);
/* Plain binary64 solve for comparison: factor a copy, solve, report the
 * backward error.  Synchronous. */
int bbk_dsysv(bbk_handle_t handle, int n, int nrhs, const double* d_A,
              int64_t lda, const double* d_B, int64_t ldb, double* d_X,
              int64_t ldx, double* berr);

/* ---- batched factorization (new; no reference counterpart) ----------
 * Factor count independent matrices of uneven sizes in one persistent
 * launch: every workgroup takes matrices from a device work queue ordered
 * largest first and factors each one by itself (same Bunch-Kaufman
 * decisions and output encoding as bbk_?sytrf).  h_n: host array of
 * orders; d_A: device array of device pointers (row-major, upper
 * triangle); d_lda: device array of leading dimensions; d_perm/d_piv:
 * device arrays of device pointers (n ints each).  The panel width of
 * each matrix comes from the handle's size-class table.
 * When h_n and the table are those of the last call (or of
 * bbk_batched_prepare), the call only queues two kernels on the handle's
 * stream (no host-device copy, no allocation) and may be captured into a
 * HIP graph; otherwise it first uploads the schedule (synchronous). */
int bbk_dsytrf_batched(bbk_handle_t handle, int count, const int* h_n,
                       double* const* d_A, const int64_t* d_lda,
                       int* const* d_perm, int* const* d_piv);
int bbk_ssytrf_batched(bbk_handle_t handle, int count, const int* h_n,
                       float* const* d_A, const int64_t* d_lda,
                       int* const* d_perm, int* const* d_piv);
/* Upload the schedule of a batch (sizes on the host) for elem_bytes 8
 * (binary64) or 4 (binary32).  Synchronous. */
int bbk_batched_prepare(bbk_handle_t handle, int count, const int* h_n,
                        int elem_bytes);

/* Size classes n <= 32, 64, 128, 256, 512, 1024, 2048 and larger; one
 * panel width (1..64) per class and precision. */
#define BBK_BATCH_CLASSES 8
int bbk_batched_get_widths(bbk_handle_t handle, int elem_bytes, int* widths);
int bbk_batched_set_widths(bbk_handle_t handle, int elem_bytes,
                           const int* widths);
/* Team per class: 64 or 256 threads per matrix (the global-memory
 * kernel), or for classes n <= 256 0: the resident kernel, which keeps the
 * matrix (packed upper triangle), its panel rows and permutation in LDS
 * with one 256-thread workgroup per matrix, or 1: the symmetric-workspace
 * kernel (one wavefront per matrix, a full symmetric copy of the matrix in
 * the handle's workspace, interchanges as index-map updates; panel widths
 * up to 32, wider ones fail with BBK_ERR_UNSUPPORTED).  A batch runs with
 * the mode of the class of its largest matrix.  All modes give the same
 * factors (bit for bit with MFMA update tiles). */
int bbk_batched_get_team_sizes(bbk_handle_t handle, int elem_bytes,
                               int* team_sizes);
int bbk_batched_set_team_sizes(bbk_handle_t handle, int elem_bytes,
                               const int* team_sizes);
/* Trailing-update tile per class: 0 = SIMT 4 x 4 register tiles, 16 or 32
 * = MFMA tiles of ts x ts outputs per wavefront (the resident kernel
 * always uses MFMA; 0 counts as 16 there). */
int bbk_batched_get_tiles(bbk_handle_t handle, int elem_bytes, int* tiles);
int bbk_batched_set_tiles(bbk_handle_t handle, int elem_bytes, const int* tiles);
/* Cap on the number of persistent workgroups of the batched kernels
 * (factorization and refinement); 0 (the default) = as many as fit the
 * GPU.  Each workgroup takes matrices from the work queue until it is
 * empty, so a cap below the batch size makes workgroups factor several
 * matrices in turn (used by the tests to exercise that path; also a way to
 * leave CUs to other work). */
int bbk_batched_set_grid_limit(bbk_handle_t handle, int max_workgroups);
/* Autotuner: for every size class up to max_n, times candidates on a
 * synthetic batch of that class and keeps the fastest in the handle's
 * table (widths written to widths_out when not NULL): for classes n <= 256
 * every (panel width, team mode 64 / 256 / resident / symmetric
 * workspace, update tile SIMT /
 * MFMA 16 / MFMA 32) combination, above that (panel width, team size).
 * Synchronous. */
int bbk_batched_tune(bbk_handle_t handle, int elem_bytes, int max_n,
                     int* widths_out);

/* HIP graph of one batched factorization (the queue reset and the
 * persistent kernel), captured on the handle's stream for fixed
 * arguments; bbk_graph_launch replays it on the same stream.  The graph
 * owns a private copy of the batch schedule and of the W slices it reads,
 * so later batched calls, bbk_batched_tune or re-planning on the same
 * handle do not affect its replays.  The device arrays passed at creation
 * (d_A, d_lda, d_perm, d_piv and the matrices they point to) must stay
 * valid while the graph lives; bbk_graph_destroy waits for queued replays
 * before freeing the graph's copy.  Replays of one graph, and a replay and
 * a direct batched call on the same matrices, must not overlap. */
typedef struct bbk_graph* bbk_graph_t;
int bbk_graph_create_dsytrf_batched(bbk_handle_t handle, int count,
                                    const int* h_n, double* const* d_A,
                                    const int64_t* d_lda, int* const* d_perm,
                                    int* const* d_piv, bbk_graph_t* graph);
int bbk_graph_create_ssytrf_batched(bbk_handle_t handle, int count,
                                    const int* h_n, float* const* d_A,
                                    const int64_t* d_lda, int* const* d_perm,
                                    int* const* d_piv, bbk_graph_t* graph);
int bbk_graph_launch(bbk_graph_t graph);
int bbk_graph_destroy(bbk_graph_t graph);

/* Batched mixed-precision solve (new): every matrix m is factored in
 * binary32 (bbk_ssytrf_batched on a binary32 copy of its upper triangle),
 * then X_m is refined in binary64 until its normwise backward error
 * max_q ||b_q - A x_q||_inf / (||A||_inf ||x_q||_inf + ||b_q||_inf) <= tol
 * or max_iter corrections were applied.  One persistent workgroup per
 * matrix runs the whole refinement loop.  nrhs: 1..4 right-hand sides per
 * matrix, row-major n x nrhs (leading dimensions ldb, ldx >= nrhs).  d_A,
 * d_B, d_X: device arrays of device pointers; d_lda, d_ldb, d_ldx: device
 * arrays; d_iters / d_berr: device arrays of count entries receiving the
 * corrections applied and the final backward error.  n * nrhs must be at
 * most 9600 (the right-hand sides live in LDS).  Synchronous; returns
 * BBK_ERR_NOT_CONVERGED when some matrix stayed above tol.  The loop of a
 * matrix stops at tol, after max_iter corrections or at a non-finite
 * error; X_m then holds the last iterate and d_berr[m] its backward error
 * (an order-0 matrix gets d_iters[m] = 0, d_berr[m] = 0). */
int bbk_dsysv_mixed_batched(bbk_handle_t handle, int count, const int* h_n,
                            int nrhs, const double* const* d_A,
                            const int64_t* d_lda, const double* const* d_B,
                            const int64_t* d_ldb, double* const* d_X,
                            const int64_t* d_ldx, double tol, int max_iter,
                            int* d_iters, double* d_berr);

/* Batched synthetic inputs: matrix m gets bbk_dgen_synth's formula with
 * seed + m (binary32 form: the same values rounded to binary32).  d_n:
 * device array of orders. */
int bbk_dgen_synth_batched(bbk_handle_t handle, int count, const int* d_n,
                           uint64_t seed, double scale, int zero_diag_stride,
                           double* const* d_A, const int64_t* d_lda);
int bbk_sgen_synth_batched(bbk_handle_t handle, int count, const int* d_n,
                           uint64_t seed, double scale, int zero_diag_stride,
                           float* const* d_A, const int64_t* d_lda);

/* ---- kernels.h counterparts (device pointers, asynchronous) ----------
 * Element strides, no return value semantics of the reference become
 * status codes.  Same arithmetic as src/kernels.cc:11-91 (non-MKL path). */
/* kernels::ScaleOneByOnePivot, include/kernels.h:6-8 */
int bbk_kernel_scale_1x1(bbk_handle_t handle, int count, double pivot,
                         const double* d_src, int src_stride, double* d_dst,
                         int dst_stride);
/* kernels::ScaleTwoByTwoPivot, include/kernels.h:10-17 */
int bbk_kernel_scale_2x2(bbk_handle_t handle, int count, double a, double b,
                         double c, const double* d_s1, int s1_stride,
                         const double* d_s2, int s2_stride, double* d_d1,
                         int d1_stride, double* d_d2, int d2_stride);
/* kernels::UpdateTrailingBlock, include/kernels.h:19-23:
 * target(i,j) -= sum_k left(k,i) * work(k,j) on MFMA. */
int bbk_kernel_update_trailing(bbk_handle_t handle, int rows, int cols,
                               int inner, const double* d_left,
                               int left_stride, const double* d_work,
                               int work_stride, double* d_target,
                               int target_stride);

/* Host-pointer forms of the three helpers (synchronous): operands are
 * copied to the device, the kernel runs, results are copied back.  This
 * is what the C++ drop-in naive_block_ldlt::kernels::* calls.  Strides
 * must be >= 0. */
int bbk_kernel_scale_1x1_host(bbk_handle_t handle, int count, double pivot,
                              const double* src, int src_stride, double* dst,
                              int dst_stride);
int bbk_kernel_scale_2x2_host(bbk_handle_t handle, int count, double a,
                              double b, double c, const double* s1,
                              int s1_stride, const double* s2, int s2_stride,
                              double* d1, int d1_stride, double* d2,
                              int d2_stride);
int bbk_kernel_update_trailing_host(bbk_handle_t handle, int rows, int cols,
                                    int inner, const double* left,
                                    int left_stride, const double* work,
                                    int work_stride, double* target,
                                    int target_stride);

/* ---- reconstruction residuals of compact factors ----------------------
 * A: the matrix that was factored, its upper triangle read as the
 * symmetric matrix it defines (row-major, lda); F, perm, piv: the output of
 * bbk_dsytrf.  Both return the residual in *residual (host) after the
 * handle's stream has finished them.
 *
 * bbk_dresidual: ||P L D L^T P^T - A||_F / max(||A||_F, 1), the relative
 * residual of ReconstructOriginal() against the input
 * (src/block_ldlt_decomposition.cc:61-78; tests/block_ldlt_decomposition_
 * test.cc:131-135, RelativeError), formed on the device with FP64 MFMA
 * from F without building L, D or the product (n^3/3 flops).
 *
 * bbk_dsampled_residual: SampledRelativeResidual(original, decomposition,
 * seed) of tests/block_ldlt_large_stress_test.cc:196-244: the same `count`
 * sample positions (the four fixed ones, then std::mt19937(seed + 170003)
 * pairs; the reference uses count = 4096), each entry reconstructed as
 * ReconstructedPermutedEntry does. */
/* bbk_dreconstruct: out = P L D L^T P^T (n x n, row-major, ldo), the matrix
 * ReconstructOriginal() returns (src/block_ldlt_decomposition.cc:61-78),
 * formed on the device from F with FP64 MFMA (queued on the handle's
 * stream).  bbk_dreconstruct_host: the same from and into host buffers
 * (returns when `out` is filled). */
int bbk_dreconstruct(bbk_handle_t handle, int n, const double* d_F, int64_t ldf,
                     const int* d_perm, const int* d_piv, double* d_out, int64_t ldo);
int bbk_dreconstruct_host(bbk_handle_t handle, int n, const double* F, int64_t ldf,
                          const int* perm, const int* piv, double* out, int64_t ldo);
int bbk_dresidual(bbk_handle_t handle, int n, const double* d_A, int64_t lda,
                  const double* d_F, int64_t ldf, const int* d_perm, const int* d_piv,
                  double* residual);
int bbk_dsampled_residual(bbk_handle_t handle, int n, const double* d_A, int64_t lda,
                          const double* d_F, int64_t ldf, const int* d_perm,
                          const int* d_piv, int seed, int count, double* residual);
/* The sample positions bbk_dsampled_residual uses (ResidualSamples,
 * tests/block_ldlt_large_stress_test.cc:212-227), into host arrays of
 * `count` entries; host only, no device work. */
int bbk_residual_samples(int n, int seed, int count, int* rows, int* cols);

/* ---- LAPACK-convention interop ---------------------------------------
 * The compact factor in place, between this library's form (every
 * interchange applied to the earlier L columns too, as the reference does,
 * src/block_ldlt_decomposition.cc:392-587) and LAPACK dsytrf('L')'s form
 * (A = P(1) L(1) ... P(s) L(s) D (...)^T with 1-based ipiv: ipiv(k) > 0 a
 * 1x1 pivot with rows/columns k, ipiv(k) interchanged; ipiv(k) = ipiv(k+1)
 * < 0 a 2x2 pivot with k+1, -ipiv(k) interchanged).  Same bytes: row-major
 * upper = column-major lower.  d_ipiv: n ints on the device.  Both return
 * when the conversion is done.
 *
 * bbk_dsytrf_to_lapack: after bbk_dsytrf, F and (perm, piv) -> LAPACK's
 * factor and ipiv (LAPACK dsytrs then solves with it).
 * bbk_dsytrf_from_lapack: LAPACK's factor and ipiv -> this library's F,
 * perm and piv (bbk_dsytrs then solves with them). */
int bbk_dsytrf_to_lapack(bbk_handle_t handle, int n, double* d_F, int64_t ldf,
                         const int* d_perm, const int* d_piv, int* d_ipiv);
int bbk_dsytrf_from_lapack(bbk_handle_t handle, int n, double* d_F, int64_t ldf,
                           const int* d_ipiv, int* d_perm, int* d_piv);

/* ---- synthetic inputs (bench/test plumbing) ---------------------------
 * A(i,j) = A(j,i) = scale * u(seed,i,j), u uniform in [-1,1) from
 * SplitMix64 keyed by (seed, i, j); A(i,i) = 0 when zero_diag_stride > 0
 * and i % zero_diag_stride == 0.  Writes the full n x n (row-major, lda). */
int bbk_dgen_synth(bbk_handle_t handle, int n, uint64_t seed, double scale,
                   int zero_diag_stride, double* d_A, int64_t lda);

/* Kernel timing accumulated over every bbk_?sytrf call since timing was
 * last enabled (milliseconds, HIP events recorded on the handle's stream
 * around each kernel group): panel kernels, trailing updates, row
 * interchanges, and the number of panels.  Any pointer may be NULL.
 * Valid after bbk_synchronize. */
int bbk_last_timing(bbk_handle_t handle, double* panel_ms, double* update_ms,
                    double* swap_ms, int* panels);
/* Enable/disable per-kernel event timing (off by default); either call
 * clears the accumulated totals. */
int bbk_set_timing(bbk_handle_t handle, int enabled);

#ifdef __cplusplus
}
#endif

#endif /* BBK_LDLT_H_ */
