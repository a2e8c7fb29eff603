/*
 * tentorch_amd: the C ABI of the MI355X device library for TenTorch's block-sparse contraction path.
 *
 * The host engine (legs, fusion trees, leg-move and contraction planning) compiles every leg move
 * and every contraction into a descriptor table; this library uploads the tables once per plan and
 * runs them with hand-written HIP kernels for gfx950.  No torch types cross this boundary: device
 * buffers are plain pointers, streams are `hipStream_t` passed as `void *`, plans are opaque handles
 * owned by the library.  Every entry point returns 0 on success and a negative status on failure;
 * `tt_last_error` then holds the message (thread-local).  Nothing here throws.
 *
 * Reference interfaces each entry point replaces (paths relative to the reference checkout):
 *   tt_gss_*     the leg-move execution of `Tensor::place`: the per-block chain walk `carried`
 *                (tentorch/tensors/place.cpp:127-180) with the host pointer walk / per-block device
 *                matmul of `TensorCaches::apply` (tentorch/tensors/cache.cpp:66-116,297-319);
 *   tt_ggemm_*   the per-root block products `out.block(r, s).addmm_(L, R)` of `sym::contract`
 *                (tentorch/linalg/graded/contract.cpp:144-163);
 *   tt_dot       the reduced scalar product `inner` and `norm`
 *                (tentorch/linalg/graded/compare.cpp:31-45), with the scalar left on the device;
 *   tt_axpby / tt_scale   the per-block `add_/sub_/mul_/div_` of `Tensor` arithmetic
 *                (tentorch/tensors/tensor.cpp:138-181) on a whole flat tensor at once;
 *   tt_graph_*   no counterpart: capture of a launch sequence (the Lanczos matvec) into a hipGraph.
 */
#ifndef TENTORCH_AMD_H
#define TENTORCH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element types of device buffers. */
enum tt_dtype { TT_F64 = 0, TT_F32 = 1, TT_C128 = 2 };

/* Largest rank of a gather-scale-scatter box (tensors of up to eight legs). */
#define TT_MAX_DIMS 8

/*
 * One target tile of a leg move: a dense box of the target buffer, `extent[d]` wide along axis d
 * and `dst_stride[d]` apart (elements), starting at `dst_offset`.  Its value is the sum, in source
 * order, of `coef * source box` over the `num_sources` sources starting at `first_source`; a tile
 * with no source is written with zeros.  Axes are listed with the destination's slowest first.
 */
typedef struct tt_gss_task {
  int64_t dst_offset;
  int32_t ndim;
  int32_t first_source;
  int32_t num_sources;
  int32_t group;          /* launch group of the per-block (plain) form */
  int64_t extent[TT_MAX_DIMS];
  int64_t dst_stride[TT_MAX_DIMS];
} tt_gss_task;

/* One source box of a tile: same extents as its task, its own offset and strides. */
typedef struct tt_gss_source {
  int64_t src_offset;
  int64_t src_stride[TT_MAX_DIMS];
  double coef_re;
  double coef_im;
} tt_gss_source;

typedef struct tt_gss_plan tt_gss_plan;

int tt_gss_plan_create(const tt_gss_task *tasks, int64_t num_tasks, const tt_gss_source *sources,
                       int64_t num_sources, tt_gss_plan **plan);
int tt_gss_plan_destroy(tt_gss_plan *plan);
/* Number of destination elements the plan writes and of source elements it reads. */
int tt_gss_plan_counts(const tt_gss_plan *plan, int64_t *written, int64_t *read);
/* All tiles in one launch.  `src` and `dst` must not overlap. */
int tt_gss_execute(const tt_gss_plan *plan, int dtype, const void *src, void *dst, void *stream);
/* The per-block form: one launch per launch group (one source block into one target block). */
int tt_gss_execute_plain(const tt_gss_plan *plan, int dtype, const void *src, void *dst, void *stream);
/* All tiles in one launch, converting between element types on the way (TT_F64 <-> TT_F32, or the
 * same type): the FP32 form stages FP64 operands into FP32 workspaces and back. */
int tt_gss_execute_typed(const tt_gss_plan *plan, int src_type, int dst_type, const void *src, void *dst,
                         void *stream);

/*
 * One output block of a grouped GEMM: C[c_offset : m x n, row-major, ldc = n] =
 * sum over its terms of A_t (m x k_t, row-major, lda = k_t) times B_t (k_t x n, row-major, ldb = n),
 * the terms added in list order; a block with no term is written with zeros.
 */
typedef struct tt_gemm_group {
  int64_t c_offset;
  int64_t m;
  int64_t n;
  int32_t first_term;
  int32_t num_terms;
} tt_gemm_group;

typedef struct tt_gemm_term {
  int64_t a_offset;
  int64_t b_offset;
  int64_t k;
} tt_gemm_term;

typedef struct tt_ggemm_plan tt_ggemm_plan;

/*
 * Limits: every block side and depth below 2^20, and every operand block (m k or k n elements) below
 * 2^27 elements, so that a tile's byte offsets and its operand ranges stay in 32 bits; larger blocks
 * are refused with an error.  A plan whose launch has few tiles splits its deep products (depth at
 * least 512 and at least twice max(m, n, 128)) into at most eight depth parts; `tt_ggemm_execute`
 * then runs a second launch that adds the parts' partial sums in part order, so the result still
 * does not depend on the configuration.  The partial sums live in a buffer the library keeps per
 * device, allocated when a plan is created (never while a stream is being captured, if the plan
 * was created before) and kept until the process ends; launches of split plans on one device share
 * it and must not overlap (one stream at a time, as everywhere in this library).
 */
int tt_ggemm_plan_create(const tt_gemm_group *groups, int64_t num_groups, const tt_gemm_term *terms,
                         int64_t num_terms, tt_ggemm_plan **plan);
/* The same with an explicit tile shape (0 <= config < tt_ggemm_num_configs(); a negative config lets
 * the library choose from the block shapes, as tt_ggemm_plan_create does).  For an autotuner. */
int tt_ggemm_plan_create_config(const tt_gemm_group *groups, int64_t num_groups, const tt_gemm_term *terms,
                                int64_t num_terms, int config, tt_ggemm_plan **plan);
/* Tile shapes the library is built with: rows x cols of the output tile, depth of one k-slab. */
/* Configurations: c % 7 picks the tile shape (BM, BN, BK), c / 7 the placement of tiles in the
 * launch: 0 longest K first; 1 XCD bands: each block product's tiles, row-major, cut into eight
 * contiguous bands, one per XCD (workgroup i runs on XCD i % 8), each XCD's tiles longest K first. */
int tt_ggemm_num_configs(void);
int tt_ggemm_config_shape(int config, int *rows, int *cols, int *depth);
/* The tile shape a plan uses. */
int tt_ggemm_plan_config(const tt_ggemm_plan *plan);
int tt_ggemm_plan_destroy(tt_ggemm_plan *plan);
/* Real multiply-add count of the plan (2 m n k summed over terms), and tile count of one launch. */
int tt_ggemm_plan_counts(const tt_ggemm_plan *plan, double *flops, int64_t *tiles);
/*
 * One launch for every output block (two when the plan splits deep products: the parts, then the
 * sum of their partial sums).  `a_type`, `b_type`, `c_type` are the element types of the three
 * buffers (TT_F64 or TT_F32); `compute_type` TT_F64 runs FP64 MFMA with FP64 accumulation, TT_F32
 * runs FP32 MFMA with FP32 accumulation (operands rounded to FP32 when staged).
 */
int tt_ggemm_execute(const tt_ggemm_plan *plan, int a_type, int b_type, int c_type, int compute_type,
                     const void *a, const void *b, void *c, void *stream);
/* The per-block form: one launch per (output block, term), the first term overwriting. */
int tt_ggemm_execute_plain(const tt_ggemm_plan *plan, int dtype, const void *a, const void *b, void *c,
                           void *stream);

/* result[0] = sum_i conj(x_i) * y_i over n elements (TT_F64 or TT_F32 inputs; the result is one
 * double, or two for TT_C128), summed in a fixed order, left in device memory. */
int tt_dot(int dtype, int64_t n, const void *x, const void *y, double *result, void *stream);
/* y = alpha * x + beta * y, with alpha and beta read from device memory when the pointers are
 * non-null (`alpha_dev`, `beta_dev`, one double each) and from the host values otherwise. */
int tt_axpby(int dtype, int64_t n, double alpha, const double *alpha_dev, const void *x, double beta,
             const double *beta_dev, void *y, void *stream);
/* y = x converted between TT_F64 and TT_F32. */
int tt_convert(int from_type, int to_type, int64_t n, const void *x, void *y, void *stream);
/* Krylov basis (FP64, rows of length n, `ldv` apart): out[i] = sum_j V[i][j] w[j] for i < k <= 32,
 * each in a fixed order, left on the device.  Replaces the k separate `inner` calls of one
 * reorthogonalisation pass (tentorch/linalg/lanczos.py:43-45). */
int tt_multidot(int64_t n, int k, const double *V, int64_t ldv, const double *w, double *out, void *stream);
/* w += scale * sum_{i < k} coef[i] V[i][:], coefficients read from device memory. */
int tt_multiaxpy(int64_t n, int k, const double *coef, double scale, const double *V, int64_t ldv, double *w,
                 void *stream);
/* The same over the first min(k, *count) rows, `count` read from device memory (a Lanczos basis past
 * an early stop may hold non-finite rows, which are then never read). */
int tt_multiaxpy_counted(int64_t n, int k, const int32_t *count, const double *coef, double scale, const double *V,
                         int64_t ldv, double *w, void *stream);
/* The Ritz pair of a Lanczos run from its device scalars, on the device (the host `eigh` of the
 * tridiagonal and the coefficient extraction of tentorch/linalg/lanczos.py:52-62): size = the first
 * j + 1 with beta[j] < tol, else k (k <= 32); the tridiagonal of alpha[0..size) and
 * beta[0..size-1) diagonalised by implicit QL; *ritz = its least eigenvalue, coef[0..k) its
 * eigenvector (zero past size, first coefficient non-negative), *size = size.  One launch, no host
 * sync. */
int tt_lanczos_ritz(int k, const double *alpha, const double *beta, double tol, double *coef, double *ritz,
                    int32_t *size, void *stream);
/* One Krylov step's vector work after the action (tentorch/linalg/lanczos.py:36-50): w, the image of
 * row k of V, is orthogonalised against rows 0..k of V by two classical Gram-Schmidt passes, each
 * subtraction fused with what follows it (the second projection, then the norm); *alpha = the two
 * passes' projections on row k, *beta = the norm left, and row k+1 of V = w / beta.  V is row-major
 * with leading dimension ldv and at least k+2 rows; k + 1 <= 32.  `coef` holds 2 (k + 1) + 1
 * doubles of device scratch; alpha and beta are device doubles.  Seven launches, no host sync. */
int tt_lanczos_orthonormalize(int64_t n, int k, double *V, int64_t ldv, double *w, double *coef, double *alpha,
                              double *beta, void *stream);

typedef struct tt_graph tt_graph;
/* Begin capturing every launch on `stream` into a graph. */
int tt_graph_begin(void *stream);
/* End the capture and instantiate it. */
int tt_graph_end(void *stream, tt_graph **graph);
int tt_graph_launch(const tt_graph *graph, void *stream);
int tt_graph_destroy(tt_graph *graph);

/* Table uploads in one copy.  No counterpart in the reference (its per-block calls upload
 * nothing).  Plans created between tt_upload_batch_begin and tt_upload_batch_end (on the current
 * device; batches nest) have their device tables copied together when the outermost batch ends,
 * one copy and one event instead of one per plan; a launch inside a batch issues what is staged
 * first.  For a caller that creates several plans at once (an einsum's moves and products). */
int tt_upload_batch_begin(void);
int tt_upload_batch_end(void);

/* Batched symmetric eigendecomposition in FP64 (the split of the two-site tensor: the Gram matrix
 * of each block).  Replaces the per-block `at::linalg_svd` of `split_over`
 * (tentorch/linalg/graded/svd.cpp:41-177) with one eigendecomposition of all blocks at once:
 * tridiagonalisation (one launch per Householder step while a trailing matrix exceeds order 192,
 * the last 192 steps in one launch with each trailing matrix in a workgroup's registers),
 * divide-and-conquer leaves and merge levels, blocked back-transformation; stream-ordered, no host
 * synchronisation, no allocation.  Matrix b is the symmetric n x n at
 * `matrices + matrix_offset` (overwritten: it keeps the Householder reflectors); its eigenvalues go
 * to `values + value_offset` in ascending order, and eigenvector j to row j of the n x n at
 * `vectors + matrix_offset`.  n is 1..2048.  `workspace` holds tt_eigh_workspace(...) doubles. */
typedef struct tt_eigh_problem {
  int64_t matrix_offset;
  int64_t value_offset;
  int64_t n;
} tt_eigh_problem;
/* The largest matrix order tt_eigh_batched takes (2048). */
int64_t tt_eigh_max_order(void);
int64_t tt_eigh_workspace(const tt_eigh_problem *problems, int64_t count);
int tt_eigh_batched(const tt_eigh_problem *problems, int64_t count, double *matrices, double *values,
                    double *vectors, double *workspace, void *stream);
/* No counterpart in the reference: tt_eigh_batched for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The same
 * kernels, with the workspace taken from a static array of TT_EIGH_STATIC_SCRATCH_DOUBLES doubles
 * (64 MiB) that the library holds on every device it runs on until the process ends; nothing is
 * allocated or freed.  A batch whose tt_eigh_workspace exceeds that size is refused with an error.
 * The scratch belongs to the current device (the stream must be on it), keeps the last call's
 * contents (every call rewrites what it reads before reading it), and is shared by every call on
 * that device: calls that overlap on one device race. */
#define TT_EIGH_STATIC_SCRATCH_DOUBLES (INT64_C(8) << 20)
int tt_eigh_batched_static(const tt_eigh_problem *problems, int64_t count, double *matrices, double *values,
                           double *vectors, void *stream);

/* Batched kernels of the graded split (`split_over`, tentorch/linalg/graded/svd.cpp:41-177: the
 * per-block `at::linalg_svd` and the factor copies into the bond, svd.cpp:59,159-172).  Each takes a
 * problem table in device memory, `count` problems sorted by `first`, whose work items are numbered
 * consecutively across problems (`first` is the number of the problem's first item).
 *
 * tt_bgemm: for each problem C = alpha op(A) op(B) diag(col_scale) + beta C (col_scale null: 1),
 * or, when row_shift is set, C_ij = alpha (op(A) op(B))_ij / (col_scale[j] - row_shift[i]) + beta C_ij
 * (the solution of a Sylvester equation with diagonal coefficients; col_scale is then required),
 * op(A) m x k, op(B) k x n, all row-major with row strides lda, ldb, ldc; trans_a / trans_b set
 * means the operand is stored transposed (A as k x m, B as n x k).  FP64 on MFMA, one work item per
 * 64 x 64 tile of C (ceil(m / 64) ceil(n / 64) per problem, row-major); `tiles` is their total.
 * With beta 0, C is not read. */
typedef struct tt_bgemm_problem {
  const double *a;
  const double *b;
  double *c;
  const double *col_scale;
  const double *row_shift;
  int64_t first;
  int32_t m, n, k;
  int32_t lda, ldb, ldc;
  int32_t trans_a, trans_b;
  double alpha;
  double beta;
} tt_bgemm_problem;
int tt_bgemm(const tt_bgemm_problem *problems, int64_t count, int64_t tiles, void *stream);
/* One work item per row.  tt_row_norms: dst[i] = the Euclidean norm of row i of src (`cols`
 * entries, row stride src_ld).  tt_copy_rows: row i of the problem is row index[i] of src (row i
 * when index is null), written as row i of dst (transpose 0, stride dst_ld) or as its column i
 * (transpose 1).  `rows` is the total over the problems. */
typedef struct tt_rows_problem {
  const double *src;
  double *dst;
  const int32_t *index;
  int64_t first;
  int32_t rows, cols;
  int32_t src_ld, dst_ld;
  int32_t transpose;
  int32_t reserved;
} tt_rows_problem;
int tt_row_norms(const tt_rows_problem *problems, int64_t count, int64_t rows, void *stream);
int tt_copy_rows(const tt_rows_problem *problems, int64_t count, int64_t rows, void *stream);
/* One work item per row: dst = alpha x + beta y + gamma I over `cols` entries of each row (y null:
 * no y term; I the identity, entry (i, i) of the problem's rows), row strides x_ld, y_ld, dst_ld.
 * dst may be x or y. */
typedef struct tt_lincomb_problem {
  const double *x;
  const double *y;
  double *dst;
  int64_t first;
  int32_t rows, cols;
  int32_t x_ld, y_ld, dst_ld;
  int32_t reserved;
  double alpha, beta, gamma;
} tt_lincomb_problem;
int tt_lincomb_rows(const tt_lincomb_problem *problems, int64_t count, int64_t rows, void *stream);
/* Batched Householder QR in FP64, replacing the per-block `at::linalg_qr` of `qr_over` and
 * `rq_over` (tentorch/linalg/graded/svd.cpp:225-249).  Problem i factors A (m x n) = its source
 * (row-major, row stride src_ld) or, when `trans` is set, the source's transpose (stored n x m), as
 * A = Q R with k = min(m, n) reflectors in LAPACK's convention (geqr2 / org2r: beta = -sign(alpha)
 * times the column norm, a column with nothing below its diagonal left as it is), so the signs of
 * Q's columns and R's rows are LAPACK's.  Q (m x k) goes to q with row stride q_ld, or as its
 * transpose (k x m) when q_trans; R (k x n, zeros below the diagonal) to r with row stride r_ld, or
 * transposed (n x k) when r_trans.  `work` holds m n + k doubles for the problem (column-major
 * matrix and reflector factors); the outputs must not overlap the source.  One workgroup per
 * problem.  tt_qr_check validates a table against device byte ranges as tt_bgemm_check does. */
typedef struct tt_qr_problem {
  const double *src;
  double *q;
  double *r;
  double *work;
  int32_t m, n;
  int32_t src_ld, q_ld, r_ld;
  int32_t trans, q_trans, r_trans;
} tt_qr_problem;
int tt_qr_batched(const tt_qr_problem *problems, int64_t count, void *stream);
int tt_qr_check(const tt_qr_problem *problems, int64_t count, const int64_t *starts, const int64_t *ends,
                int64_t spans);
/* Host-side checks of the three table kinds above before they are uploaded (no counterpart in the
 * reference, whose per-block calls take sized tensors): every operand a problem names must lie in
 * one of the `spans` device byte ranges [starts[i], ends[i]) (sorted by start, not overlapping),
 * sizes must be non-negative, row strides no shorter than their rows, and the work items numbered
 * consecutively from 0 in `first`.  On success `items` receives the launch's work-item count (tiles
 * or rows).  tt_rows_check: with an index, source_rows[i] is the number of rows problem i's source
 * holds (the caller keeps every index below it); vector_dst: dst takes one value per row. */
int tt_bgemm_check(const tt_bgemm_problem *problems, int64_t count, const int64_t *starts, const int64_t *ends,
                   int64_t spans, int64_t *items);
int tt_rows_check(const tt_rows_problem *problems, int64_t count, const int64_t *source_rows, int vector_dst,
                  const int64_t *starts, const int64_t *ends, int64_t spans, int64_t *items);
int tt_lincomb_check(const tt_lincomb_problem *problems, int64_t count, const int64_t *starts, const int64_t *ends,
                     int64_t spans, int64_t *items);

/* The learned tile schedule's cost model (no counterpart in the reference): a multilayer
 * perceptron x (features) -> relu(W1^T x + b1) (hidden) -> relu(W2^T h1 + b2) (hidden) -> w3 . h2 + b3
 * in FP32, parameters packed as W1 [features][hidden], b1, W2 [hidden][hidden], b2, w3, b3.
 * hidden must be 64 and features 1..64.  `x` is [batch][features] row-major.
 * tt_mlp_gradient writes the gradient of mean((y - target)^2) over the batch (fixed summation order)
 * and, when `loss` is not null, that mean; tt_mlp_adam takes one Adam step (step counts from 1). */
int64_t tt_mlp_param_count(int features, int hidden);
int tt_mlp_forward(int features, int hidden, int64_t batch, const float *x, const float *params, float *y,
                   void *stream);
int tt_mlp_gradient(int features, int hidden, int64_t batch, const float *x, const float *target,
                    const float *params, float *gradient, float *loss, void *stream);
int tt_mlp_adam(int64_t n, float *params, const float *gradient, float *m, float *v, float lr, float beta1,
                float beta2, float eps, int step, void *stream);

/* Kernel timing (no counterpart in the reference): while enabled on the calling thread, every
 * leg-move and block-product launch is issued with a start and a stop event attached to the kernel
 * dispatch itself (hipExtLaunchKernelGGL), so the interval is the kernel's execution alone.
 * `tt_timing_collect` sets `count` to the number recorded; given an output of at least that many,
 * it synchronises on them, writes the durations in launch order and forgets them (with a null output
 * it only counts). */
int tt_timing_enable(int on);
int tt_timing_collect(double *milliseconds, int64_t capacity, int64_t *count);

/* Copy the last error of this thread into `buffer` (NUL-terminated, truncated to `size`). */
int tt_last_error(char *buffer, size_t size);
/* Library version, for a loader to check. */
int tt_version(void);
/* Leg-move and block-product kernel launches issued by this process so far (diagnostics: the
 * per-block form's launch count against the grouped form's).  No counterpart in the reference. */
int64_t tt_launch_count(void);

#ifdef __cplusplus
}
#endif

#endif /* TENTORCH_AMD_H */
