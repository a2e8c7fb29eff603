// Batched kernels of the graded split (tentorch/linalg/graded/svd.cpp:41-177): the Gram products,
// projections and rotations of every block of a split in one launch each, the row norms the
// ranking reads, and the copies that assemble the factors.  Every launch takes a problem table in
// device memory (the caller uploads it), so its work does not depend on how many blocks there are.
//
// bgemm_kernel: C = alpha op(A) op(B) diag(s) + beta C for each problem (or, with row shifts r,
// C_ij = alpha (op(A) op(B))_ij / (s_j - r_i) + beta C_ij), FP64 MFMA
// (v_mfma_f64_16x16x4f64).  One workgroup per 64 x 64 tile of one problem; four waves, each a
// 32 x 32 quarter as 2 x 2 MFMA blocks.  The 16-deep slabs of A and B go through LDS (A as
// As[depth][row], B as Bs[depth][col]), which is also where a transposed operand is turned:
// each operand is loaded along its contiguous axis (a row of A or of B^T: 16 consecutive depths;
// a row of A^T or of B: 64 consecutive rows or columns), the next slab in registers while the MFMAs
// run on the current one.  The MFMA result layout: lane l holds rows 4 r + l / 16 of a 16 x 16
// block, column l % 16 (measured on gfx950; the eigensolver's merge products use the same).

#include <algorithm>
#include <cstdio>
#include <string>

#include "common.h"

using ttdev::as_stream;
using ttdev::fail;

namespace {

constexpr int kTile = 64;
constexpr int kSlab = 16;
constexpr int kThreads = 256;

typedef double d4 __attribute__((ext_vector_type(4)));

// The problem of a global tile index: the last problem whose first tile is at or before it.
template <typename Problem>
__device__ int find_problem(const Problem *__restrict__ problems, int count, int64_t tile) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (problems[mid].first <= tile) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(kThreads) void bgemm_kernel(const tt_bgemm_problem *__restrict__ problems, int count) {
  const int index = find_problem(problems, count, blockIdx.x);
  const tt_bgemm_problem p = problems[index];
  const int64_t local = int64_t(blockIdx.x) - p.first;
  const int col_tiles = (p.n + kTile - 1) / kTile;
  const int r0 = int(local / col_tiles) * kTile, c0 = int(local % col_tiles) * kTile;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  __shared__ double As[2][kSlab][kTile + 1], Bs[2][kSlab][kTile + 1];

  // staging: each thread brings four elements of each operand per slab
  double ra[4], rb[4];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int i, kk;
      if (p.trans_a) { i = t & 63; kk = (t >> 6) + 4 * q; }      // A^T rows: 64 consecutive i
      else { kk = t & 15; i = (t >> 4) + 16 * q; }              // A rows: 16 consecutive depths
      const int row = r0 + i, depth = k0 + kk;
      ra[q] = (row < p.m && depth < p.k)
                  ? (p.trans_a ? p.a[int64_t(depth) * p.lda + row] : p.a[int64_t(row) * p.lda + depth])
                  : 0.0;
      int j, kb;
      if (p.trans_b) { kb = t & 15; j = (t >> 4) + 16 * q; }      // B^T rows: 16 consecutive depths
      else { j = t & 63; kb = (t >> 6) + 4 * q; }               // B rows: 64 consecutive columns
      const int col = c0 + j, depth_b = k0 + kb;
      rb[q] = (col < p.n && depth_b < p.k)
                  ? (p.trans_b ? p.b[int64_t(col) * p.ldb + depth_b] : p.b[int64_t(depth_b) * p.ldb + col])
                  : 0.0;
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (p.trans_a) As[buf][(t >> 6) + 4 * q][t & 63] = ra[q];
      else As[buf][t & 15][(t >> 4) + 16 * q] = ra[q];
      if (p.trans_b) Bs[buf][t & 15][(t >> 4) + 16 * q] = rb[q];
      else Bs[buf][(t >> 6) + 4 * q][t & 63] = rb[q];
    }
  };

  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32, lr = lane & 15, lk = lane >> 4;
  d4 acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = d4{0.0, 0.0, 0.0, 0.0};
  int buf = 0;
  if (p.k > 0) fetch(0);
  for (int k0 = 0; k0 < p.k; k0 += kSlab) {
    stage(buf);
    __syncthreads();
    if (k0 + kSlab < p.k) fetch(k0 + kSlab);
#pragma unroll
    for (int k4 = 0; k4 < kSlab; k4 += 4) {
      double a[2], bv[2];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi) a[bi] = As[buf][k4 + lk][wr + 16 * bi + lr];
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) bv[bj] = Bs[buf][k4 + lk][wc + 16 * bj + lr];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
          acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], bv[bj], acc[bi][bj], 0, 0, 0);
    }
    buf ^= 1;
    // the next iteration stages into the other buffer; the one just read is rewritten two slabs on,
    // after the barrier of the next iteration
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) {
      const int col = c0 + wc + 16 * bj + lr;
      if (col >= p.n) continue;
      const double column = p.col_scale ? p.col_scale[col] : 1.0;
      const double scale = p.alpha * column;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = r0 + wr + 16 * bi + 4 * r + lk;
        if (row >= p.m) continue;
        double *const out = p.c + int64_t(row) * p.ldc + col;
        const double value = p.row_shift ? p.alpha * acc[bi][bj][r] / (column - p.row_shift[row])
                                         : scale * acc[bi][bj][r];
        *out = p.beta == 0.0 ? value : value + p.beta * *out;
      }
    }
}

// One wave per row: the Euclidean norm of its `cols` entries, summed in a fixed order (lane
// partial sums over a stride of 64, then a butterfly over the wave).
constexpr int kRowWaves = 4;

__global__ __launch_bounds__(64 * kRowWaves) void row_norms_kernel(const tt_rows_problem *__restrict__ problems,
                                                                  int count, int64_t rows) {
  const int64_t row = int64_t(blockIdx.x) * kRowWaves + (threadIdx.x >> 6);
  if (row >= rows) return;
  const int lane = threadIdx.x & 63;
  const tt_rows_problem p = problems[find_problem(problems, count, row)];
  const int64_t local = row - p.first;
  const double *__restrict__ x = p.src + local * p.src_ld;
  double sum = 0.0;
  for (int j = lane; j < p.cols; j += 64) sum = fma(x[j], x[j], sum);
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) sum += __shfl_xor(sum, offset);
  if (lane == 0) p.dst[local] = sqrt(sum);
}

// One wave per destination row: row i of the problem takes source row index[i] (i when there is no
// index), written as row i of dst (`transpose` 0) or as column i of dst (`transpose` 1).
__global__ __launch_bounds__(64 * kRowWaves) void copy_rows_kernel(const tt_rows_problem *__restrict__ problems,
                                                                  int count, int64_t rows) {
  const int64_t row = int64_t(blockIdx.x) * kRowWaves + (threadIdx.x >> 6);
  if (row >= rows) return;
  const int lane = threadIdx.x & 63;
  const tt_rows_problem p = problems[find_problem(problems, count, row)];
  const int64_t local = row - p.first;
  const int64_t from = p.index ? p.index[local] : local;
  const double *__restrict__ x = p.src + from * p.src_ld;
  if (p.transpose) {
    for (int j = lane; j < p.cols; j += 64) p.dst[int64_t(j) * p.dst_ld + local] = x[j];
  } else {
    double *__restrict__ y = p.dst + local * p.dst_ld;
    for (int j = lane; j < p.cols; j += 64) y[j] = x[j];
  }
}

// One wave per row: dst = alpha x + beta y + gamma I (the identity's entry on the row's diagonal),
// y and the identity optional.
__global__ __launch_bounds__(64 * kRowWaves) void lincomb_rows_kernel(const tt_lincomb_problem *__restrict__ problems,
                                                                     int count, int64_t rows) {
  const int64_t row = int64_t(blockIdx.x) * kRowWaves + (threadIdx.x >> 6);
  if (row >= rows) return;
  const int lane = threadIdx.x & 63;
  const tt_lincomb_problem p = problems[find_problem(problems, count, row)];
  const int64_t local = row - p.first;
  const double *__restrict__ x = p.x + local * p.x_ld;
  const double *__restrict__ y = p.y ? p.y + local * p.y_ld : nullptr;
  double *__restrict__ out = p.dst + local * p.dst_ld;
  for (int j = lane; j < p.cols; j += 64) {
    double value = p.alpha * x[j];
    if (y) value += p.beta * y[j];
    if (j == local) value += p.gamma;
    out[j] = value;
  }
}

// ---------------------------------------------------------------------------------------------
// Batched Householder QR (qr_over / rq_over, svd.cpp:225-249, which run at::linalg_qr per block).
// One workgroup per problem.  The matrix is copied into the problem's column-major workspace and
// reduced column by column as LAPACK's geqr2 does: wave 0 forms the column's reflector (dlarfg's
// convention: beta = -sign(alpha) ||(alpha, x)||, tau = (beta - alpha) / beta, v = x / (alpha - beta),
// a column with nothing below its diagonal left as it is), then the trailing columns are dealt out
// one per wave, each a dot product with the reflector (rows over the lanes) and an update.  R is
// written out, and Q is formed in place from the reflectors, last first (org2r).  Every step's
// column work is one wave's, so a step costs two barriers.

constexpr int kQrThreads = 1024;
constexpr int kQrWaves = kQrThreads / 64;

// Sum over the wave's 64 lanes by butterfly shuffles: every lane ends with the same total, summed
// in the same order whatever the data.
__device__ inline double wave_total(double v) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
  return v;
}

// Apply the reflector of column j (v = 1 at row j, W[j+1.., j] below it, factor tj) to columns
// [from, to) of W over rows j..m-1, one column per wave in turn.
__device__ inline void qr_apply(double *__restrict__ W, int m, int j, double tj, int from, int to, int wave, int lane) {
  const double *__restrict__ v = W + int64_t(j) * m;
  for (int c = from + wave; c < to; c += kQrWaves) {
    double *__restrict__ y = W + int64_t(c) * m;
    double s = lane == 0 ? y[j] : 0.0;
    for (int i = j + 1 + lane; i < m; i += 64) s += v[i] * y[i];
    const double f = tj * wave_total(s);
    if (lane == 0) y[j] -= f;
    for (int i = j + 1 + lane; i < m; i += 64) y[i] -= f * v[i];
  }
}

__global__ __launch_bounds__(kQrThreads) void qr_kernel(const tt_qr_problem *__restrict__ problems) {
  const tt_qr_problem p = problems[blockIdx.x];
  const int m = p.m, n = p.n, k = m < n ? m : n;
  if (m == 0 || n == 0) return;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  double *__restrict__ W = p.work;
  double *__restrict__ tau = p.work + int64_t(m) * n;
  const int64_t size = int64_t(m) * n;
  // W[i + j m] = A(i, j): A is the source (rows x cols, m = rows) or its transpose (m = cols); the
  // source is read along its rows
  for (int64_t e = t; e < size; e += kQrThreads) {
    if (!p.trans) {
      const int64_t i = e / n, j = e % n;
      W[i + j * m] = p.src[i * p.src_ld + j];
    } else {
      const int64_t j = e / m, i = e % m;
      W[i + j * m] = p.src[j * p.src_ld + i];
    }
  }
  __syncthreads();
  for (int j = 0; j < k; ++j) {
    if (wave == 0) {
      double *__restrict__ col = W + int64_t(j) * m;
      double s = 0.0;
      for (int i = j + 1 + lane; i < m; i += 64) s += col[i] * col[i];
      const double sigma = wave_total(s), alpha = col[j];
      double tj = 0.0, beta = alpha, scale = 1.0;
      if (sigma != 0.0) {
        const double norm = sqrt(alpha * alpha + sigma);
        beta = alpha >= 0.0 ? -norm : norm;
        tj = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
      for (int i = j + 1 + lane; i < m; i += 64) col[i] *= scale;
      if (lane == 0) {
        col[j] = beta;
        tau[j] = tj;
      }
    }
    __syncthreads();
    const double tj = tau[j];
    if (tj != 0.0) qr_apply(W, m, j, tj, j + 1, n, wave, lane);
    __syncthreads();
  }
  // R (k x n, zeros below the diagonal), or its transpose
  for (int64_t e = t; e < int64_t(k) * n; e += kQrThreads) {
    if (!p.r_trans) {
      const int64_t i = e / n, c = e % n;
      p.r[i * p.r_ld + c] = c >= i ? W[i + c * m] : 0.0;
    } else {
      const int64_t c = e / k, i = e % k;
      p.r[c * p.r_ld + i] = c >= i ? W[i + c * m] : 0.0;
    }
  }
  __syncthreads();
  // Q = H_0 ... H_{k-1} [I; 0] in the first k columns, the last reflector first (org2r)
  for (int j = k - 1; j >= 0; --j) {
    const double tj = tau[j];
    if (tj != 0.0) qr_apply(W, m, j, tj, j + 1, k, wave, lane);
    __syncthreads();
    double *__restrict__ col = W + int64_t(j) * m;
    for (int i = t; i < m; i += kQrThreads) col[i] = i < j ? 0.0 : (i == j ? 1.0 - tj : -tj * col[i]);
    __syncthreads();
  }
  // Q (m x k), or its transpose
  for (int64_t e = t; e < int64_t(m) * k; e += kQrThreads) {
    if (!p.q_trans) {
      const int64_t i = e / k, c = e % k;
      p.q[i * p.q_ld + c] = W[i + c * m];
    } else {
      const int64_t c = e / m, i = e % m;
      p.q[c * p.q_ld + i] = W[i + c * m];
    }
  }
}

template <typename Problem>
int check_table(const Problem *problems, int64_t count, const char *what) {
  if (count < 0 || (count > 0 && !problems)) return fail(std::string(what) + ": bad problem table");
  if (count > (int64_t(1) << 30)) return fail(std::string(what) + ": too many problems");
  return 0;
}

}  // namespace

extern "C" int tt_bgemm(const tt_bgemm_problem *problems, int64_t count, int64_t tiles, void *stream) {
  if (int rc = check_table(problems, count, "tt_bgemm")) return rc;
  if (tiles < 0 || tiles > (int64_t(1) << 31) - 1) return fail("tt_bgemm: tile count out of range");
  if (count == 0 || tiles == 0) return 0;
  hipLaunchKernelGGL(bgemm_kernel, dim3(static_cast<uint32_t>(tiles)), dim3(kThreads), 0, as_stream(stream),
                     problems, static_cast<int>(count));
  TT_LAUNCHED();
  return 0;
}

extern "C" int tt_row_norms(const tt_rows_problem *problems, int64_t count, int64_t rows, void *stream) {
  if (int rc = check_table(problems, count, "tt_row_norms")) return rc;
  if (rows < 0) return fail("tt_row_norms: negative row count");
  if (count == 0 || rows == 0) return 0;
  hipLaunchKernelGGL(row_norms_kernel, dim3(static_cast<uint32_t>((rows + kRowWaves - 1) / kRowWaves)),
                     dim3(64 * kRowWaves), 0, as_stream(stream), problems, static_cast<int>(count), rows);
  TT_LAUNCHED();
  return 0;
}

extern "C" int tt_copy_rows(const tt_rows_problem *problems, int64_t count, int64_t rows, void *stream) {
  if (int rc = check_table(problems, count, "tt_copy_rows")) return rc;
  if (rows < 0) return fail("tt_copy_rows: negative row count");
  if (count == 0 || rows == 0) return 0;
  hipLaunchKernelGGL(copy_rows_kernel, dim3(static_cast<uint32_t>((rows + kRowWaves - 1) / kRowWaves)),
                     dim3(64 * kRowWaves), 0, as_stream(stream), problems, static_cast<int>(count), rows);
  TT_LAUNCHED();
  return 0;
}

extern "C" int tt_lincomb_rows(const tt_lincomb_problem *problems, int64_t count, int64_t rows, void *stream) {
  if (int rc = check_table(problems, count, "tt_lincomb_rows")) return rc;
  if (rows < 0) return fail("tt_lincomb_rows: negative row count");
  if (count == 0 || rows == 0) return 0;
  hipLaunchKernelGGL(lincomb_rows_kernel, dim3(static_cast<uint32_t>((rows + kRowWaves - 1) / kRowWaves)),
                     dim3(64 * kRowWaves), 0, as_stream(stream), problems, static_cast<int>(count), rows);
  TT_LAUNCHED();
  return 0;
}

extern "C" int tt_qr_batched(const tt_qr_problem *problems, int64_t count, void *stream) {
  if (int rc = check_table(problems, count, "tt_qr_batched")) return rc;
  if (count > (int64_t(1) << 31) - 1) return fail("tt_qr_batched: too many problems");
  if (count == 0) return 0;
  hipLaunchKernelGGL(qr_kernel, dim3(static_cast<uint32_t>(count)), dim3(kQrThreads), 0, as_stream(stream), problems);
  TT_LAUNCHED();
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Host-side table checks.

namespace {

struct Spans {
  const int64_t *starts, *ends;
  int64_t count;
  // bytes [pointer, pointer + last] lie in one span
  bool holds(const void *pointer, int64_t last) const {
    const int64_t lo = static_cast<int64_t>(reinterpret_cast<uintptr_t>(pointer)), hi = lo + last;
    int64_t a = 0, b = count;   // the last span starting at or before lo
    while (a < b) {
      const int64_t mid = (a + b) / 2;
      if (starts[mid] <= lo) a = mid + 1; else b = mid;
    }
    return a > 0 && hi < ends[a - 1];
  }
};

// the last byte offset a row-major rows x cols operand with row stride ld touches
int64_t touched(int64_t rows, int64_t cols, int64_t ld) { return 8 * ((rows - 1) * ld + cols - 1); }

int outside(const char *what, const void *pointer) {
  char text[64];
  std::snprintf(text, sizeof text, "%p", pointer);
  return fail(std::string(what) + ": the operand at " + text + " reaches outside every buffer of the launch");
}

int check_spans(const int64_t *starts, const int64_t *ends, int64_t spans, const char *what) {
  if (spans < 0 || (spans > 0 && (!starts || !ends))) return fail(std::string(what) + ": bad buffer list");
  for (int64_t i = 0; i < spans; ++i)
    if (ends[i] < starts[i] || (i > 0 && starts[i] < ends[i - 1]))
      return fail(std::string(what) + ": buffer ranges unsorted or overlapping");
  return 0;
}

}  // namespace

extern "C" int tt_bgemm_check(const tt_bgemm_problem *problems, int64_t count, const int64_t *starts,
                              const int64_t *ends, int64_t spans, int64_t *items) {
  if (int rc = check_table(problems, count, "bgemm")) return rc;
  if (int rc = check_spans(starts, ends, spans, "bgemm")) return rc;
  const Spans in{starts, ends, spans};
  int64_t tiles = 0;
  for (int64_t i = 0; i < count; ++i) {
    const tt_bgemm_problem &p = problems[i];
    const int64_t m = p.m, n = p.n, k = p.k;
    if (m < 0 || n < 0 || k < 0) return fail("bgemm: negative problem sizes");
    if (p.first != tiles) return fail("bgemm: work items are not numbered consecutively");
    const int64_t a_rows = p.trans_a ? k : m, a_cols = p.trans_a ? m : k;
    const int64_t b_rows = p.trans_b ? n : k, b_cols = p.trans_b ? k : n;
    if (a_cols > p.lda || b_cols > p.ldb || n > p.ldc) return fail("bgemm: a row stride is shorter than its row");
    if (m > 0 && k > 0 && !in.holds(p.a, touched(a_rows, a_cols, p.lda))) return outside("bgemm A", p.a);
    if (n > 0 && k > 0 && !in.holds(p.b, touched(b_rows, b_cols, p.ldb))) return outside("bgemm B", p.b);
    if (m > 0 && n > 0) {
      if (!in.holds(p.c, touched(m, n, p.ldc))) return outside("bgemm C", p.c);
      if (p.col_scale && !in.holds(p.col_scale, 8 * (n - 1))) return outside("bgemm column scale", p.col_scale);
      if (p.row_shift && !p.col_scale) return fail("bgemm: row shifts need column shifts");
      if (p.row_shift && !in.holds(p.row_shift, 8 * (m - 1))) return outside("bgemm row shift", p.row_shift);
    }
    tiles += ((m + kTile - 1) / kTile) * ((n + kTile - 1) / kTile);
  }
  if (items) *items = tiles;
  return 0;
}

extern "C" int tt_rows_check(const tt_rows_problem *problems, int64_t count, const int64_t *source_rows, int vector_dst,
                             const int64_t *starts, const int64_t *ends, int64_t spans, int64_t *items) {
  if (int rc = check_table(problems, count, "row problem")) return rc;
  if (int rc = check_spans(starts, ends, spans, "row problem")) return rc;
  const Spans in{starts, ends, spans};
  int64_t rows = 0;
  for (int64_t i = 0; i < count; ++i) {
    const tt_rows_problem &p = problems[i];
    const int64_t r = p.rows, c = p.cols;
    if (r < 0 || c < 0 || c > p.src_ld) return fail("row problem: bad sizes");
    if (p.first != rows) return fail("row problem: work items are not numbered consecutively");
    if (r > 0 && c > 0) {
      int64_t held = r;
      if (p.index) {
        if (!source_rows) return fail("row problem: an index needs the source row counts");
        held = source_rows[i];
        if (!in.holds(p.index, 4 * (r - 1))) return outside("row index", p.index);
      }
      if (!in.holds(p.src, touched(held, c, p.src_ld))) return outside("row source", p.src);
      const int64_t last = vector_dst ? 8 * (r - 1) : (p.transpose ? touched(c, r, p.dst_ld) : touched(r, c, p.dst_ld));
      if (!in.holds(p.dst, last)) return outside("row destination", p.dst);
    }
    rows += r;
  }
  if (items) *items = rows;
  return 0;
}

extern "C" int tt_lincomb_check(const tt_lincomb_problem *problems, int64_t count, const int64_t *starts,
                                const int64_t *ends, int64_t spans, int64_t *items) {
  if (int rc = check_table(problems, count, "lincomb")) return rc;
  if (int rc = check_spans(starts, ends, spans, "lincomb")) return rc;
  const Spans in{starts, ends, spans};
  int64_t rows = 0;
  for (int64_t i = 0; i < count; ++i) {
    const tt_lincomb_problem &p = problems[i];
    const int64_t r = p.rows, c = p.cols;
    if (r < 0 || c < 0 || c > p.x_ld || c > p.dst_ld || (p.y && c > p.y_ld))
      return fail("lincomb: bad sizes or row strides");
    if (p.first != rows) return fail("lincomb: work items are not numbered consecutively");
    if (r > 0 && c > 0) {
      if (!in.holds(p.x, touched(r, c, p.x_ld))) return outside("lincomb x", p.x);
      if (p.y && !in.holds(p.y, touched(r, c, p.y_ld))) return outside("lincomb y", p.y);
      if (!in.holds(p.dst, touched(r, c, p.dst_ld))) return outside("lincomb destination", p.dst);
    }
    rows += r;
  }
  if (items) *items = rows;
  return 0;
}

extern "C" int tt_qr_check(const tt_qr_problem *problems, int64_t count, const int64_t *starts, const int64_t *ends,
                           int64_t spans) {
  if (int rc = check_table(problems, count, "qr")) return rc;
  if (int rc = check_spans(starts, ends, spans, "qr")) return rc;
  const Spans in{starts, ends, spans};
  for (int64_t i = 0; i < count; ++i) {
    const tt_qr_problem &p = problems[i];
    const int64_t m = p.m, n = p.n, k = std::min(m, n);
    if (m < 0 || n < 0) return fail("qr: negative problem sizes");
    if (m == 0 || n == 0) continue;
    // the source is rows x cols as stored: m x n, or n x m when it is factored transposed
    const int64_t rows = p.trans ? n : m, cols = p.trans ? m : n;
    if (cols > p.src_ld) return fail("qr: the source's row stride is shorter than its row");
    if ((p.q_trans ? m : k) > p.q_ld || (p.r_trans ? k : n) > p.r_ld)
      return fail("qr: an output row stride is shorter than its row");
    if (!in.holds(p.src, touched(rows, cols, p.src_ld))) return outside("qr source", p.src);
    if (!in.holds(p.q, p.q_trans ? touched(k, m, p.q_ld) : touched(m, k, p.q_ld))) return outside("qr Q", p.q);
    if (!in.holds(p.r, p.r_trans ? touched(n, k, p.r_ld) : touched(k, n, p.r_ld))) return outside("qr R", p.r);
    if (!in.holds(p.work, 8 * (m * n + k - 1))) return outside("qr workspace", p.work);
  }
  return 0;
}
