// The device library, part four: batched symmetric eigendecomposition in FP64 for the split of the
// two-site tensor (the Gram matrices of its blocks), for a whole batch of matrices of different
// sizes at once:
//
//   1. Householder tridiagonalisation.  The rank-2 update of step k is deferred into step k+1's
//      symmetric matrix-vector pass, so each step reads and writes the trailing matrix once.  The
//      reflectors stay in the matrix's lower part.  While a trailing matrix exceeds 192 rows, one
//      launch per step spreads every matrix over one wave per trailing column; the last 192 steps
//      of every matrix run in one launch, one workgroup per matrix, the trailing matrix held in
//      the workgroup's registers.
//   2. Cuppen divide and conquer on the tridiagonal: leaves of 8 rows by implicit QL (one wave
//      each), then merges bottom up, four launches per level for every matrix of the batch: one
//      workgroup per merge deflates small coupling components and close poles (Givens rotations);
//      one wave per secular root takes rational ("middle way") steps from the nearer pole; one
//      workgroup per merge forms the eigenvectors of the rank-one update by the Gu-Eisenstat
//      formula (orthogonal to working precision however close the roots); the merge products run
//      as 64 x 64 FP64 MFMA tiles spread over the GPU.
//   3. Back-transformation of the eigenvectors by the reflectors in blocks of 32 (compact WY
//      form), one workgroup per 16-column panel held in registers, on FP64 MFMA.
//
// Eigenvalues come out ascending; eigenvector j is row j of the matrix's n x n output.  gfx950 only.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "common.h"

using ttdev::as_stream;
using ttdev::fail;

namespace {

constexpr int kEigThreads = 256;
constexpr int kEigWaves = kEigThreads / 64;
constexpr int kEigMaxN = 2048;      // largest matrix (LDS vectors are sized for it)
constexpr int kEigBatch = 64;       // problems per launch (kernel-argument table)
constexpr int kPanel = 16;          // columns per back-transformation workgroup
constexpr int kLeaf = 8;            // order of the divide-and-conquer leaves
constexpr double kEps = 2.220446049250313e-16;

struct Problem {
  int64_t matrix;       // offset of the n x n matrix (and of its eigenvectors)
  int64_t values;       // offset of its n eigenvalues
  int64_t work;         // offset of its workspace (3 n^2 + 7 n + 1024 ceil((n - 1) / 32) doubles)
  int32_t n;
  int32_t first_panel;  // index of its first back-transformation panel in the launch
  int32_t first_leaf;   // index of its first divide-and-conquer leaf in the launch
  int32_t pad;
};

struct Batch {
  Problem p[kEigBatch];
  int32_t count;
};

typedef double d4 __attribute__((ext_vector_type(4)));

// One DPP move of a double (both halves); lanes of rows outside kRowMask receive 0.
template <int kCtrl, int kRowMask>
__device__ inline double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), kCtrl, kRowMask, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), kCtrl, kRowMask, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Largest of non-negative values over the wave's 64 lanes (the same DPP pattern as wave_sum).
__device__ double wave_max(double v) {
  v = fmax(v, dpp_move<0xb1, 0xf>(v));
  v = fmax(v, dpp_move<0x4e, 0xf>(v));
  v = fmax(v, dpp_move<0x141, 0xf>(v));
  v = fmax(v, dpp_move<0x140, 0xf>(v));
  v = fmax(v, dpp_move<0x142, 0xa>(v));
  v = fmax(v, dpp_move<0x143, 0xc>(v));
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// Sum over the wave's 64 lanes, the same value in every lane: quad and mirror moves total each
// row of 16 lanes, two row broadcasts carry the totals into lane 63, which every lane then reads.
__device__ double wave_sum(double v) {
  v += dpp_move<0xb1, 0xf>(v);    // quad_perm [1, 0, 3, 2]
  v += dpp_move<0x4e, 0xf>(v);    // quad_perm [2, 3, 0, 1]
  v += dpp_move<0x141, 0xf>(v);   // row_half_mirror
  v += dpp_move<0x140, 0xf>(v);   // row_mirror
  v += dpp_move<0x142, 0xa>(v);   // row_bcast:15, into rows 1 and 3
  v += dpp_move<0x143, 0xc>(v);   // row_bcast:31, into rows 2 and 3
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// Lane `lane`'s value of a double (uniform lane index).
__device__ inline double readlane_double(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// The problem a wave of a batched launch works on, when problem b owns count(n_b) consecutive
// work items (workgroups, or waves for the leaves) in batch order.  Lane b of the wave loads problem
// b (the table always holds kEigBatch entries, zeroed past the count), a prefix sum over the lanes'
// counts and a ballot then give the problem: one memory round trip, where walking the table would
// be a chain of dependent loads ahead of every launch's own.  All 64 lanes must take part.
// Returns false past the last problem; `first` is the index of the problem's first work item.
struct Located {
  Problem pr;
  int b, first;
};

static_assert(kEigBatch == 64, "one lane per problem of a batch");

template <class Count>
__device__ inline bool locate(const Batch *__restrict__ batch, int item, Count count, Located *out) {
  const int lane = threadIdx.x & 63;
  const Problem mine = batch->p[lane];
  const int problems = batch->count;
  int incl = lane < problems ? count(mine.n) : 0;
#pragma unroll
  for (int offset = 1; offset < 64; offset <<= 1) {
    const int below = __shfl_up(incl, offset, 64);
    if (lane >= offset) incl += below;
  }
  const int b = __popcll(__ballot(incl <= item));
  if (b >= problems) return false;
  out->b = b;
  out->first = b > 0 ? __shfl(incl, b - 1, 64) : 0;
  out->pr.matrix = __shfl(static_cast<long long>(mine.matrix), b, 64);
  out->pr.values = __shfl(static_cast<long long>(mine.values), b, 64);
  out->pr.work = __shfl(static_cast<long long>(mine.work), b, 64);
  out->pr.n = __shfl(mine.n, b, 64);
  out->pr.first_panel = __shfl(mine.first_panel, b, 64);
  out->pr.first_leaf = __shfl(mine.first_leaf, b, 64);
  out->pr.pad = 0;
  return true;
}

// ---------------------------------------------------------------------------------------------
// 1. Tridiagonalisation.  A is column-major n x n (symmetric, so either order), in place.
//    Output: diag[n], off[n-1], tau[n-1]; reflector k is A[k+1.., k] with A[k+1, k] = 1.  The
//    rank-2 update of step k is deferred into step k+1's pass over the trailing matrix.  While a
//    trailing matrix exceeds kRegOrder rows the steps run one launch each over many workgroups
//    (1'); the last kRegOrder steps of every matrix run in one launch, one workgroup per matrix
//    with the trailing matrix in registers (1'').

// 1'. The steps while trailing matrices are large: one launch per Householder step k, one wave per
//     trailing column of every matrix, four waves a workgroup.  Each wave works alone, with no
//     barrier and no LDS: its lanes hold rows k + lane + 64 q of the step's vectors (v and p of
//     step k-1, column k as it stands), so the step's O(n) part (w of step k-1, column k with that
//     step's update, the reflector of step k) costs two wave reductions, redone by every wave.
//     Then the wave applies step k-1's update to its column (rows k+1..) and forms the column's
//     entry of p = tau A v (A is symmetric, so a column dot product).  A step's time is then the
//     launch and one round trip to memory, where a workgroup-wide prologue paid its barriers on
//     every step.  Between launches the state lives in the workspace: v and p of the last two
//     steps (by step parity, so a launch never writes what it reads).  The first wave of a matrix
//     writes the step's scalars, v, and the reflector of step k-1 into column k-1, which no wave
//     reads at step k; a last launch (k = n-1, one wave) forms the last diagonal entry and stores
//     the last reflector.  The kernel is instantiated for 4 to 32 row chunks; each step takes the
//     smallest that covers its largest trailing matrix.

constexpr int kStepThreads = 256;
constexpr int kStepCols = kStepThreads / 64;
static_assert(32 * 64 >= kEigMaxN, "the largest instance covers every row");

// workgroups of a matrix of order n at step k
__host__ __device__ inline int step_blocks(int n, int k) {
  return k + 1 < n ? (n - k - 1 + kStepCols - 1) / kStepCols : (k + 1 == n ? 1 : 0);
}

template <int kChunks>
__global__ __launch_bounds__(kStepThreads) void tridiag_step_kernel(const Batch *__restrict__ batch,
                                                                    double *__restrict__ matrices,
                                                                    double *__restrict__ work, int k) {
  Located at;
  if (!locate(batch, static_cast<int>(blockIdx.x), [k](int order) { return step_blocks(order, k); }, &at)) return;
  const int n = at.pr.n, block = static_cast<int>(blockIdx.x) - at.first;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c = k + 1 + block * kStepCols + wave;   // this wave's column
  const bool last = k + 1 == n;
  if (last ? wave != 0 : c >= n) return;
  double *__restrict__ A = matrices + at.pr.matrix;
  double *diag = work + at.pr.work + 3 * int64_t(n) * n, *off = diag + n, *tau = off + n;
  double *vbuf = tau + n, *pbuf = vbuf + 2 * n;   // [2][n] each, by step parity
  const double *vprev = vbuf + ((k - 1) & 1) * n, *pprev = pbuf + ((k - 1) & 1) * n;
  // every load of the step at once: the vectors of step k-1, column k, this wave's column
  double v[kChunks], p[kChunks], a[kChunks], y[kChunks];
#pragma unroll
  for (int q = 0; q < kChunks; ++q) {
    const int i = k + lane + 64 * q;
    const bool in = i < n;
    v[q] = in && k > 0 ? vprev[i] : 0.0;
    p[q] = in && k > 0 ? pprev[i] : 0.0;
    a[q] = in ? A[i + int64_t(k) * n] : 0.0;
    y[q] = in && !last ? A[i + int64_t(c) * n] : 0.0;
  }
  const double tprev = k > 0 ? tau[k - 1] : 0.0;
  const double vcol = k > 0 && !last ? vprev[c] : 0.0, pcol = k > 0 && !last ? pprev[c] : 0.0;
  // w = p - tau/2 (p . v) v of step k-1, and column k with step k-1's update
  double partial = 0.0;
#pragma unroll
  for (int q = 0; q < kChunks; ++q) partial += p[q] * v[q];
  const double half = 0.5 * tprev * wave_sum(partial);
  double *const w = p, *const x = a;   // w replaces p, and the updated column k replaces a
#pragma unroll
  for (int q = 0; q < kChunks; ++q) w[q] = p[q] - half * v[q];
  const double vk = readlane_double(v[0], 0), wk = readlane_double(w[0], 0);   // row k: lane 0, chunk 0
  partial = 0.0;
#pragma unroll
  for (int q = 0; q < kChunks; ++q) {
    const int i = k + lane + 64 * q;
    x[q] = i < n ? a[q] - (v[q] * wk + w[q] * vk) : 0.0;
    if (i >= k + 2) partial += x[q] * x[q];
  }
  const double sigma = wave_sum(partial);
  const double dk = readlane_double(x[0], 0);
  const bool leader = block == 0 && wave == 0;
  if (last) {
    if (lane == 0) diag[k] = dk;
#pragma unroll
    for (int q = 0; q < kChunks; ++q) {
      const int i = k + lane + 64 * q;
      if (k > 0 && i < n) A[i + int64_t(k - 1) * n] = v[q];
    }
    return;
  }
  const double alpha = readlane_double(x[0], 1);   // row k + 1
  double beta = alpha, tk = 0.0, scale = 0.0;
  if (sigma != 0.0) {
    const double norm = sqrt(alpha * alpha + sigma);
    beta = alpha >= 0.0 ? -norm : norm;
    tk = (beta - alpha) / beta;
    scale = 1.0 / (alpha - beta);
  }
  // the reflector of step k, and this wave's column with step k-1's update, dotted with it
  const double wcol = pcol - half * vcol;
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < kChunks; ++q) {
    const int i = k + lane + 64 * q;
    const double vnew = i == k + 1 ? 1.0 : (i >= k + 2 && i < n ? x[q] * scale : 0.0);
    if (leader && i < n) {
      vbuf[(k & 1) * n + i] = vnew;
      if (k > 0) A[i + int64_t(k - 1) * n] = v[q];
    }
    if (i >= k + 1 && i < n) {
      const double z = y[q] - (v[q] * wcol + w[q] * vcol);
      A[i + int64_t(c) * n] = z;
      acc += z * vnew;
    }
  }
  acc = wave_sum(acc);
  if (lane == 0) {
    pbuf[(k & 1) * n + c] = tk * acc;
    if (leader) {
      diag[k] = dk;
      off[k] = beta;
      tau[k] = tk;
    }
  }
}

// 1''. The last kRegOrder steps of every matrix in one launch, one workgroup of eight waves per
//      matrix, with the trailing matrix in registers.  Wave w holds rows 24w..24w+23 of it: lane
//      (row group g = lane / 16, column lane l = lane % 16) holds rows 24w + 6g + a (a < 6) of
//      columns l + 16b (b < 12), 72 entries.  Each step takes three barriers: the lanes holding
//      column kk write it, with the pending update, to LDS; every wave then forms the step's
//      reflector from it by itself (one wave reduction), applies the pending update to its
//      entries and writes its rows' partial products with the reflector to LDS, one per column
//      lane; after the second barrier one thread per row sums its 16 partials in order (p = tau
//      A v), and after the third every wave forms p . v by itself.  The pending update (v, p and
//      tau/2 (p . v) of the last step) sits in LDS, double-buffered by step parity.  Finished rows
//      and columns carry zero vectors, so the update leaves them unchanged; padding past the
//      matrix's order is zero, and the update skips waves and column groups wholly finished or
//      past the order.  Against 64 stepped launches and the 128-step LDS launch this kernel
//      replaces, a converged chi = 1024 sweep ran 3.4 % faster (same saved state, same box).
//      Takes over from the stepped launches with the state they left (k0 > 0), or reduces a whole
//      matrix of order <= kRegOrder (k0 = 0).

constexpr int kRegOrder = 192;
constexpr int kRegWaves = 8;
constexpr int kRegThreads = 64 * kRegWaves;
constexpr int kRegRows = 6;                   // rows per lane
constexpr int kRegCols = 12;                  // columns per lane
static_assert(kRegWaves * 4 * kRegRows == kRegOrder, "the waves' four row groups cover the rows");
static_assert(16 * kRegCols == kRegOrder, "sixteen column lanes cover the columns");
constexpr int kRegSpan = 3 * 64;              // LDS vector length: a wave reads three entries a lane
static_assert(kRegOrder <= kRegSpan, "a wave holds a column in three rows per lane");

// The workgroup barrier of the register kernel: its waves share only LDS, so the fences cover LDS
// alone and the barrier does not wait for the reflectors' global stores.
__device__ inline void reg_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__global__ __launch_bounds__(kRegThreads) void tridiag_reg_kernel(const Batch *__restrict__ batch,
                                                                   double *__restrict__ matrices,
                                                                   double *__restrict__ work, int k0) {
  const Problem pr = batch->p[blockIdx.x];
  const int n = pr.n;
  if (n - 1 < k0) return;   // finished by the stepped launches
  const int m = n - k0;     // at most kRegOrder (the host starts this launch late enough)
  double *__restrict__ A = matrices + pr.matrix;
  double *diag = work + pr.work + 3 * int64_t(n) * n, *off = diag + n, *tau = off + n;
  const double *vbuf = tau + n, *pbuf = vbuf + 2 * n;
  __shared__ double X[2][kRegSpan], V[2][kRegSpan], P[2][kRegSpan];
  __shared__ double S[kRegOrder][16];   // the rows' partial products, by column lane
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int group = lane >> 4, col_lane = lane & 15;
  const int row0 = wave * 4 * kRegRows + group * kRegRows;   // this lane's first row
  // the trailing matrix, zero past its order (entries past it load the first entry, then drop it:
  // the loads stay unconditional)
  double T[kRegRows][kRegCols];
#pragma unroll
  for (int b = 0; b < kRegCols; ++b) {
    const int c = col_lane + 16 * b;
#pragma unroll
    for (int a = 0; a < kRegRows; ++a) {
      const int r = row0 + a;
      const bool in = r < m && c < m;
      const double y = A[in ? (k0 + r) + int64_t(k0 + c) * n : k0 + int64_t(k0) * n];
      T[a][b] = in ? y : 0.0;
    }
  }
  // the pending update of step k0-1 (none before the first step) in parity 1, zero past the order
  // (the vectors' entries up to kRegSpan are read); its reflector goes to column k0-1
  for (int r = t; r < kRegSpan; r += kRegThreads) {
    const bool on = k0 > 0 && r < m;
    const double v = on ? vbuf[((k0 - 1) & 1) * n + k0 + r] : 0.0;
    V[1][r] = v;
    P[1][r] = on ? pbuf[((k0 - 1) & 1) * n + k0 + r] : 0.0;
    V[0][r] = 0.0;
    P[0][r] = 0.0;
    X[0][r] = 0.0;
    X[1][r] = 0.0;
    if (on) A[(k0 + r) + int64_t(k0 - 1) * n] = v;
  }
  for (int e = t; e < kRegOrder * 16; e += kRegThreads) S[e >> 4][e & 15] = 0.0;
  const double tprev = k0 > 0 ? tau[k0 - 1] : 0.0;
  __syncthreads();
  // tau/2 (p . v) of the pending step in buffer `buf`, over rows `from`.., by every wave for itself
  // (every LDS entry holds a finite value, so the reads are unconditional and the rows masked)
  auto half_of = [&](int buf, int from, double tk) {
    double partial = 0.0;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      const int r = lane + 64 * s;
      const double term = P[buf][r] * V[buf][r];
      partial += r >= from ? term : 0.0;
    }
    return 0.5 * tk * wave_sum(partial);
  };
  double half = half_of(1, 0, tprev);
  int pending = 1;   // the parity of the pending update
  const int wave_end = (wave + 1) * 4 * kRegRows;   // one past this wave's last row
  // The steps in blocks of 16, one per column group B: the column of step kk sits at b = B in the
  // lanes with l = kk % 16, and the groups before B are finished, so with B a constant (the loop
  // over B is unrolled) the column is a plain register and the update skips the finished groups.
#pragma unroll
  for (int B = 0; B < kRegCols; ++B) {
    const int block_end = min(16 * B + 16, m - 1);
    for (int kk = 16 * B; kk < block_end; ++kk) {
      const int k = k0 + kk, par = kk & 1, prev = pending;
      // the pending vectors at this lane's rows, and at column kk
      double vr[kRegRows], wr[kRegRows];
#pragma unroll
      for (int a = 0; a < kRegRows; ++a) {
        vr[a] = V[prev][row0 + a];
        wr[a] = P[prev][row0 + a] - half * vr[a];
      }
      const double vk = V[prev][kk], wk = P[prev][kk] - half * vk;
      // (a) the lanes holding column kk write it, rows kk.., with the pending update
      if (col_lane == (kk & 15)) {
#pragma unroll
        for (int a = 0; a < kRegRows; ++a) {
          const int r = row0 + a;
          if (r >= kk && r < m) X[par][r] = T[a][B] - (vr[a] * wk + wr[a] * vk);
        }
      }
      reg_barrier();
      // (b) every wave: the reflector of this step from column kk
      double x[3], partial = 0.0;
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const int r = lane + 64 * s;
        const double y = X[par][r];
        x[s] = r >= kk ? y : 0.0;
        partial += r >= kk + 2 ? y * y : 0.0;
      }
      const double sigma = wave_sum(partial);
      const double dk = X[par][kk], alpha = X[par][kk + 1];
      double beta = alpha, tk = 0.0, scale = 0.0;
      if (sigma != 0.0) {
        const double norm = sqrt(alpha * alpha + sigma);
        beta = alpha >= 0.0 ? -norm : norm;
        tk = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
      if (wave == 0) {
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const int r = lane + 64 * s;
          if (r < m) {
            const double vnew = r == kk + 1 ? 1.0 : (r >= kk + 2 ? x[s] * scale : 0.0);
            V[par][r] = vnew;
            if (r >= kk + 1) A[(k0 + r) + int64_t(k) * n] = vnew;
          }
        }
        if (lane == 0) {
          diag[k] = dk;
          off[k] = beta;
          tau[k] = tk;
        }
      }
      // (c) the pending update on this lane's entries of the live column groups, and their
      //     products with the reflector; a wave whose rows are all finished (their entries are
      //     never read again) or all past the order skips it, and so do column groups past the
      //     order.  The columns go in pairs, each pair's reads fenced off from the next so the
      //     compiler cannot hoist them all (their registers would push the entries out).
      if (wave_end > kk + 1 && wave_end - 4 * kRegRows < m) {
        double acc[kRegRows];
#pragma unroll
        for (int a = 0; a < kRegRows; ++a) acc[a] = 0.0;
#pragma unroll
        for (int b0 = 0; b0 < kRegCols; b0 += 2) {
          if (b0 + 1 < B || 16 * b0 >= m) continue;   // finished, or past the order
          double vc[2], wc[2], vn[2];
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            const int c = col_lane + 16 * (b0 + j);
            // the reflector's entry c as arithmetic on 0/1 factors, so the read stays
            // unconditional (a select around it becomes a branch and a wait per column)
            const double xc = X[par][c];
            vc[j] = V[prev][c];
            wc[j] = P[prev][c] - half * vc[j];
            const int d = c - (kk + 1);   // 0 at the reflector's leading 1, positive below it
            const double below = static_cast<double>(min(max(d, 0), 1));
            const double lead = static_cast<double>(1 - min(abs(d), 1));
            vn[j] = fma(xc * scale, below, lead);
          }
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            if (b0 + j < B) continue;
#pragma unroll
            for (int a = 0; a < kRegRows; ++a) {
              const double y = fma(-vr[a], wc[j], fma(-wr[a], vc[j], T[a][b0 + j]));
              T[a][b0 + j] = y;
              acc[a] += y * vn[j];
            }
          }
          __builtin_amdgcn_sched_barrier(0);
        }
        // (d) the rows' partial sums to LDS, one per column lane
#pragma unroll
        for (int a = 0; a < kRegRows; ++a) S[row0 + a][col_lane] = acc[a];
      }
      reg_barrier();
      // p of this step at rows kk+1.., each row's 16 partials summed in order, zero elsewhere (the
      // rows of waves that skipped the update hold stale partials, and are finished)
      if (t < kRegOrder) {
        double total = 0.0;
#pragma unroll
        for (int l = 0; l < 16; ++l) total += S[t][l];
        P[par][t] = t >= kk + 1 && t < m ? tk * total : 0.0;
      }
      reg_barrier();
      // (e) this step's update becomes the pending one
      half = half_of(par, kk + 1, tk);
      pending = par;
    }
  }
  // the last diagonal entry, with the pending update, from the lane holding it
  const int last = m - 1;
  if (col_lane == (last & 15) && last >= row0 && last < row0 + kRegRows) {
    const int bk = last >> 4, ak = last - row0;
    double y = 0.0;
#pragma unroll
    for (int a = 0; a < kRegRows; ++a)
#pragma unroll
      for (int b = 0; b < kRegCols; ++b)
        if (a == ak && b == bk) y = T[a][b];
    const double v = V[pending][last], w = P[pending][last] - half * v;
    diag[n - 1] = y - 2.0 * v * w;
  }
}

// ---------------------------------------------------------------------------------------------
// 2a. Leaves of the divide and conquer: every block of kLeaf rows of the tridiagonal (with the
//     couplings of its two cuts taken off its end diagonals, Cuppen's split) solved by implicit QL
//     with Wilkinson-type shifts and eigenvector accumulation, one wave per leaf.  Every lane runs
//     the scalar recurrence on the wave's copy of the leaf in LDS; lane r applies the rotations to
//     row r of the leaf's eigenvectors.  The eigenpairs are then sorted ascending and written to the
//     leaf's diagonal block of Q, one entry per lane (the rest of Q was zeroed with the batch table).

static_assert(kLeaf * kLeaf == 64, "a leaf's eigenvector block is written by one wave, one entry per lane");

__global__ __launch_bounds__(kEigThreads) void leaf_kernel(const Batch *__restrict__ batch, double *__restrict__ values,
                                                            double *__restrict__ work, int32_t leaves) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int leaf = static_cast<int>(blockIdx.x) * kEigWaves + wave;
  __shared__ double dl[kEigWaves][kLeaf], el[kEigWaves][kLeaf], Z[kEigWaves][kLeaf][kLeaf + 1];
  Located at;
  if (leaf >= leaves || !locate(batch, leaf, [](int order) { return (order + kLeaf - 1) / kLeaf; }, &at)) return;
  const Problem &pr = at.pr;
  const int n = pr.n;
  const int lo = (leaf - pr.first_leaf) * kLeaf, hi = min(lo + kLeaf, n), m = hi - lo;
  double *Q = work + pr.work;
  const double *diag = Q + 3 * int64_t(n) * n, *off = diag + n;
  double *d = dl[wave], *e = el[wave];
  double (*z)[kLeaf + 1] = Z[wave];
  if (lane < m) {
    double v = diag[lo + lane];
    if (lane == 0 && lo > 0) v -= off[lo - 1];
    if (lane == m - 1 && hi < n) v -= off[hi - 1];
    d[lane] = v;
    e[lane] = lane + 1 < m ? off[lo + lane] : 0.0;
    for (int c = 0; c < m; ++c) z[lane][c] = lane == c ? 1.0 : 0.0;
  }
  __builtin_amdgcn_wave_barrier();
  for (int l = 0; l < m; ++l) {
    for (int iter = 0; iter < 60; ++iter) {
      int mm = l;
      for (; mm < m - 1; ++mm)
        if (fabs(e[mm]) <= kEps * (fabs(d[mm]) + fabs(d[mm + 1]))) break;
      if (mm == l) break;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[mm] - d[l] + e[l] / (g + copysign(r, g));
      double sn = 1.0, cs = 1.0, p = 0.0;
      bool split = false;
      for (int i = mm - 1; i >= l; --i) {
        const double f = sn * e[i], bb = cs * e[i];
        r = hypot(f, g);
        __builtin_amdgcn_wave_barrier();
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[mm] = 0.0;
          split = true;
          break;
        }
        sn = f / r;
        cs = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * sn + 2.0 * cs * bb;
        p = sn * r;
        __builtin_amdgcn_wave_barrier();
        d[i + 1] = g + p;
        g = cs * r - bb;
        if (lane < m) {
          const double zi = z[lane][i], zi1 = z[lane][i + 1];
          z[lane][i + 1] = sn * zi + cs * zi1;
          z[lane][i] = cs * zi - sn * zi1;
        }
      }
      __builtin_amdgcn_wave_barrier();
      if (split) continue;
      d[l] -= p;
      e[l] = g;
      e[mm] = 0.0;
      __builtin_amdgcn_wave_barrier();
    }
  }
  // ascending order, vectors along
  for (int i = 0; i < m; ++i) {
    int best = i;
    for (int j = i + 1; j < m; ++j)
      if (d[j] < d[best]) best = j;
    __builtin_amdgcn_wave_barrier();
    if (best != i) {
      const double di = d[i], db = d[best];
      __builtin_amdgcn_wave_barrier();
      d[i] = db;
      d[best] = di;
      if (lane < m) {
        const double a = z[lane][i];
        z[lane][i] = z[lane][best];
        z[lane][best] = a;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  double *lam = values + pr.values;
  if (lane < m) lam[lo + lane] = d[lane];
  // the leaf's eigenvectors in its diagonal block of Q (both buffers were zeroed with the batch table)
  const int r = lane & (kLeaf - 1), c = lane / kLeaf;
  if (r < m && c < m) Q[lo + r + int64_t(lo + c) * n] = z[r][c];
}

// ---------------------------------------------------------------------------------------------
// 2b. Divide and conquer.  Q and Qt are column-major n x n (Q[r + c n]); U is a merge's k x k.

struct Rotation {
  int32_t a, b;  // local columns
  double c, s;
};

// 1/d from the hardware reciprocal estimate and one Newton-Raphson correction r + r (1 - d r).
__device__ inline double reciprocal(double d) {
  const double r = __builtin_amdgcn_rcp(d);
  return fma(r, fma(-d, r, 1.0), r);
}

// A level of the divide and conquer is a list of units, blocks of up to 2 * width rows: the merge of
// two solved halves, or (the last block of a matrix, when it has no partner at this level) a copy.
// Level l reads the eigenvectors from buffer l % 2 (Q or Qt) and writes the other, in four
// launches: deflation (one workgroup per unit), secular roots (one wave per root), finishing (one
// workgroup per unit: the eigenvectors of the rank-one update, the merged eigenvalues, the column
// maps) and the merge products.  A unit leaves in the workspace its k x k matrix U (the
// eigenvectors of the rank-one update in the basis of the kept poles, at 2 n^2 + lo * min(2 width,
// n)) and, per output column j of the unit, the source column colsrc[lo + j] and the final column
// dst[lo + j]: columns j < k are Q[:, colsrc] U[:, j], the others copies of Q[:, colsrc[j]];
// unitk[lo] = k.  Between the launches a unit's poles, coupling components and roots wait in the
// scratch arrays of dc_scratch, indexed from lo.

__host__ __device__ inline int dc_units(int n, int width) { return n > width ? (n + 2 * width - 1) / (2 * width) : 0; }

// merge levels a matrix of order n goes through (the widths kLeaf, 2 kLeaf, ... below n)
__device__ inline int dc_levels(int n) {
  int levels = 0;
  for (int width = kLeaf; width < n; width *= 2) ++levels;
  return levels;
}

struct DcMaps {
  int32_t *colsrc, *dst, *unitk;
};

// integer maps of a level, in the step vectors' space (free once the matrix is tridiagonal)
__device__ inline DcMaps dc_maps(double *base, int n) {
  int32_t *ints = reinterpret_cast<int32_t *>(base + 3 * int64_t(n) * n + 3 * n);
  return DcMaps{ints, ints + n, ints + 2 * n};
}

// Per-unit arrays between the launches of a level, in the space of the back-transformation's T
// matrices (at least 32 (n - 1) doubles, unused until the divide and conquer is done): kept poles
// dk, their squared and signed coupling components z2 and zk, root offsets tau, all sorted poles,
// rho (at lo); root origins org, merge order perm, kept positions keep, deflation flags defl.
struct DcScratch {
  double *dk, *z2, *zk, *tau, *pole, *rho;
  int32_t *org, *perm, *keep, *defl;
};

__device__ inline DcScratch dc_scratch(double *base, int n) {
  double *d = base + 3 * int64_t(n) * n + 7 * n;
  int32_t *i = reinterpret_cast<int32_t *>(d + 6 * n);
  return DcScratch{d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n, i, i + n, i + 2 * n, i + 3 * n};
}

// unit of workgroup `block` among the units of a level: problem, lo (-1 past the batch)
struct UnitRef {
  int b, lo;
  Problem pr;
};

__device__ inline UnitRef find_unit(const Batch *__restrict__ batch, int width, int block) {
  Located at;
  if (!locate(batch, block, [width](int order) { return dc_units(order, width); }, &at)) return UnitRef{-1, 0, Problem{}};
  return UnitRef{at.b, (block - at.first) * 2 * width, at.pr};
}

// Exclusive prefix sum over the workgroup (kEigThreads threads) of one value per thread; the total
// is left in *total.  `scratch` holds kEigThreads + 1 ints.
__device__ int block_exclusive_scan(int value, int *scratch, int *total) {
  const int t = threadIdx.x;
  scratch[t] = value;
  __syncthreads();
  for (int offset = 1; offset < kEigThreads; offset *= 2) {
    const int add = t >= offset ? scratch[t - offset] : 0;
    __syncthreads();
    scratch[t] += add;
    __syncthreads();
  }
  const int inclusive = scratch[t];
  *total = scratch[kEigThreads - 1];
  __syncthreads();
  return inclusive - value;
}

// Per thread, the elements [4 t, 4 t + 4) of an s-long list (s <= kEigMaxN = 4 kEigThreads).
constexpr int kPerThread = kEigMaxN / kEigThreads;
static_assert(kPerThread * kEigThreads == kEigMaxN, "four list elements per thread");

__global__ __launch_bounds__(kEigThreads) void dc_deflate_kernel(const Batch *__restrict__ batch, double *__restrict__ values,
                                                                  double *__restrict__ work, int width, int parity) {
  const UnitRef u = find_unit(batch, width, static_cast<int>(blockIdx.x));
  if (u.b < 0) return;
  const Problem &pr = u.pr;
  const int n = pr.n, lo = u.lo;
  double *const base = work + pr.work;
  double *__restrict__ Q = base + (parity ? int64_t(n) * n : 0);
  const double *diag = base + 3 * int64_t(n) * n, *off = diag + n;
  const DcMaps maps = dc_maps(base, n);
  const DcScratch sc = dc_scratch(base, n);
  const double *lam = values + pr.values;
  const int t = threadIdx.x;
  const int mid = lo + width, hi = min(lo + 2 * width, n), s = hi - lo, left = mid - lo;
  if (mid >= n) {   // no partner at this level: the block moves over unchanged
    for (int j = t; j < s; j += kEigThreads) {
      maps.colsrc[lo + j] = j;
      maps.dst[lo + j] = j;
    }
    if (t == 0) maps.unitk[lo] = 0;
    return;
  }
  __shared__ double dloc[kEigMaxN], zloc[kEigMaxN], pole[kEigMaxN], zsort[kEigMaxN];
  __shared__ int32_t perm[kEigMaxN], defl[kEigMaxN], cand[kEigMaxN], keep[kEigMaxN];
  __shared__ Rotation rot[kEigMaxN];  // each rotation deflates one pole: at most s - 1
  __shared__ int scan[kEigThreads + 1];
  __shared__ double red[kEigWaves];
  __shared__ int s_rot;
  const double beta = off[mid - 1];
  const double sign = beta < 0.0 ? -1.0 : 1.0, rho = 2.0 * beta * sign;
  // (a) coupling vector (last row of the left block, first row of the right) and poles
  for (int i = t; i < s; i += kEigThreads) {
    const double q = i < left ? Q[(mid - 1) + int64_t(lo + i) * n] : Q[mid + int64_t(lo + i) * n];
    zloc[i] = q * 0.7071067811865476;
    dloc[i] = sign * lam[lo + i];
    perm[i] = i;   // overwritten below; keeps every entry a valid index whatever the values
  }
  __syncthreads();
  // (b) merge the two ascending halves (a negated half is descending and walked from its end):
  //     an element's position is its rank in its half plus the count of the other half's elements
  //     before it, ties going to the left half first
  for (int i = t; i < s; i += kEigThreads) {
    const double x = dloc[i];
    const bool in_a = i < left;
    const int rank = in_a ? (sign > 0 ? i : left - 1 - i) : (sign > 0 ? i - left : s - 1 - i);
    const int other = in_a ? s - left : left;
    int lo_b = 0, hi_b = other;   // count of the other half's elements before x
    while (lo_b < hi_b) {
      const int m = (lo_b + hi_b) >> 1;
      const int at = in_a ? (sign > 0 ? left + m : s - 1 - m) : (sign > 0 ? m : left - 1 - m);
      const bool before = in_a ? dloc[at] < x : dloc[at] <= x;
      if (before) lo_b = m + 1; else hi_b = m;
    }
    const int position = rank + lo_b;
    if (position < s) perm[position] = i;
  }
  __syncthreads();
  double dmax = 0.0;
  for (int p = t; p < s; p += kEigThreads) {
    pole[p] = dloc[perm[p]];
    zsort[p] = zloc[perm[p]];
    dmax = fmax(dmax, fabs(pole[p]));
  }
  dmax = wave_max(dmax);
  if ((t & 63) == 0) red[t >> 6] = dmax;
  __syncthreads();
  dmax = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  const double tol = 8.0 * kEps * fmax(dmax, rho);
  // small coupling components deflate at once; the rest are candidates for the close-pole test
  int mine = 0;
  for (int e = 0; e < kPerThread; ++e) {
    const int p = kPerThread * t + e;
    const int d = p < s && rho * fabs(zsort[p]) <= tol ? 1 : 0;
    if (p < s) defl[p] = d;
    mine += p < s && !d ? 1 : 0;
  }
  int candidates;
  int at = block_exclusive_scan(mine, scan, &candidates);
  for (int e = 0; e < kPerThread; ++e) {
    const int p = kPerThread * t + e;
    if (p < s && !defl[p]) cand[at++] = p;
  }
  __syncthreads();
  // one thread: close poles, in order; a pair is rotated so one of its components vanishes when
  // |(d_p - d_prev) c s| <= tol, which is tested after a cheaper bound rules most pairs out
  if (t == 0) {
    int nrot = 0, prev = -1;
    for (int c = 0; c < candidates; ++c) {
      const int p = cand[c];
      if (prev >= 0) {
        const double zp = zsort[p], zq = zsort[prev], gap = fabs(pole[p] - pole[prev]);
        if (gap * fabs(zp * zq) <= 2.0 * tol * (zp * zp + zq * zq)) {
          const double r = hypot(zp, zq);
          const double cs = zp / r, sn = -zq / r;
          if (fabs((pole[p] - pole[prev]) * cs * sn) <= tol) {
            zsort[p] = r;
            zsort[prev] = 0.0;
            rot[nrot++] = Rotation{perm[prev], perm[p], cs, sn};
            const double dp = pole[prev] * cs * cs + pole[p] * sn * sn;
            pole[p] = pole[prev] * sn * sn + pole[p] * cs * cs;
            pole[prev] = dp;
            defl[prev] = 1;
          }
        }
      }
      prev = p;
    }
    s_rot = nrot;
  }
  __syncthreads();
  const int nrot = s_rot;
  mine = 0;
  for (int e = 0; e < kPerThread; ++e) {
    const int p = kPerThread * t + e;
    mine += p < s && !defl[p] ? 1 : 0;
  }
  int k;
  at = block_exclusive_scan(mine, scan, &k);
  for (int e = 0; e < kPerThread; ++e) {
    const int p = kPerThread * t + e;
    if (p < s && !defl[p]) keep[at++] = p;
  }
  __syncthreads();
  // (c) the deflating rotations on Q's columns, in order; each thread owns its rows
  for (int r = lo + t; r < hi; r += kEigThreads)
    for (int q = 0; q < nrot; ++q) {
      double *ca = Q + r + int64_t(lo + rot[q].a) * n, *cb = Q + r + int64_t(lo + rot[q].b) * n;
      const double x = *ca, y = *cb;
      *ca = rot[q].c * x + rot[q].s * y;
      *cb = -rot[q].s * x + rot[q].c * y;
    }
  // what the next launches read
  for (int j = t; j < k; j += kEigThreads) {
    const double z = zsort[keep[j]];
    sc.dk[lo + j] = pole[keep[j]];
    sc.zk[lo + j] = z;
    sc.z2[lo + j] = z * z;
    sc.keep[lo + j] = keep[j];
  }
  for (int p = t; p < s; p += kEigThreads) {
    sc.pole[lo + p] = pole[p];
    sc.perm[lo + p] = perm[p];
    sc.defl[lo + p] = defl[p];
  }
  if (t == 0) {
    sc.rho[lo] = rho;
    maps.unitk[lo] = k;
  }
}

// Secular roots, one wave per root j of a unit: lane l holds the kept poles l, l + 64, ... in
// registers; every evaluation of f(x) = 1 + psi(x) + phi(x) (x measured from the root's origin
// pole; psi sums over the poles 0..j, phi over the rest) sums them lane-wise and then over the
// wave, so every lane holds the same f and takes the same step.  The root starts from the nearer
// pole of its interval; each step models psi and phi by one pole each, matching value and slope
// at the iterate, and takes the root of that model (the rational "middle way" step); steps that
// leave the bracket bisect it.  Converged when f is within its own rounding error, the bracket is
// at the resolution of x, or a step no longer changes x.

constexpr int kPoleChunks = kEigMaxN / 64;

// root blocks (one per workgroup, kEigWaves roots) of a matrix of order n at a level
__host__ __device__ inline int dc_root_blocks(int n, int width) {
  const int units = dc_units(n, width);
  if (units == 0) return 0;
  const int last = n - (units - 1) * 2 * width;
  return (units - 1) * ((2 * width + kEigWaves - 1) / kEigWaves) + (last + kEigWaves - 1) / kEigWaves;
}

struct Secular {
  double f, psi, phi, dpsi, dphi;
};

__device__ inline Secular secular_wave(const double (&d)[kPoleChunks], const double (&w2)[kPoleChunks], int k, int j,
                                       double origin, double rho, double x) {
  const int lane = threadIdx.x & 63;
  double psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0;
#pragma unroll
  for (int m = 0; m < kPoleChunks; ++m) {
    const int i = lane + 64 * m;
    if (i < k) {
      const double inv = reciprocal((d[m] - origin) - x);
      const double w = w2[m] * inv;
      if (i <= j) {
        psi += w;
        dpsi += w * inv;
      } else {
        phi += w;
        dphi += w * inv;
      }
    }
  }
  psi = wave_sum(psi);
  dpsi = wave_sum(dpsi);
  phi = wave_sum(phi);
  dphi = wave_sum(dphi);
  return Secular{1.0 + rho * (psi + phi), rho * psi, rho * phi, rho * dpsi, rho * dphi};
}

__global__ __launch_bounds__(kEigThreads) void dc_secular_kernel(const Batch *__restrict__ batch, double *__restrict__ work, int width) {
  Located at;
  if (!locate(batch, static_cast<int>(blockIdx.x), [width](int order) { return dc_root_blocks(order, width); }, &at))
    return;
  const Problem &pr = at.pr;
  const int n = pr.n;
  const int units = dc_units(n, width), full = (2 * width + kEigWaves - 1) / kEigWaves;
  int local = static_cast<int>(blockIdx.x) - at.first, unit = local / full;
  if (unit >= units - 1) {
    unit = units - 1;
    local -= (units - 1) * full;
  } else {
    local -= unit * full;
  }
  const int lo = unit * 2 * width;
  double *const base = work + pr.work;
  const DcMaps maps = dc_maps(base, n);
  const DcScratch sc = dc_scratch(base, n);
  const int k = maps.unitk[lo];
  const int j = local * kEigWaves + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (j >= k) return;
  const double rho = sc.rho[lo];
  double d[kPoleChunks], w2[kPoleChunks];
#pragma unroll
  for (int m = 0; m < kPoleChunks; ++m) {
    const int i = lane + 64 * m;
    d[m] = i < k ? sc.dk[lo + i] : 0.0;
    w2[m] = i < k ? sc.z2[lo + i] : 0.0;
  }
  const double dj = sc.dk[lo + j];
  int o = j;
  double a = 0.0, bnd;
  if (j + 1 < k) {
    const double gap = sc.dk[lo + j + 1] - dj;
    if (secular_wave(d, w2, k, j, dj, rho, 0.5 * gap).f > 0.0) {
      bnd = 0.5 * gap;
    } else {
      o = j + 1;
      a = -0.5 * gap;
      bnd = 0.0;
    }
  } else {
    double zs = 0.0;
#pragma unroll
    for (int m = 0; m < kPoleChunks; ++m) zs += w2[m];
    bnd = rho * wave_sum(zs);
  }
  const double origin = sc.dk[lo + o];
  const double lo_pole = dj - origin;
  const double hi_pole = j + 1 < k ? sc.dk[lo + j + 1] - origin : 0.0;
  double x = 0.5 * (a + bnd);
  for (int it = 0; it < 80; ++it) {
    const Secular v = secular_wave(d, w2, k, j, origin, rho, x);
    if (v.f > 0.0) bnd = x; else a = x;
    if (fabs(v.f) <= 8.0 * kEps * k * (1.0 + fabs(v.psi) + fabs(v.phi)) ||
        bnd - a <= 4.0 * kEps * fmax(fabs(a), fabs(bnd)))
      break;
    const double dlo = lo_pole - x;
    double eta;
    if (j + 1 < k) {
      const double dhi = hi_pole - x;
      const double S = dlo * dlo * v.dpsi, T = dhi * dhi * v.dphi;
      const double c = v.f - dlo * v.dpsi - dhi * v.dphi;
      const double B = c * (dlo + dhi) + S + T, C = c * dlo * dhi + S * dhi + T * dlo;
      const double root = sqrt(fmax(B * B - 4.0 * c * C, 0.0));
      if (B >= 0.0)
        eta = 2.0 * C / (B + root);
      else if (c != 0.0)
        eta = (B - root) / (2.0 * c);
      else
        eta = C / B;
    } else {
      const double c = v.f - dlo * v.dpsi;
      eta = c != 0.0 ? dlo + dlo * dlo * v.dpsi / c : 0.5 * (bnd - a);
    }
    double next = x + eta;
    if (!(next > a && next < bnd)) next = 0.5 * (a + bnd);
    if (fabs(next - x) <= 2.0 * kEps * fabs(x)) {
      x = next;
      break;
    }
    x = next;
  }
  if (lane == 0) {
    sc.org[lo + j] = o;
    sc.tau[lo + j] = x;
  }
}

__global__ __launch_bounds__(kEigThreads) void dc_finish_kernel(const Batch *__restrict__ batch, double *__restrict__ values,
                                                                 double *__restrict__ work, int width) {
  const UnitRef u = find_unit(batch, width, static_cast<int>(blockIdx.x));
  if (u.b < 0) return;
  const Problem &pr = u.pr;
  const int n = pr.n, lo = u.lo;
  const int mid = lo + width, hi = min(lo + 2 * width, n), s = hi - lo;
  if (mid >= n) return;   // a copy: its maps are written
  double *const base = work + pr.work;
  const DcMaps maps = dc_maps(base, n);
  const DcScratch sc = dc_scratch(base, n);
  double *lam = values + pr.values;
  double *__restrict__ U = base + 2 * int64_t(n) * n + int64_t(lo) * min(2 * width, n);
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  __shared__ double dk[kEigMaxN], zk[kEigMaxN], tau[kEigMaxN], zh[kEigMaxN], val[kEigMaxN];
  __shared__ int32_t org[kEigMaxN], src[kEigMaxN];
  __shared__ int scan[kEigThreads + 1];
  const int k = maps.unitk[lo];
  const double rho = sc.rho[lo];
  const double *off = base + 3 * int64_t(n) * n + n;
  const double sign = off[mid - 1] < 0.0 ? -1.0 : 1.0;
  for (int j = t; j < k; j += kEigThreads) {
    dk[j] = sc.dk[lo + j];
    zk[j] = sc.zk[lo + j];
    tau[j] = sc.tau[lo + j];
    org[j] = sc.org[lo + j];
  }
  // slots: 0..k-1 the roots, then the deflated poles in order
  int mine = 0;
  for (int e = 0; e < kPerThread; ++e) {
    const int p = kPerThread * t + e;
    mine += p < s && sc.defl[lo + p] ? 1 : 0;
  }
  int deflated;
  int at = k + block_exclusive_scan(mine, scan, &deflated);
  for (int e = 0; e < kPerThread; ++e) {
    const int p = kPerThread * t + e;
    if (p < s && sc.defl[lo + p]) {
      val[at] = sc.pole[lo + p];
      src[at] = sc.perm[lo + p];
      ++at;
    }
  }
  __syncthreads();
  // (e) Gu-Eisenstat: zh_i^2 = (lam_i - d_i)/rho prod_{j != i} (lam_j - d_i)/(d_j - d_i)
  for (int i = t; i < k; i += kEigThreads) {
    double w = ((dk[org[i]] - dk[i]) + tau[i]) / rho;
    for (int j = 0; j < k; ++j)
      if (j != i) w *= ((dk[org[j]] - dk[i]) + tau[j]) / (dk[j] - dk[i]);
    zh[i] = copysign(sqrt(fabs(w)), zk[i]);
    val[i] = dk[org[i]] + tau[i];
    src[i] = sc.perm[lo + sc.keep[lo + i]];
  }
  __syncthreads();
  // (f) U[:, j] = zh / (d - lam_j), normalised (k x k, column-major): one wave per column
  for (int j = wave; j < k; j += kEigWaves) {
    const double dj = dk[org[j]], tj = tau[j];
    double norm = 0.0;
    for (int i = lane; i < k; i += 64) {
      const double v = zh[i] / ((dk[i] - dj) - tj);
      norm += v * v;
    }
    const double inv = 1.0 / sqrt(wave_sum(norm));
    for (int i = lane; i < k; i += 64) U[i + int64_t(j) * k] = inv * zh[i] / ((dk[i] - dj) - tj);
  }
  // (h) ascending rank of every slot (ties by slot): its final column and eigenvalue
  for (int x = t; x < s; x += kEigThreads) {
    int r = 0;
    for (int y = 0; y < s; ++y) r += (val[y] < val[x] || (val[y] == val[x] && y < x)) ? 1 : 0;
    r = sign > 0 ? r : s - 1 - r;
    lam[lo + r] = sign * val[x];
    maps.colsrc[lo + x] = src[x];
    maps.dst[lo + x] = r;
  }
}

// (i) the unit's output columns: one workgroup per 64 x 64 tile of (rows, output columns).  Copies
//     move over directly; root columns are the product Q[:, colsrc[0..k)] U on FP64 MFMA
//     (16x16x4: four waves, each a 32 x 32 quarter of the tile), with the next 16 columns of the
//     sum loaded into registers while the current ones are multiplied from LDS.

constexpr int kMergeTile = 64;
constexpr int kMergeK = 16;

__host__ __device__ inline int merge_tiles(int s) {
  const int per = (s + kMergeTile - 1) / kMergeTile;
  return per * per;
}

// workgroups of a matrix of order n at a level: full units, then the last
__host__ __device__ inline int dc_tiles(int n, int width) {
  const int units = dc_units(n, width);
  if (units == 0) return 0;
  const int last = n - (units - 1) * 2 * width;
  return (units - 1) * merge_tiles(2 * width) + merge_tiles(last);
}

__global__ __launch_bounds__(kEigThreads) void dc_gemm_kernel(const Batch *__restrict__ batch, double *__restrict__ work, int width,
                                                               int parity) {
  Located at;
  if (!locate(batch, static_cast<int>(blockIdx.x), [width](int order) { return dc_tiles(order, width); }, &at)) return;
  const Problem &pr = at.pr;
  const int n = pr.n, first = at.first;
  double *const base = work + pr.work;
  const double *__restrict__ Q = base + (parity ? int64_t(n) * n : 0);
  double *__restrict__ Qt = base + (parity ? 0 : int64_t(n) * n);
  const DcMaps maps = dc_maps(base, n);
  const int units = dc_units(n, width), full = merge_tiles(2 * width);
  int local = static_cast<int>(blockIdx.x) - first, unit = local / full;
  if (unit >= units - 1) {
    unit = units - 1;
    local -= (units - 1) * full;
  } else {
    local -= unit * full;
  }
  const int lo = unit * 2 * width, s = min(2 * width, n - lo);
  const int per = (s + kMergeTile - 1) / kMergeTile;
  const int r0 = (local / per) * kMergeTile, j0 = (local % per) * kMergeTile;
  const int k = maps.unitk[lo];
  const double *__restrict__ U = base + 2 * int64_t(n) * n + int64_t(lo) * min(2 * width, n);
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  __shared__ int32_t src[kEigMaxN];
  __shared__ double As[2][kMergeK][kMergeTile + 2], Bs[2][kMergeK][kMergeTile + 2];
  __shared__ double Cs[kMergeTile][kMergeTile + 1];
  for (int j = t; j < k; j += kEigThreads) src[j] = maps.colsrc[lo + j];
  // copies among this tile's output columns
  for (int e = t; e < kMergeTile * kMergeTile; e += kEigThreads) {
    const int rr = e & (kMergeTile - 1), j = j0 + e / kMergeTile, r = r0 + rr;
    if (j >= k && j < s && r < s)
      Qt[lo + r + int64_t(lo + maps.dst[lo + j]) * n] = Q[lo + r + int64_t(lo + maps.colsrc[lo + j]) * n];
  }
  if (j0 >= k) return;
  __syncthreads();
  // register staging: A element (row rr, sum index kq + 4 q), B element (sum index t % 16, column jq + 16 q)
  const int rr = t & 63, kq = t >> 6, kb = t & 15, jq = t >> 4;
  double ra[4], rb[4];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int kk = k0 + kq + 4 * q;
      ra[q] = (r0 + rr < s && kk < k) ? Q[lo + r0 + rr + int64_t(lo + src[kk]) * n] : 0.0;
      const int kk2 = k0 + kb, j = j0 + jq + 16 * q;
      rb[q] = (kk2 < k && j < k) ? U[kk2 + int64_t(j) * k] : 0.0;
    }
  };
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32, lr = lane & 15, lk = lane >> 4;
  d4 acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = d4{0.0, 0.0, 0.0, 0.0};
  fetch(0);
  int buf = 0;
  for (int k0 = 0; k0 < k; k0 += kMergeK) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      As[buf][kq + 4 * q][rr] = ra[q];
      Bs[buf][kb][jq + 16 * q] = rb[q];
    }
    __syncthreads();
    if (k0 + kMergeK < k) fetch(k0 + kMergeK);
#pragma unroll
    for (int k4 = 0; k4 < kMergeK; k4 += 4) {
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
  }
  // through LDS (column-major) so each output column is written as 64 consecutive rows
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) Cs[wc + 16 * bj + lr][wr + 16 * bi + 4 * r + lk] = acc[bi][bj][r];
  __syncthreads();
  for (int e = t; e < kMergeTile * kMergeTile; e += kEigThreads) {
    const int row = e & (kMergeTile - 1), jj = e / kMergeTile, j = j0 + jj, r = r0 + row;
    if (j < k && r < s) Qt[lo + r + int64_t(lo + maps.dst[lo + j]) * n] = Cs[jj][row];
  }
}

// ---------------------------------------------------------------------------------------------
// 3. Back-transformation Y = H_0 H_1 ... H_{n-2} Z, with the reflectors in blocks of kRefl:
//    H_j ... H_{j+m-1} = I - V T V^T (compact WY form, T upper triangular).  One workgroup per
//    block forms T from the block's Gram matrix V^T V; then one workgroup per 16-column panel of Z
//    keeps the panel in registers and applies the blocks last to first, each as W = V^T Y,
//    W = T W, Y -= V W on FP64 MFMA.  Wave w of eight holds the 16-row blocks w, w + 8, ... of the panel in
//    the MFMA result layout (lane l: rows 4 r + l / 16 of a block, column l % 16), which is also
//    the layout of the 4 x 16 right-hand operand of W = V^T Y.  Eigenvector c is written as row c
//    of the output.

constexpr int kRefl = 32;
constexpr int kBackThreads = 512;
constexpr int kBackWaves = kBackThreads / 64;
constexpr int kRowBlocks = kEigMaxN / (16 * kBackWaves);   // 16-row blocks per wave

__host__ __device__ inline int reflector_blocks(int n) { return n > 1 ? (n - 1 + kRefl - 1) / kRefl : 0; }

__global__ __launch_bounds__(kEigThreads) void block_t_kernel(const Batch *__restrict__ batch, const double *__restrict__ matrices,
                                                               double *__restrict__ work) {
  Located at;
  if (!locate(batch, static_cast<int>(blockIdx.x), [](int order) { return reflector_blocks(order); }, &at)) return;
  const Problem &pr = at.pr;
  const int n = pr.n, q = static_cast<int>(blockIdx.x) - at.first;
  const int j = q * kRefl, m = min(kRefl, n - 1 - j);
  const double *__restrict__ A = matrices + pr.matrix;
  const double *tau = work + pr.work + 3 * int64_t(n) * n + 2 * n;
  double *T = work + pr.work + 3 * int64_t(n) * n + 7 * n + int64_t(q) * kRefl * kRefl;
  __shared__ double Vs[64][kRefl + 1];
  __shared__ double G[kRefl][kRefl + 1], Ts[kRefl][kRefl + 1];
  const int t = threadIdx.x, ga = t >> 3, gc = (t & 7) * 4;
  // G = V^T V over rows j+1..: reflector a of the block is nonzero from row j + a + 1 (where it is 1)
  double g[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i0 = j + 1; i0 < n; i0 += 64) {
    __syncthreads();
    for (int e = t; e < 64 * kRefl; e += kEigThreads) {
      const int row = e & 63, a = e >> 6, i = i0 + row;
      Vs[row][a] = (a < m && i < n && i > j + a) ? A[i + int64_t(j + a) * n] : 0.0;
    }
    __syncthreads();
    for (int row = 0; row < 64; ++row) {
      const double va = Vs[row][ga];
#pragma unroll
      for (int u = 0; u < 4; ++u) g[u] += va * Vs[row][gc + u];
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) G[ga][gc + u] = g[u];
  __syncthreads();
  // T column by column (forward order): T[i][i] = tau_i, T[a][i] = -tau_i sum_c T[a][c] G[c][i];
  // thread a owns row a of T
  if (t < kRefl) {
    const int a = t;
    for (int i = 0; i < kRefl; ++i) {
      double value = 0.0;
      if (i < m && a <= i) {
        const double ti = tau[j + i];
        if (a == i) {
          value = ti;
        } else {
          double sum = 0.0;
          for (int c = a; c < i; ++c) sum += Ts[a][c] * G[c][i];
          value = -ti * sum;
        }
      }
      Ts[a][i] = value;
    }
    for (int i = 0; i < kRefl; ++i) T[a * kRefl + i] = Ts[a][i];
  }
}

__global__ __launch_bounds__(kBackThreads) void back_kernel(const Batch *__restrict__ batch, const double *__restrict__ matrices,
                                                            double *__restrict__ vectors, const double *__restrict__ work) {
  // the problem this panel belongs to
  Located at;
  if (!locate(batch, static_cast<int>(blockIdx.x), [](int order) { return (order + kPanel - 1) / kPanel; }, &at)) return;
  const Problem &pr = at.pr;
  const int n = pr.n;
  const int c0 = (static_cast<int>(blockIdx.x) - at.first) * kPanel;
  const double *__restrict__ A = matrices + pr.matrix;
  const double *__restrict__ Z = work + pr.work + (dc_levels(n) & 1 ? int64_t(n) * n : 0);
  const double *__restrict__ Tall = work + pr.work + 3 * int64_t(n) * n + 7 * n;
  __shared__ double Wp[kBackWaves][kRefl][kPanel + 1];
  __shared__ double Wq[kRefl][kPanel + 1], Wr[kRefl][kPanel + 1];
  __shared__ double Ts[kRefl][kRefl + 1];
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int col = c0 + (lane & 15);
  double y[kRowBlocks][4];
  const int lk0 = lane >> 4;
#pragma unroll
  for (int mb = 0; mb < kRowBlocks; ++mb) {
    const int base_row = 16 * (kBackWaves * mb + wave);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = base_row + 4 * r + lk0;
      y[mb][r] = (row < n && col < n) ? Z[row + int64_t(col) * n] : 0.0;
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  for (int q = reflector_blocks(n) - 1; q >= 0; --q) {
    const int j = q * kRefl, m = min(kRefl, n - 1 - j);
    // the lane's row and column indices, recomputed in every block: kept live across the loop
    // (the compiler's choice otherwise) they do not fit beside the panel and spill
    int lane_q = lane;
    asm volatile("" : "+v"(lane_q));
    const int lr = lane_q & 15, lk = lane_q >> 4;
    __syncthreads();
    for (int e = t; e < kRefl * kRefl; e += kBackThreads) Ts[e / kRefl][e % kRefl] = Tall[int64_t(q) * kRefl * kRefl + e];
    // W = V^T Y: this wave's rows, both halves of the block's reflectors
    d4 w0 = d4{0.0, 0.0, 0.0, 0.0}, w1 = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int mb = 0; mb < kRowBlocks; ++mb) {
      const int base_row = 16 * (kBackWaves * mb + wave);
      if (base_row + 16 <= j + 1 || base_row >= n) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = base_row + 4 * r + lk, a0 = lr, a1 = 16 + lr;
        const double v0 = (a0 < m && i < n && i > j + a0) ? A[i + int64_t(j + a0) * n] : 0.0;
        const double v1 = (a1 < m && i < n && i > j + a1) ? A[i + int64_t(j + a1) * n] : 0.0;
        w0 = __builtin_amdgcn_mfma_f64_16x16x4f64(v0, y[mb][r], w0, 0, 0, 0);
        w1 = __builtin_amdgcn_mfma_f64_16x16x4f64(v1, y[mb][r], w1, 0, 0, 0);
      }
      // one row block's operands at a time: hoisting the next block's loads would spill
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      Wp[wave][4 * r + lk][lr] = w0[r];
      Wp[wave][16 + 4 * r + lk][lr] = w1[r];
    }
    __syncthreads();
    for (int e = t; e < kRefl * kPanel; e += kBackThreads) {
      const int a = e / kPanel, c = e % kPanel;
      double sum = 0.0;
#pragma unroll
      for (int w = 0; w < kBackWaves; ++w) sum += Wp[w][a][c];
      Wq[a][c] = sum;
    }
    __syncthreads();
    // W = T W (T upper triangular)
    for (int e = t; e < kRefl * kPanel; e += kBackThreads) {
      const int a = e / kPanel, c = e % kPanel;
      double sum = 0.0;
      for (int i = a; i < kRefl; ++i) sum += Ts[a][i] * Wq[i][c];
      Wr[a][c] = sum;
    }
    __syncthreads();
    // Y -= V W
#pragma unroll
    for (int mb = 0; mb < kRowBlocks; ++mb) {
      const int base_row = 16 * (kBackWaves * mb + wave);
      if (base_row + 16 <= j + 1 || base_row >= n) continue;
      d4 acc = d4{0.0, 0.0, 0.0, 0.0};
      const int i = base_row + lr;
#pragma unroll
      for (int s4 = 0; s4 < kRefl / 4; ++s4) {
        const int a = 4 * s4 + lk;
        const double v = (a < m && i < n && i > j + a) ? A[i + int64_t(j + a) * n] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(v, Wr[a][lr], acc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) y[mb][r] -= acc[r];
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  double *out = vectors + pr.matrix;
  if (col < n)
#pragma unroll
    for (int mb = 0; mb < kRowBlocks; ++mb) {
      const int base_row = 16 * (kBackWaves * mb + wave);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = base_row + 4 * r + lk0;
        if (row < n) out[int64_t(col) * n + row] = y[mb][r];
      }
      __builtin_amdgcn_sched_barrier(0);
    }
}

// The batch table, written once per batch into the workspace, where every launch reads it (a
// kernel argument of its size would be copied for each of the many launches of a batch).  The same
// launch zeroes the two eigenvector buffers of every problem (2 n^2 doubles at its workspace, unused
// until the leaves): the merges read whole diagonal blocks of them, whose parts outside the solved
// sub-blocks must be zero.  Grid (kZeroBlocks, problems), coalesced 16-byte stores.
constexpr int kZeroBlocks = 64;
constexpr int kZeroThreads = 256;

__global__ __launch_bounds__(kZeroThreads) void store_batch_kernel(Batch batch, Batch *__restrict__ table,
                                                                  double *__restrict__ work) {
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *table = batch;
  const Problem &pr = batch.p[blockIdx.y];
  double *const Q = work + pr.work;
  const int64_t count = 2 * int64_t(pr.n) * pr.n;
  const int64_t stride = int64_t(kZeroBlocks) * kZeroThreads;
  const int64_t first = int64_t(blockIdx.x) * kZeroThreads + threadIdx.x;
  if ((reinterpret_cast<uintptr_t>(Q) & 15) == 0) {
    double2 *const Q2 = reinterpret_cast<double2 *>(Q);
    for (int64_t i = first; i < count / 2; i += stride) Q2[i] = double2{0.0, 0.0};
    if (first == 0 && (count & 1)) Q[count - 1] = 0.0;
  } else {
    for (int64_t i = first; i < count; i += stride) Q[i] = 0.0;
  }
}

constexpr int64_t kTableDoubles = (sizeof(Batch) + 255) / 256 * 32;   // one table, 256-byte aligned

}  // namespace

extern "C" int64_t tt_eigh_max_order(void) { return kEigMaxN; }

extern "C" int64_t tt_eigh_workspace(const tt_eigh_problem *problems, int64_t count) {
  if (count < 0 || (count > 0 && !problems)) {
    fail("tt_eigh_workspace: bad problem list");
    return -1;
  }
  int64_t total = kTableDoubles * ((count + kEigBatch - 1) / kEigBatch);
  for (int64_t b = 0; b < count; ++b) {
    const int64_t n = problems[b].n;
    total += 3 * n * n + 7 * n + (n >= 1 && n <= kEigMaxN ? int64_t(kRefl) * kRefl * reflector_blocks(static_cast<int>(n)) : 0);
  }
  return total;
}

namespace {

int eigh_batched(const tt_eigh_problem *problems, int64_t count, double *matrices, double *values, double *vectors,
                 double *workspace, void *stream) {
  if (count < 0 || (count > 0 && !problems)) return fail("tt_eigh_batched: bad problem list");
  if (count == 0) return 0;
  if (!matrices || !values || !vectors || !workspace) return fail("tt_eigh_batched: null buffer");
  for (int64_t b = 0; b < count; ++b) {
    const tt_eigh_problem &p = problems[b];
    if (p.n < 1 || p.n > kEigMaxN)
      return fail("tt_eigh_batched: matrix order " + std::to_string(p.n) + " outside 1.." + std::to_string(kEigMaxN));
    if (p.matrix_offset < 0 || p.value_offset < 0) return fail("tt_eigh_batched: negative offset");
  }
  // the batch tables first, then every problem's workspace
  int64_t work = kTableDoubles * ((count + kEigBatch - 1) / kEigBatch);
  for (int64_t first = 0; first < count; first += kEigBatch) {
    Batch *table = reinterpret_cast<Batch *>(workspace + kTableDoubles * (first / kEigBatch));
    Batch batch{};
    batch.count = static_cast<int32_t>(std::min<int64_t>(kEigBatch, count - first));
    int32_t panels = 0, leaves = 0;
    for (int32_t b = 0; b < batch.count; ++b) {
      const tt_eigh_problem &p = problems[first + b];
      batch.p[b] = Problem{p.matrix_offset, p.value_offset, work, static_cast<int32_t>(p.n), panels, leaves, 0};
      work += 3 * p.n * p.n + 7 * p.n + int64_t(kRefl) * kRefl * reflector_blocks(static_cast<int>(p.n));
      panels += static_cast<int32_t>((p.n + kPanel - 1) / kPanel);
      leaves += static_cast<int32_t>((p.n + kLeaf - 1) / kLeaf);
    }
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(store_batch_kernel, dim3(kZeroBlocks, batch.count), dim3(kZeroThreads), 0, s, batch, table,
                       workspace);
    TT_LAUNCHED();
    int64_t largest = 0;
    for (int32_t b = 0; b < batch.count; ++b) largest = std::max<int64_t>(largest, batch.p[b].n);
    // steps over many workgroups while the largest trailing matrix exceeds kRegOrder rows, then
    // the rest of every matrix in one workgroup each
    const int reg_from = static_cast<int>(std::max<int64_t>(0, largest - kRegOrder));
    for (int k = 0; k < reg_from; ++k) {
      int grid = 0;
      for (int32_t b = 0; b < batch.count; ++b) grid += step_blocks(batch.p[b].n, k);
      // the step's longest vectors are rows k..n-1; in timing mode each step launch is timed
      const int64_t rows = largest - k;
      const dim3 blocks(grid), threads(kStepThreads);
      int rc;
      if (rows <= 4 * 64) {
        rc = ttdev::launch(tridiag_step_kernel<4>, blocks, threads, 0, s, table, matrices, workspace, k);
      } else if (rows <= 8 * 64) {
        rc = ttdev::launch(tridiag_step_kernel<8>, blocks, threads, 0, s, table, matrices, workspace, k);
      } else if (rows <= 16 * 64) {
        rc = ttdev::launch(tridiag_step_kernel<16>, blocks, threads, 0, s, table, matrices, workspace, k);
      } else {
        rc = ttdev::launch(tridiag_step_kernel<32>, blocks, threads, 0, s, table, matrices, workspace, k);
      }
      if (rc) return rc;
      TT_LAUNCHED();
    }
    hipLaunchKernelGGL(tridiag_reg_kernel, dim3(batch.count), dim3(kRegThreads), 0, s, table, matrices, workspace,
                       reg_from);
    TT_LAUNCHED();
    hipLaunchKernelGGL(leaf_kernel, dim3((leaves + kEigWaves - 1) / kEigWaves), dim3(kEigThreads), 0, s, table, values,
                       workspace, leaves);
    TT_LAUNCHED();
    for (int width = kLeaf, level = 0; width < largest; width *= 2, ++level) {
      int units = 0, tiles = 0;
      for (int32_t b = 0; b < batch.count; ++b) {
        units += dc_units(batch.p[b].n, width);
        tiles += dc_tiles(batch.p[b].n, width);
      }
      int roots = 0;
      for (int32_t b = 0; b < batch.count; ++b) roots += dc_root_blocks(batch.p[b].n, width);
      hipLaunchKernelGGL(dc_deflate_kernel, dim3(units), dim3(kEigThreads), 0, s, table, values, workspace, width, level & 1);
      TT_LAUNCHED();
      hipLaunchKernelGGL(dc_secular_kernel, dim3(roots), dim3(kEigThreads), 0, s, table, workspace, width);
      TT_LAUNCHED();
      hipLaunchKernelGGL(dc_finish_kernel, dim3(units), dim3(kEigThreads), 0, s, table, values, workspace, width);
      TT_LAUNCHED();
      hipLaunchKernelGGL(dc_gemm_kernel, dim3(tiles), dim3(kEigThreads), 0, s, table, workspace, width, level & 1);
      TT_LAUNCHED();
    }
    int blocks = 0;
    for (int32_t b = 0; b < batch.count; ++b) blocks += reflector_blocks(batch.p[b].n);
    if (blocks > 0) {
      hipLaunchKernelGGL(block_t_kernel, dim3(blocks), dim3(kEigThreads), 0, s, table, matrices, workspace);
      TT_LAUNCHED();
    }
    hipLaunchKernelGGL(back_kernel, dim3(panels), dim3(kBackThreads), 0, s, table, matrices, vectors, workspace);
    TT_LAUNCHED();
  }
  return 0;
}

}  // namespace

extern "C" int tt_eigh_batched(const tt_eigh_problem *problems, int64_t count, double *matrices, double *values,
                               double *vectors, double *workspace, void *stream) {
  return eigh_batched(problems, count, matrices, values, vectors, workspace, stream);
}

// The workspace of tt_eigh_batched_static: one array per device, in the code object's device
// memory, held until the process ends.  In host code `scratch` only names the symbol; its device
// address comes from hipGetSymbolAddress, once per device (each device holds its own copy).
static __device__ __align__(256) double scratch[TT_EIGH_STATIC_SCRATCH_DOUBLES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
double *scratch_address[kMaxDevices] = {};

// The device address of the current device's scratch array, looked up on the first call there.
int current_scratch(double **out) {
  int device = -1;
  TT_HIP(hipGetDevice(&device));
  if (device < 0 || device >= kMaxDevices)
    return fail("tt_eigh_batched_static: device " + std::to_string(device) + " beyond the " +
                std::to_string(kMaxDevices) + " the library tracks");
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_address[device]) {
    void *address = nullptr;
    TT_HIP(hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)));
    scratch_address[device] = static_cast<double *>(address);
  }
  *out = scratch_address[device];
  return 0;
}

}  // namespace

extern "C" int tt_eigh_batched_static(const tt_eigh_problem *problems, int64_t count, double *matrices,
                                      double *values, double *vectors, void *stream) {
  const int64_t need = tt_eigh_workspace(problems, count);
  if (need < 0) return -1;
  if (need > TT_EIGH_STATIC_SCRATCH_DOUBLES)
    return fail("tt_eigh_batched_static: the batch needs " + std::to_string(need) +
                " doubles of workspace, more than the " + std::to_string(TT_EIGH_STATIC_SCRATCH_DOUBLES) +
                " of the library's static scratch; pass a workspace to tt_eigh_batched instead");
  if (count == 0) return 0;
  double *workspace = nullptr;
  if (int rc = current_scratch(&workspace)) return rc;
  return eigh_batched(problems, count, matrices, values, vectors, workspace, stream);
}

