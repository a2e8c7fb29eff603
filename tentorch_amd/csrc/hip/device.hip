// The MI355X device library behind include/tentorch_amd.h: descriptor-table upload, the
// gather-scale-scatter leg-move kernel, the grouped block GEMM on FP64/FP32 MFMA, fixed-order
// reductions and hipGraph capture.  gfx950 only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/tentorch_amd.h"

namespace {

thread_local std::string g_error;

int fail(const std::string &message) {
  g_error = message;
  return -1;
}

#define TT_HIP(call)                                                                           \
  do {                                                                                         \
    hipError_t status_ = (call);                                                               \
    if (status_ != hipSuccess)                                                                 \
      return fail(std::string(#call) + " failed: " + hipGetErrorString(status_));              \
  } while (0)

#define TT_LAUNCHED()                                                                          \
  do {                                                                                         \
    hipError_t status_ = hipGetLastError();                                                    \
    if (status_ != hipSuccess) return fail(std::string("kernel launch failed: ") + hipGetErrorString(status_)); \
  } while (0)

hipStream_t as_stream(void *stream) { return reinterpret_cast<hipStream_t>(stream); }

struct cdouble {
  double re, im;
};

// ---------------------------------------------------------------------------------------------
// Gather-scale-scatter

// One work unit of a leg move: a run of destination elements of one tile.
struct GssChunk {
  int32_t task;
  int32_t count;
  int64_t start;
};

constexpr int kGssThreads = 256;
constexpr int kGssChunk = 4096;

template <typename T>
struct Scale;
template <>
struct Scale<double> {
  static __device__ double zero() { return 0.0; }
  static __device__ double fma(double acc, double re, double, double x) { return acc + re * x; }
};
template <>
struct Scale<float> {
  static __device__ float zero() { return 0.0f; }
  static __device__ float fma(float acc, double re, double, float x) { return acc + static_cast<float>(re) * x; }
};
template <>
struct Scale<cdouble> {
  static __device__ cdouble zero() { return {0.0, 0.0}; }
  static __device__ cdouble fma(cdouble acc, double re, double im, cdouble x) {
    return {acc.re + (re * x.re - im * x.im), acc.im + (re * x.im + im * x.re)};
  }
};

// Each workgroup takes one chunk of one tile; threads stride the chunk in destination order, so
// consecutive lanes write consecutive destination elements along the tile's fastest axis.
template <typename T>
__global__ __launch_bounds__(kGssThreads) void gss_kernel(const GssChunk *__restrict__ chunks,
                                                          const tt_gss_task *__restrict__ tasks,
                                                          const tt_gss_source *__restrict__ sources,
                                                          const T *__restrict__ src, T *__restrict__ dst,
                                                          int32_t chunk_base) {
  const GssChunk chunk = chunks[chunk_base + blockIdx.x];
  const tt_gss_task &task = tasks[chunk.task];
  const int nd = task.ndim;
  uint32_t ext[TT_MAX_DIMS];
  int64_t dstr[TT_MAX_DIMS];
#pragma unroll
  for (int d = 0; d < TT_MAX_DIMS; ++d) {
    ext[d] = d < nd ? static_cast<uint32_t>(task.extent[d]) : 1u;
    dstr[d] = d < nd ? task.dst_stride[d] : 0;
  }
  const int ns = task.num_sources;
  const tt_gss_source *first = sources + task.first_source;
  for (int32_t at = threadIdx.x; at < chunk.count; at += kGssThreads) {
    uint32_t rem = static_cast<uint32_t>(chunk.start + at);
    uint32_t idx[TT_MAX_DIMS];
    int64_t doff = task.dst_offset;
#pragma unroll
    for (int d = TT_MAX_DIMS - 1; d >= 0; --d) {
      if (d < nd) {
        idx[d] = rem % ext[d];
        rem /= ext[d];
        doff += static_cast<int64_t>(idx[d]) * dstr[d];
      } else {
        idx[d] = 0;
      }
    }
    T acc = Scale<T>::zero();
    for (int s = 0; s < ns; ++s) {
      const tt_gss_source &source = first[s];
      int64_t soff = source.src_offset;
#pragma unroll
      for (int d = 0; d < TT_MAX_DIMS; ++d)
        if (d < nd) soff += static_cast<int64_t>(idx[d]) * source.src_stride[d];
      acc = Scale<T>::fma(acc, source.coef_re, source.coef_im, src[soff]);
    }
    dst[doff] = acc;
  }
}

}  // namespace

struct tt_gss_plan {
  tt_gss_task *tasks = nullptr;
  tt_gss_source *sources = nullptr;
  GssChunk *chunks = nullptr;
  int64_t num_tasks = 0, num_sources = 0, num_chunks = 0;
  int64_t written = 0, read = 0;
  // Chunk ranges of the launch groups of the per-block form.
  std::vector<std::pair<int32_t, int32_t>> groups;
};

extern "C" int tt_gss_plan_create(const tt_gss_task *tasks, int64_t num_tasks, const tt_gss_source *sources,
                                  int64_t num_sources, tt_gss_plan **plan) {
  if (!plan) return fail("tt_gss_plan_create: null output handle");
  if (num_tasks < 0 || num_sources < 0) return fail("tt_gss_plan_create: negative count");
  std::vector<GssChunk> chunks;
  std::vector<std::pair<int32_t, int32_t>> groups;
  int64_t written = 0, read = 0;
  int32_t current_group = -1;
  for (int64_t t = 0; t < num_tasks; ++t) {
    const tt_gss_task &task = tasks[t];
    if (task.ndim < 0 || task.ndim > TT_MAX_DIMS) return fail("tt_gss_plan_create: task rank out of range");
    if (task.num_sources < 0 || task.first_source < 0 || task.first_source + task.num_sources > num_sources)
      return fail("tt_gss_plan_create: task sources out of range");
    int64_t elements = 1;
    for (int d = 0; d < task.ndim; ++d) {
      if (task.extent[d] <= 0) return fail("tt_gss_plan_create: non-positive extent");
      elements *= task.extent[d];
    }
    if (elements >= (int64_t(1) << 31)) return fail("tt_gss_plan_create: a tile exceeds 2^31 elements");
    written += elements;
    read += elements * task.num_sources;
    if (task.group != current_group || groups.empty()) {
      groups.emplace_back(static_cast<int32_t>(chunks.size()), 0);
      current_group = task.group;
    }
    for (int64_t start = 0; start < elements; start += kGssChunk) {
      chunks.push_back(GssChunk{static_cast<int32_t>(t), static_cast<int32_t>(std::min<int64_t>(kGssChunk, elements - start)), start});
      groups.back().second++;
    }
  }
  auto *p = new tt_gss_plan();
  p->num_tasks = num_tasks;
  p->num_sources = num_sources;
  p->num_chunks = static_cast<int64_t>(chunks.size());
  p->written = written;
  p->read = read;
  p->groups = std::move(groups);
  auto upload = [&](void **where, const void *data, size_t bytes) -> int {
    if (!bytes) return 0;
    TT_HIP(hipMalloc(where, bytes));
    TT_HIP(hipMemcpy(*where, data, bytes, hipMemcpyHostToDevice));
    return 0;
  };
  if (upload(reinterpret_cast<void **>(&p->tasks), tasks, sizeof(tt_gss_task) * num_tasks) ||
      upload(reinterpret_cast<void **>(&p->sources), sources, sizeof(tt_gss_source) * num_sources) ||
      upload(reinterpret_cast<void **>(&p->chunks), chunks.data(), sizeof(GssChunk) * chunks.size())) {
    std::string keep = g_error;
    tt_gss_plan_destroy(p);
    return fail(keep);
  }
  *plan = p;
  return 0;
}

extern "C" int tt_gss_plan_destroy(tt_gss_plan *plan) {
  if (!plan) return 0;
  if (plan->tasks) (void)hipFree(plan->tasks);
  if (plan->sources) (void)hipFree(plan->sources);
  if (plan->chunks) (void)hipFree(plan->chunks);
  delete plan;
  return 0;
}

extern "C" int tt_gss_plan_counts(const tt_gss_plan *plan, int64_t *written, int64_t *read) {
  if (!plan) return fail("tt_gss_plan_counts: null plan");
  if (written) *written = plan->written;
  if (read) *read = plan->read;
  return 0;
}

namespace {
template <typename T>
int gss_launch(const tt_gss_plan *plan, const void *src, void *dst, hipStream_t stream, int32_t first, int32_t count) {
  if (count <= 0) return 0;
  hipLaunchKernelGGL(gss_kernel<T>, dim3(count), dim3(kGssThreads), 0, stream, plan->chunks, plan->tasks,
                     plan->sources, static_cast<const T *>(src), static_cast<T *>(dst), first);
  TT_LAUNCHED();
  return 0;
}

int gss_dispatch(const tt_gss_plan *plan, int dtype, const void *src, void *dst, hipStream_t stream, int32_t first,
                 int32_t count) {
  switch (dtype) {
    case TT_F64: return gss_launch<double>(plan, src, dst, stream, first, count);
    case TT_F32: return gss_launch<float>(plan, src, dst, stream, first, count);
    case TT_C128: return gss_launch<cdouble>(plan, src, dst, stream, first, count);
    default: return fail("gather-scale-scatter: unsupported dtype " + std::to_string(dtype));
  }
}
}  // namespace

extern "C" int tt_gss_execute(const tt_gss_plan *plan, int dtype, const void *src, void *dst, void *stream) {
  if (!plan) return fail("tt_gss_execute: null plan");
  if (plan->num_chunks == 0) return 0;
  if (!src || !dst) return fail("tt_gss_execute: null buffer");
  return gss_dispatch(plan, dtype, src, dst, as_stream(stream), 0, static_cast<int32_t>(plan->num_chunks));
}

extern "C" int tt_gss_execute_plain(const tt_gss_plan *plan, int dtype, const void *src, void *dst, void *stream) {
  if (!plan) return fail("tt_gss_execute_plain: null plan");
  if (plan->num_chunks == 0) return 0;
  if (!src || !dst) return fail("tt_gss_execute_plain: null buffer");
  for (const auto &g : plan->groups)
    if (int rc = gss_dispatch(plan, dtype, src, dst, as_stream(stream), g.first, g.second)) return rc;
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Grouped block GEMM on MFMA

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename TC>
struct Mfma;
// Result layout of one 16x16x4 MFMA (measured on gfx950): lane l holds column l % 16 in its four
// result registers r; the f64 form holds rows 4 r + l / 16, the f32 form rows 4 (l / 16) + r.
template <>
struct Mfma<double> {
  using acc_t = d4;
  static __device__ acc_t step(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  static __device__ int row(int lk, int r) { return 4 * r + lk; }
};
template <>
struct Mfma<float> {
  using acc_t = f4;
  static __device__ acc_t step(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ int row(int lk, int r) { return 4 * lk + r; }
};

// Output tile of one workgroup: which block, and the tile's first row and column.
struct GemmTile {
  int32_t group;
  int32_t m0;
  int32_t n0;
  int32_t pad;
};

constexpr int kBM = 64, kBN = 64, kBK = 16, kPad = 1;
constexpr int kGemmThreads = 256;

// A 64x64 tile per workgroup: four waves in a 2x2 grid, each wave 32x32 as 2x2 MFMA 16x16x4 tiles.
// Operands are staged through LDS (A transposed to k-major) with the next k-slab prefetched into
// registers while the current one is multiplied; the terms of a block are summed in list order.
// Lane l of an MFMA supplies A[l % 16][l / 16] and B[l / 16][l % 16]; its results follow Mfma::row.
template <typename TA, typename TB, typename TO, typename TC>
__global__ __launch_bounds__(kGemmThreads) void ggemm_kernel(const GemmTile *__restrict__ tiles,
                                                             const tt_gemm_group *__restrict__ groups,
                                                             const tt_gemm_term *__restrict__ terms,
                                                             const TA *__restrict__ A, const TB *__restrict__ B,
                                                             TO *__restrict__ C, int32_t plain_group,
                                                             int32_t plain_term, int32_t accumulate) {
  using M = Mfma<TC>;
  __shared__ TC As[2][kBK][kBM + kPad];
  __shared__ TC Bs[2][kBK][kBN + kPad];

  int32_t gi, m0, n0, t_begin, t_end;
  if (plain_group < 0) {
    const GemmTile tile = tiles[blockIdx.x];
    gi = tile.group;
    m0 = tile.m0;
    n0 = tile.n0;
    t_begin = groups[gi].first_term;
    t_end = t_begin + groups[gi].num_terms;
  } else {
    gi = plain_group;
    const int32_t tiles_n = static_cast<int32_t>((groups[gi].n + kBN - 1) / kBN);
    m0 = static_cast<int32_t>(blockIdx.x / tiles_n) * kBM;
    n0 = static_cast<int32_t>(blockIdx.x % tiles_n) * kBN;
    t_begin = plain_term < 0 ? 0 : plain_term;
    t_end = plain_term < 0 ? 0 : plain_term + 1;
  }
  const tt_gemm_group group = groups[gi];
  const int64_t m = group.m, n = group.n;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int li = lane & 15, lk = lane >> 4;

  typename M::acc_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = typename M::acc_t{0, 0, 0, 0};

  // Loader coordinates: A slab 64 rows x 16 k (4 k per thread), B slab 16 k x 64 cols (4 cols).
  const int a_row = tid >> 2, a_k = (tid & 3) * 4;
  const int b_k = tid >> 4, b_col = (tid & 15) * 4;

  // Flattened (term, k0) steps.
  int32_t term = t_begin;
  int64_t k0 = 0;
  // Skip empty terms.
  while (term < t_end && terms[term].k <= 0) ++term;

  TC ra[4], rb[4];
  auto load = [&](int32_t t, int64_t kb) {
    const tt_gemm_term tm = terms[t];
    const int64_t K = tm.k;
    const int64_t gr = m0 + a_row;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t kk = kb + a_k + j;
      ra[j] = (gr < m && kk < K) ? static_cast<TC>(A[tm.a_offset + gr * K + kk]) : TC(0);
    }
    const int64_t kr = kb + b_k;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t gc = n0 + b_col + j;
      rb[j] = (kr < K && gc < n) ? static_cast<TC>(B[tm.b_offset + kr * n + gc]) : TC(0);
    }
  };
  auto advance = [&](int32_t &t, int64_t &kb) {
    kb += kBK;
    if (kb >= terms[t].k) {
      kb = 0;
      ++t;
      while (t < t_end && terms[t].k <= 0) ++t;
    }
  };

  int buf = 0;
  if (term < t_end) load(term, k0);
  while (term < t_end) {
#pragma unroll
    for (int j = 0; j < 4; ++j) As[buf][a_k + j][a_row] = ra[j];
#pragma unroll
    for (int j = 0; j < 4; ++j) Bs[buf][b_k][b_col + j] = rb[j];
    __syncthreads();
    int32_t next_term = term;
    int64_t next_k = k0;
    advance(next_term, next_k);
    if (next_term < t_end) load(next_term, next_k);
#pragma unroll
    for (int kk = 0; kk < kBK / 4; ++kk) {
      const int k = kk * 4 + lk;
      TC af[2], bf[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) af[i] = As[buf][k][wm * 32 + i * 16 + li];
#pragma unroll
      for (int j = 0; j < 2; ++j) bf[j] = Bs[buf][k][wn * 32 + j * 16 + li];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = M::step(af[i], bf[j], acc[i][j]);
    }
    term = next_term;
    k0 = next_k;
    buf ^= 1;
  }

#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm * 32 + i * 16 + M::row(lk, r);
        const int64_t col = n0 + wn * 32 + j * 16 + li;
        if (row < m && col < n) {
          TO *out = C + group.c_offset + row * n + col;
          const TC value = acc[i][j][r];
          if (accumulate)
            *out = static_cast<TO>(static_cast<TC>(*out) + value);
          else
            *out = static_cast<TO>(value);
        }
      }
}

}  // namespace

struct tt_ggemm_plan {
  tt_gemm_group *groups = nullptr;
  tt_gemm_term *terms = nullptr;
  GemmTile *tiles = nullptr;
  int64_t num_groups = 0, num_terms = 0, num_tiles = 0;
  double flops = 0.0;
  std::vector<tt_gemm_group> host_groups;
};

extern "C" int tt_ggemm_plan_create(const tt_gemm_group *groups, int64_t num_groups, const tt_gemm_term *terms,
                                    int64_t num_terms, tt_ggemm_plan **plan) {
  if (!plan) return fail("tt_ggemm_plan_create: null output handle");
  if (num_groups < 0 || num_terms < 0) return fail("tt_ggemm_plan_create: negative count");
  struct Costed {
    double cost;
    GemmTile tile;
  };
  std::vector<Costed> costed;
  double flops = 0.0;
  for (int64_t g = 0; g < num_groups; ++g) {
    const tt_gemm_group &grp = groups[g];
    if (grp.m < 0 || grp.n < 0) return fail("tt_ggemm_plan_create: negative block shape");
    if (grp.num_terms < 0 || grp.first_term < 0 || grp.first_term + grp.num_terms > num_terms)
      return fail("tt_ggemm_plan_create: group terms out of range");
    if (grp.m >= (int64_t(1) << 31) || grp.n >= (int64_t(1) << 31)) return fail("tt_ggemm_plan_create: block too large");
    int64_t ksum = 0;
    for (int32_t t = grp.first_term; t < grp.first_term + grp.num_terms; ++t) {
      if (terms[t].k < 0) return fail("tt_ggemm_plan_create: negative k");
      ksum += terms[t].k;
    }
    flops += 2.0 * static_cast<double>(grp.m) * static_cast<double>(grp.n) * static_cast<double>(ksum);
    for (int64_t mt = 0; mt < grp.m; mt += kBM)
      for (int64_t nt = 0; nt < grp.n; nt += kBN)
        costed.push_back({static_cast<double>(ksum) + 1.0,
                          GemmTile{static_cast<int32_t>(g), static_cast<int32_t>(mt), static_cast<int32_t>(nt), 0}});
  }
  // Longest tiles first, so the last wave of the launch is made of short ones.
  std::stable_sort(costed.begin(), costed.end(), [](const Costed &a, const Costed &b) { return a.cost > b.cost; });
  std::vector<GemmTile> tiles;
  tiles.reserve(costed.size());
  for (const Costed &c : costed) tiles.push_back(c.tile);
  auto *p = new tt_ggemm_plan();
  p->num_groups = num_groups;
  p->num_terms = num_terms;
  p->num_tiles = static_cast<int64_t>(tiles.size());
  p->flops = flops;
  p->host_groups.assign(groups, groups + num_groups);
  auto upload = [&](void **where, const void *data, size_t bytes) -> int {
    if (!bytes) return 0;
    TT_HIP(hipMalloc(where, bytes));
    TT_HIP(hipMemcpy(*where, data, bytes, hipMemcpyHostToDevice));
    return 0;
  };
  if (upload(reinterpret_cast<void **>(&p->groups), groups, sizeof(tt_gemm_group) * num_groups) ||
      upload(reinterpret_cast<void **>(&p->terms), terms, sizeof(tt_gemm_term) * num_terms) ||
      upload(reinterpret_cast<void **>(&p->tiles), tiles.data(), sizeof(GemmTile) * tiles.size())) {
    std::string keep = g_error;
    tt_ggemm_plan_destroy(p);
    return fail(keep);
  }
  *plan = p;
  return 0;
}

extern "C" int tt_ggemm_plan_destroy(tt_ggemm_plan *plan) {
  if (!plan) return 0;
  if (plan->groups) (void)hipFree(plan->groups);
  if (plan->terms) (void)hipFree(plan->terms);
  if (plan->tiles) (void)hipFree(plan->tiles);
  delete plan;
  return 0;
}

extern "C" int tt_ggemm_plan_counts(const tt_ggemm_plan *plan, double *flops, int64_t *tiles) {
  if (!plan) return fail("tt_ggemm_plan_counts: null plan");
  if (flops) *flops = plan->flops;
  if (tiles) *tiles = plan->num_tiles;
  return 0;
}

namespace {
template <typename TA, typename TB, typename TO, typename TC>
int ggemm_launch(const tt_ggemm_plan *plan, const void *a, const void *b, void *c, hipStream_t stream, int64_t grid,
                 int32_t plain_group, int32_t plain_term, int32_t accumulate) {
  if (grid <= 0) return 0;
  hipLaunchKernelGGL((ggemm_kernel<TA, TB, TO, TC>), dim3(static_cast<uint32_t>(grid)), dim3(kGemmThreads), 0, stream,
                     plan->tiles, plan->groups, plan->terms, static_cast<const TA *>(a), static_cast<const TB *>(b),
                     static_cast<TO *>(c), plain_group, plain_term, accumulate);
  TT_LAUNCHED();
  return 0;
}

int ggemm_dispatch(const tt_ggemm_plan *plan, int at, int bt, int ct, int comp, const void *a, const void *b, void *c,
                   hipStream_t stream, int64_t grid, int32_t pg, int32_t pt, int32_t acc) {
#define TT_CASE(A_, B_, C_, K_, TA, TB, TO, TC) \
  if (at == A_ && bt == B_ && ct == C_ && comp == K_) return ggemm_launch<TA, TB, TO, TC>(plan, a, b, c, stream, grid, pg, pt, acc);
  TT_CASE(TT_F64, TT_F64, TT_F64, TT_F64, double, double, double, double)
  TT_CASE(TT_F32, TT_F32, TT_F32, TT_F32, float, float, float, float)
  TT_CASE(TT_F32, TT_F64, TT_F32, TT_F32, float, double, float, float)
  TT_CASE(TT_F32, TT_F32, TT_F64, TT_F32, float, float, double, float)
  TT_CASE(TT_F64, TT_F64, TT_F64, TT_F32, double, double, double, float)
  TT_CASE(TT_F64, TT_F64, TT_F32, TT_F32, double, double, float, float)
  TT_CASE(TT_F32, TT_F64, TT_F64, TT_F32, float, double, double, float)
#undef TT_CASE
  return fail("grouped GEMM: unsupported type combination (a " + std::to_string(at) + ", b " + std::to_string(bt) +
              ", c " + std::to_string(ct) + ", compute " + std::to_string(comp) + ")");
}
}  // namespace

extern "C" int tt_ggemm_execute(const tt_ggemm_plan *plan, int a_type, int b_type, int c_type, int compute_type,
                                const void *a, const void *b, void *c, void *stream) {
  if (!plan) return fail("tt_ggemm_execute: null plan");
  if (plan->num_tiles == 0) return 0;
  if (!c || ((!a || !b) && plan->num_terms)) return fail("tt_ggemm_execute: null buffer");
  return ggemm_dispatch(plan, a_type, b_type, c_type, compute_type, a, b, c, as_stream(stream), plan->num_tiles, -1, 0, 0);
}

extern "C" int tt_ggemm_execute_plain(const tt_ggemm_plan *plan, int dtype, const void *a, const void *b, void *c,
                                      void *stream) {
  if (!plan) return fail("tt_ggemm_execute_plain: null plan");
  if (dtype != TT_F64 && dtype != TT_F32) return fail("tt_ggemm_execute_plain: unsupported dtype");
  for (int64_t g = 0; g < plan->num_groups; ++g) {
    const tt_gemm_group &grp = plan->host_groups[static_cast<size_t>(g)];
    const int64_t grid = ((grp.m + kBM - 1) / kBM) * ((grp.n + kBN - 1) / kBN);
    if (grp.num_terms == 0) {
      // An empty term list still owes the block its zeros: one launch over an empty K.
      continue;
    }
    for (int32_t t = 0; t < grp.num_terms; ++t)
      if (int rc = ggemm_dispatch(plan, dtype, dtype, dtype, dtype, a, b, c, as_stream(stream), grid,
                                  static_cast<int32_t>(g), grp.first_term + t, t > 0 ? 1 : 0))
        return rc;
  }
  // Blocks without terms: zero them with the batched kernel restricted by an empty term range.
  for (int64_t g = 0; g < plan->num_groups; ++g) {
    const tt_gemm_group &grp = plan->host_groups[static_cast<size_t>(g)];
    if (grp.num_terms != 0) continue;
    const int64_t grid = ((grp.m + kBM - 1) / kBM) * ((grp.n + kBN - 1) / kBN);
    if (int rc = ggemm_dispatch(plan, dtype, dtype, dtype, dtype, a, b, c, as_stream(stream), grid,
                                static_cast<int32_t>(g), -1, 0))
      return rc;
  }
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Reductions and elementwise

namespace {

constexpr int kDotBlocks = 512;
constexpr int kDotThreads = 256;

struct DeviceScratch {
  double *partial = nullptr;
};

std::mutex g_scratch_mutex;
std::vector<DeviceScratch> g_scratch;

int scratch_for_current_device(double **out) {
  int device = 0;
  TT_HIP(hipGetDevice(&device));
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (static_cast<int>(g_scratch.size()) <= device) g_scratch.resize(static_cast<size_t>(device) + 1);
  DeviceScratch &s = g_scratch[static_cast<size_t>(device)];
  if (!s.partial) TT_HIP(hipMalloc(reinterpret_cast<void **>(&s.partial), sizeof(double) * 2 * kDotBlocks));
  *out = s.partial;
  return 0;
}

template <typename T>
__device__ double as_double(T x) { return static_cast<double>(x); }

// Stage one: block b sums x[i] y[i] over i = b*T + t (mod stride) in a fixed order, then a fixed
// tree in LDS.  Stage two: one block sums the partials the same way.
template <typename T>
__global__ __launch_bounds__(kDotThreads) void dot_stage1(int64_t n, const T *__restrict__ x, const T *__restrict__ y,
                                                          double *__restrict__ partial) {
  __shared__ double sh[kDotThreads];
  double acc = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kDotThreads + threadIdx.x; i < n;
       i += static_cast<int64_t>(kDotBlocks) * kDotThreads)
    acc += as_double(x[i]) * as_double(y[i]);
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kDotThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(kDotThreads) void dot_stage1_complex(int64_t n, const cdouble *__restrict__ x,
                                                                  const cdouble *__restrict__ y,
                                                                  double *__restrict__ partial) {
  __shared__ double sr[kDotThreads], si[kDotThreads];
  double re = 0.0, im = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kDotThreads + threadIdx.x; i < n;
       i += static_cast<int64_t>(kDotBlocks) * kDotThreads) {
    const cdouble a = x[i], b = y[i];
    re += a.re * b.re + a.im * b.im;
    im += a.re * b.im - a.im * b.re;
  }
  sr[threadIdx.x] = re;
  si[threadIdx.x] = im;
  __syncthreads();
  for (int s = kDotThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sr[threadIdx.x] += sr[threadIdx.x + s];
      si[threadIdx.x] += si[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sr[0];
    partial[2 * blockIdx.x + 1] = si[0];
  }
}

__global__ __launch_bounds__(kDotThreads) void dot_stage2(const double *__restrict__ partial, int width,
                                                          double *__restrict__ result) {
  __shared__ double sh[kDotThreads];
  for (int c = 0; c < width; ++c) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < kDotBlocks; i += kDotThreads) acc += partial[width * i + c];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kDotThreads / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) result[c] = sh[0];
    __syncthreads();
  }
}

template <typename T>
__global__ void axpby_kernel(int64_t n, double alpha, const double *alpha_dev, const T *__restrict__ x, double beta,
                             const double *beta_dev, T *__restrict__ y) {
  const double a = alpha_dev ? *alpha_dev : alpha;
  const double b = beta_dev ? *beta_dev : beta;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double yi = b == 0.0 ? 0.0 : b * static_cast<double>(y[i]);
    y[i] = static_cast<T>((x ? a * static_cast<double>(x[i]) : 0.0) + yi);
  }
}

template <typename S, typename D>
__global__ void convert_kernel(int64_t n, const S *__restrict__ x, D *__restrict__ y) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    y[i] = static_cast<D>(x[i]);
}

int64_t elementwise_grid(int64_t n) { return std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096)); }

}  // namespace

extern "C" int tt_dot(int dtype, int64_t n, const void *x, const void *y, double *result, void *stream) {
  if (!result) return fail("tt_dot: null result");
  if (n < 0) return fail("tt_dot: negative length");
  double *partial = nullptr;
  if (int rc = scratch_for_current_device(&partial)) return rc;
  hipStream_t s = as_stream(stream);
  int width = 1;
  switch (dtype) {
    case TT_F64:
      hipLaunchKernelGGL(dot_stage1<double>, dim3(kDotBlocks), dim3(kDotThreads), 0, s, n,
                         static_cast<const double *>(x), static_cast<const double *>(y), partial);
      break;
    case TT_F32:
      hipLaunchKernelGGL(dot_stage1<float>, dim3(kDotBlocks), dim3(kDotThreads), 0, s, n,
                         static_cast<const float *>(x), static_cast<const float *>(y), partial);
      break;
    case TT_C128:
      width = 2;
      hipLaunchKernelGGL(dot_stage1_complex, dim3(kDotBlocks), dim3(kDotThreads), 0, s, n,
                         static_cast<const cdouble *>(x), static_cast<const cdouble *>(y), partial);
      break;
    default: return fail("tt_dot: unsupported dtype");
  }
  TT_LAUNCHED();
  hipLaunchKernelGGL(dot_stage2, dim3(1), dim3(kDotThreads), 0, s, partial, width, result);
  TT_LAUNCHED();
  return 0;
}

extern "C" int tt_axpby(int dtype, int64_t n, double alpha, const double *alpha_dev, const void *x, double beta,
                        const double *beta_dev, void *y, void *stream) {
  if (n < 0) return fail("tt_axpby: negative length");
  if (n == 0) return 0;
  if (!y) return fail("tt_axpby: null output");
  const int64_t grid = elementwise_grid(n);
  switch (dtype) {
    case TT_F64:
      hipLaunchKernelGGL(axpby_kernel<double>, dim3(grid), dim3(256), 0, as_stream(stream), n, alpha, alpha_dev,
                         static_cast<const double *>(x), beta, beta_dev, static_cast<double *>(y));
      break;
    case TT_F32:
      hipLaunchKernelGGL(axpby_kernel<float>, dim3(grid), dim3(256), 0, as_stream(stream), n, alpha, alpha_dev,
                         static_cast<const float *>(x), beta, beta_dev, static_cast<float *>(y));
      break;
    default: return fail("tt_axpby: unsupported dtype");
  }
  TT_LAUNCHED();
  return 0;
}

extern "C" int tt_convert(int from_type, int to_type, int64_t n, const void *x, void *y, void *stream) {
  if (n < 0) return fail("tt_convert: negative length");
  if (n == 0) return 0;
  const int64_t grid = elementwise_grid(n);
  if (from_type == TT_F64 && to_type == TT_F32)
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid), dim3(256), 0, as_stream(stream), n,
                       static_cast<const double *>(x), static_cast<float *>(y));
  else if (from_type == TT_F32 && to_type == TT_F64)
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(grid), dim3(256), 0, as_stream(stream), n,
                       static_cast<const float *>(x), static_cast<double *>(y));
  else
    return fail("tt_convert: unsupported conversion");
  TT_LAUNCHED();
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Graphs

struct tt_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" int tt_graph_begin(void *stream) {
  TT_HIP(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal));
  return 0;
}

extern "C" int tt_graph_end(void *stream, tt_graph **graph) {
  if (!graph) return fail("tt_graph_end: null output handle");
  auto *g = new tt_graph();
  hipError_t st = hipStreamEndCapture(as_stream(stream), &g->graph);
  if (st != hipSuccess) {
    delete g;
    return fail(std::string("hipStreamEndCapture failed: ") + hipGetErrorString(st));
  }
  st = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (st != hipSuccess) {
    (void)hipGraphDestroy(g->graph);
    delete g;
    return fail(std::string("hipGraphInstantiate failed: ") + hipGetErrorString(st));
  }
  *graph = g;
  return 0;
}

extern "C" int tt_graph_launch(const tt_graph *graph, void *stream) {
  if (!graph) return fail("tt_graph_launch: null graph");
  TT_HIP(hipGraphLaunch(graph->exec, as_stream(stream)));
  return 0;
}

extern "C" int tt_graph_destroy(tt_graph *graph) {
  if (!graph) return 0;
  if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
  if (graph->graph) (void)hipGraphDestroy(graph->graph);
  delete graph;
  return 0;
}

extern "C" int tt_last_error(char *buffer, size_t size) {
  if (!buffer || size == 0) return -1;
  std::snprintf(buffer, size, "%s", g_error.c_str());
  return 0;
}

extern "C" int tt_version(void) { return 1; }
