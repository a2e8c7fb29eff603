// The MI355X device library behind include/tentorch_amd.h (part one): descriptor-table upload and
// the gather-scale-scatter leg-move kernel, fixed-order reductions, elementwise updates, hipGraph
// capture and the error channel.  The grouped block GEMM lives in ggemm.hip.  gfx950 only.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <string>
#include <vector>

#include "common.h"

namespace ttdev {

namespace {
thread_local std::string g_error;
}

int fail(const std::string &message) {
  g_error = message;
  return -1;
}

}  // namespace ttdev

using ttdev::as_stream;
using ttdev::cdouble;
using ttdev::fail;

namespace {

// ---------------------------------------------------------------------------------------------
// Gather-scale-scatter

// One work unit of a leg move: a run of destination elements of one tile.
struct GssChunk {
  int32_t task;
  int32_t count;
  int64_t start;
};

// The device form of a tile: extents with their division constants (n / d == (n * magic) >> shift
// for every n < 2^31, Granlund-Montgomery with N = 31), and the first source resolved in place so
// a single-source tile (every abelian move) reads no second record.
struct DevTask {
  int64_t dst_offset;
  int64_t src_offset;
  int32_t ndim;
  int32_t num_sources;
  int32_t first_source;
  int32_t pad;
  uint32_t extent[TT_MAX_DIMS];
  uint32_t shift[TT_MAX_DIMS];
  uint64_t magic[TT_MAX_DIMS];
  int64_t dst_stride[TT_MAX_DIMS];
  int64_t src_stride[TT_MAX_DIMS];
  double coef_re;
  double coef_im;
};

constexpr int kGssThreads = 256;
constexpr int kGssUnroll = 4;
// One pass per workgroup: a chunk is one element per (thread, unroll slot), so a large move runs as
// many short workgroups (≈ 2700 for the DMRG matvec's 2.8 M-element move at χ = 1024) and its
// memory latency is hidden by occupancy, where four passes per workgroup left under three
// workgroups per CU.
constexpr int kGssChunk = kGssThreads * kGssUnroll;

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

// Each workgroup takes one chunk of one tile.  A thread handles kGssUnroll elements per pass, 256
// apart, so consecutive lanes write consecutive destination elements along the tile's fastest axis
// and each thread keeps several independent loads in flight.
template <typename TS, typename T>
__global__ __launch_bounds__(kGssThreads) void gss_kernel(const GssChunk *__restrict__ chunks,
                                                          const DevTask *__restrict__ tasks,
                                                          const tt_gss_source *__restrict__ sources,
                                                          const TS *__restrict__ src, T *__restrict__ dst,
                                                          int32_t chunk_base) {
  const GssChunk chunk = chunks[chunk_base + blockIdx.x];
  const DevTask &task = tasks[chunk.task];
  const int nd = task.ndim;
  const int ns = task.num_sources;
  for (int32_t at = threadIdx.x; at < chunk.count; at += kGssThreads * kGssUnroll) {
    int64_t doff[kGssUnroll], soff[kGssUnroll];
    uint32_t idx[kGssUnroll][TT_MAX_DIMS];
#pragma unroll
    for (int u = 0; u < kGssUnroll; ++u) {
      uint32_t rem = static_cast<uint32_t>(chunk.start + at + u * kGssThreads);
      doff[u] = task.dst_offset;
      soff[u] = task.src_offset;
#pragma unroll
      for (int d = TT_MAX_DIMS - 1; d >= 0; --d) {
        if (d < nd) {
          const uint32_t q = static_cast<uint32_t>((static_cast<uint64_t>(rem) * task.magic[d]) >> task.shift[d]);
          idx[u][d] = rem - q * task.extent[d];
          rem = q;
          doff[u] += static_cast<int64_t>(idx[u][d]) * task.dst_stride[d];
          soff[u] += static_cast<int64_t>(idx[u][d]) * task.src_stride[d];
        } else {
          idx[u][d] = 0;
        }
      }
    }
    // Every slot loads, a slot past the chunk's end from the tile's first source element (an
    // address inside the source), so the loads carry no branch and are all in flight together;
    // only the stores are masked.
    bool live[kGssUnroll];
#pragma unroll
    for (int u = 0; u < kGssUnroll; ++u) live[u] = at + u * kGssThreads < chunk.count;
    T value[kGssUnroll];
    if (ns > 0) {
      TS loaded[kGssUnroll];
#pragma unroll
      for (int u = 0; u < kGssUnroll; ++u) loaded[u] = src[live[u] ? soff[u] : task.src_offset];
#pragma unroll
      for (int u = 0; u < kGssUnroll; ++u)
        value[u] = Scale<T>::fma(Scale<T>::zero(), task.coef_re, task.coef_im, static_cast<T>(loaded[u]));
    } else {
#pragma unroll
      for (int u = 0; u < kGssUnroll; ++u) value[u] = Scale<T>::zero();
    }
    for (int s = 1; s < ns; ++s) {
      const tt_gss_source &source = sources[task.first_source + s];
      TS loaded[kGssUnroll];
#pragma unroll
      for (int u = 0; u < kGssUnroll; ++u) {
        int64_t off = source.src_offset;
#pragma unroll
        for (int d = 0; d < TT_MAX_DIMS; ++d)
          if (d < nd) off += static_cast<int64_t>(idx[u][d]) * source.src_stride[d];
        loaded[u] = src[live[u] ? off : source.src_offset];
      }
#pragma unroll
      for (int u = 0; u < kGssUnroll; ++u)
        value[u] = Scale<T>::fma(value[u], source.coef_re, source.coef_im, static_cast<T>(loaded[u]));
    }
#pragma unroll
    for (int u = 0; u < kGssUnroll; ++u)
      if (live[u]) dst[doff[u]] = value[u];
  }
}

// The division constants of one extent d (1 <= d < 2^31): shift = 31 + ceil(log2 d),
// magic = ceil(2^shift / d); then (n * magic) >> shift == n / d for all n < 2^31.
void division_constants(uint32_t d, uint64_t *magic, uint32_t *shift) {
  uint32_t l = 0;
  while ((uint64_t(1) << l) < d) ++l;
  *shift = 31 + l;
  const unsigned __int128 numerator = static_cast<unsigned __int128>(1) << *shift;
  *magic = static_cast<uint64_t>((numerator + d - 1) / d);
}

}  // namespace

struct tt_gss_plan {
  void *block = nullptr;   // the plan's range of the table arena, holding its three tables
  DevTask *tasks = nullptr;
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
  std::vector<DevTask> dev(static_cast<size_t>(num_tasks));
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
    DevTask &d = dev[static_cast<size_t>(t)];
    std::memset(&d, 0, sizeof(DevTask));
    d.dst_offset = task.dst_offset;
    d.ndim = task.ndim;
    d.num_sources = task.num_sources;
    d.first_source = task.first_source;
    for (int k = 0; k < task.ndim; ++k) {
      d.extent[k] = static_cast<uint32_t>(task.extent[k]);
      division_constants(d.extent[k], &d.magic[k], &d.shift[k]);
      d.dst_stride[k] = task.dst_stride[k];
    }
    if (task.num_sources > 0) {
      const tt_gss_source &s0 = sources[task.first_source];
      d.src_offset = s0.src_offset;
      for (int k = 0; k < task.ndim; ++k) d.src_stride[k] = s0.src_stride[k];
      d.coef_re = s0.coef_re;
      d.coef_im = s0.coef_im;
    }
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
  if (ttdev::upload_tables(&p->block, {{reinterpret_cast<void **>(&p->tasks), dev.data(), sizeof(DevTask) * dev.size()},
                                        {reinterpret_cast<void **>(&p->sources), sources, sizeof(tt_gss_source) * num_sources},
                                        {reinterpret_cast<void **>(&p->chunks), chunks.data(), sizeof(GssChunk) * chunks.size()}})) {
    std::string keep = ttdev::g_error;
    tt_gss_plan_destroy(p);
    return fail(keep);
  }
  *plan = p;
  return 0;
}

extern "C" int tt_gss_plan_destroy(tt_gss_plan *plan) {
  if (!plan) return 0;
  ttdev::release_tables(plan->block);
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
template <typename TS, typename T>
int gss_launch(const tt_gss_plan *plan, const void *src, void *dst, hipStream_t stream, int32_t first, int32_t count) {
  if (count <= 0) return 0;
  if (int rc = ttdev::await_uploads(stream)) return rc;
  if (int rc = ttdev::launch(gss_kernel<TS, T>, dim3(count), dim3(kGssThreads), 0, stream, plan->chunks, plan->tasks,
                             plan->sources, static_cast<const TS *>(src), static_cast<T *>(dst), first))
    return rc;
  TT_LAUNCHED();
  ttdev::note_launch();
  return 0;
}

int gss_dispatch(const tt_gss_plan *plan, int src_type, int dst_type, const void *src, void *dst, hipStream_t stream,
                 int32_t first, int32_t count) {
  if (src_type == TT_F64 && dst_type == TT_F64) return gss_launch<double, double>(plan, src, dst, stream, first, count);
  if (src_type == TT_F32 && dst_type == TT_F32) return gss_launch<float, float>(plan, src, dst, stream, first, count);
  if (src_type == TT_C128 && dst_type == TT_C128) return gss_launch<cdouble, cdouble>(plan, src, dst, stream, first, count);
  if (src_type == TT_F64 && dst_type == TT_F32) return gss_launch<double, float>(plan, src, dst, stream, first, count);
  if (src_type == TT_F32 && dst_type == TT_F64) return gss_launch<float, double>(plan, src, dst, stream, first, count);
  return fail("gather-scale-scatter: unsupported types " + std::to_string(src_type) + " -> " + std::to_string(dst_type));
}

int gss_dispatch(const tt_gss_plan *plan, int dtype, const void *src, void *dst, hipStream_t stream, int32_t first,
                 int32_t count) {
  return gss_dispatch(plan, dtype, dtype, src, dst, stream, first, count);
}
}  // namespace

extern "C" int tt_gss_execute_typed(const tt_gss_plan *plan, int src_type, int dst_type, const void *src, void *dst,
                                    void *stream) {
  if (!plan) return fail("tt_gss_execute_typed: null plan");
  if (plan->num_chunks == 0) return 0;
  if (!src || !dst) return fail("tt_gss_execute_typed: null buffer");
  return gss_dispatch(plan, src_type, dst_type, src, dst, as_stream(stream), 0, static_cast<int32_t>(plan->num_chunks));
}

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
// Krylov-basis kernels: the k scalar products of one vector with k basis rows, and the update of a
// vector by a combination of basis rows, both with coefficients that stay on the device.

namespace {

constexpr int kMultiBlocks = 1024;   // enough waves to keep the memory-bound passes fed
constexpr int kMultiMax = 32;

std::vector<double *> g_multi_scratch;

int multi_scratch(double **out) {
  int device = 0;
  TT_HIP(hipGetDevice(&device));
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (static_cast<int>(g_multi_scratch.size()) <= device) g_multi_scratch.resize(static_cast<size_t>(device) + 1, nullptr);
  double *&s = g_multi_scratch[static_cast<size_t>(device)];
  if (!s) TT_HIP(hipMalloc(reinterpret_cast<void **>(&s), sizeof(double) * kMultiMax * kMultiBlocks));
  *out = s;
  return 0;
}

// Block b sums V[i, j] w[j] over its strided share of j for every row i < k (fixed order), then a
// fixed LDS tree per row; partial[i * kMultiBlocks + b].
// kRows (a compile-time bound on k: 4, 8, 16 or 32) keeps the per-row registers to what k needs.
template <int kRows>
__global__ __launch_bounds__(kDotThreads) void multidot_stage1(int64_t n, int k, const double *__restrict__ V, int64_t ldv,
                                                               const double *__restrict__ w, double *__restrict__ partial) {
  __shared__ double sh[kDotThreads];
  double acc[kRows];
#pragma unroll
  for (int i = 0; i < kRows; ++i) acc[i] = 0.0;
  for (int64_t j = static_cast<int64_t>(blockIdx.x) * kDotThreads + threadIdx.x; j < n;
       j += static_cast<int64_t>(kMultiBlocks) * kDotThreads) {
    const double x = w[j];
#pragma unroll
    for (int i = 0; i < kRows; ++i)
      if (i < k) acc[i] += V[i * ldv + j] * x;
  }
  // unrolled over kRows (the guard is uniform), so acc stays in registers
#pragma unroll
  for (int i = 0; i < kRows; ++i) {
    if (i < k) {
      sh[threadIdx.x] = acc[i];
      __syncthreads();
#pragma unroll
      for (int s = kDotThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
      }
      if (threadIdx.x == 0) partial[i * kMultiBlocks + blockIdx.x] = sh[0];
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(kDotThreads) void multidot_stage2(const double *__restrict__ partial, double *__restrict__ out) {
  __shared__ double sh[kDotThreads];
  const int i = blockIdx.x;
  double acc = 0.0;
  for (int b = threadIdx.x; b < kMultiBlocks; b += kDotThreads) acc += partial[i * kMultiBlocks + b];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kDotThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i] = sh[0];
}

// `count`, when given, caps k from device memory: rows of V past it are never read (a Lanczos basis
// past an early stop may hold non-finite rows).
__global__ void multiaxpy_kernel(int64_t n, int k, const int32_t *__restrict__ count, const double *__restrict__ coef,
                                 double scale, const double *__restrict__ V, int64_t ldv, double *__restrict__ w) {
  const int rows = count ? min(k, *count) : k;
  for (int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < n;
       j += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double acc = 0.0;
    for (int i = 0; i < rows; ++i) acc += coef[i] * V[i * ldv + j];
    w[j] += scale * acc;
  }
}

// The Ritz pair of a Lanczos run from its device scalars (lanczos.py:52-62): the tridiagonal of
// alpha[0..size) and beta[0..size-1), where size stops at the first beta below `tol` (the
// reference's break), is diagonalised by implicit QL with Wilkinson-type shifts and eigenvector
// accumulation (the tql2 iteration of EISPACK).  One wave: every lane runs the scalar recurrence on
// the wave's copy of the tridiagonal in LDS, and lane r applies the rotations to row r of the
// eigenvector matrix.  Writes the least eigenvalue, its eigenvector as the coefficients of the
// basis (zero past size; the sign makes the first coefficient non-negative, the Ritz vector's
// overlap with the start), and size.
constexpr int kRitzMax = 32;
__global__ __launch_bounds__(64) void lanczos_ritz_kernel(int k, const double *__restrict__ alpha,
                                                          const double *__restrict__ beta, double tol,
                                                          double *__restrict__ coef, double *__restrict__ ritz,
                                                          int32_t *__restrict__ size_out) {
  __shared__ double d[kRitzMax], e[kRitzMax], z[kRitzMax][kRitzMax + 1];
  const int lane = threadIdx.x;
  int m = k;
  for (int j = 0; j < k; ++j)
    if (beta[j] < tol) {
      m = j + 1;
      break;
    }
  if (lane < m) {
    d[lane] = alpha[lane];
    e[lane] = lane + 1 < m ? beta[lane] : 0.0;
    for (int c = 0; c < m; ++c) z[lane][c] = lane == c ? 1.0 : 0.0;
  }
  __builtin_amdgcn_wave_barrier();
  constexpr double kEps = 2.220446049250313e-16;
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
          const double zi1 = z[lane][i + 1], zi = z[lane][i];
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
  __builtin_amdgcn_wave_barrier();
  int least = 0;
  for (int c = 1; c < m; ++c)
    if (d[c] < d[least]) least = c;
  const double sign = z[0][least] < 0.0 ? -1.0 : 1.0;
  if (lane < k) coef[lane] = lane < m ? sign * z[lane][least] : 0.0;
  if (lane == 0) {
    ritz[0] = d[least];
    size_out[0] = m;
  }
}

// One Gram-Schmidt pass fused with what follows it: w -= V[0..k) coef, then, over the updated w,
// either the partial products with V[0..k) (next_rows = k, the next pass's projection) or the
// partial sum of squares (next_rows = 0, the norm); partials laid out as multidot_stage1's.
template <int kRows>
__global__ __launch_bounds__(kDotThreads) void project_kernel(int64_t n, int k, const double *__restrict__ V, int64_t ldv,
                                                              const double *__restrict__ coef, double *__restrict__ w,
                                                              int next_rows, double *__restrict__ partial) {
  __shared__ double sh[kDotThreads];
  double c[kRows], acc[kRows];
#pragma unroll
  for (int i = 0; i < kRows; ++i) {
    c[i] = i < k ? coef[i] : 0.0;
    acc[i] = 0.0;
  }
  double squares = 0.0;
  for (int64_t j = static_cast<int64_t>(blockIdx.x) * kDotThreads + threadIdx.x; j < n;
       j += static_cast<int64_t>(kMultiBlocks) * kDotThreads) {
    double v[kRows];
    double x = w[j];
#pragma unroll
    for (int i = 0; i < kRows; ++i) {
      v[i] = i < k ? V[i * ldv + j] : 0.0;
      x -= c[i] * v[i];
    }
    w[j] = x;
    squares += x * x;
#pragma unroll
    for (int i = 0; i < kRows; ++i) acc[i] += v[i] * x;
  }
  const int rows = next_rows > 0 ? next_rows : 1;
  // unrolled over kRows (the guard is uniform), so acc stays in registers
#pragma unroll
  for (int i = 0; i < kRows; ++i) {
    if (i < rows) {
      sh[threadIdx.x] = next_rows > 0 ? acc[i] : squares;
      __syncthreads();
#pragma unroll
      for (int s = kDotThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
      }
      if (threadIdx.x == 0) partial[i * kMultiBlocks + blockIdx.x] = sh[0];
      __syncthreads();
    }
  }
}

// The step's scalars from the norm's partials: beta = |w|, 1/beta, and alpha = the two passes'
// projections on row k (first[k] + second[k]).
__global__ __launch_bounds__(kDotThreads) void lanczos_scalars_kernel(const double *__restrict__ partial, int k,
                                                                      const double *__restrict__ first,
                                                                      const double *__restrict__ second,
                                                                      double *__restrict__ alpha,
                                                                      double *__restrict__ beta,
                                                                      double *__restrict__ inverse) {
  __shared__ double sh[kDotThreads];
  double acc = 0.0;
  for (int b = threadIdx.x; b < kMultiBlocks; b += kDotThreads) acc += partial[b];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kDotThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double norm = sqrt(sh[0]);
    *beta = norm;
    *inverse = 1.0 / norm;
    *alpha = first[k] + second[k];
  }
}

// The smallest row bound that covers k rows.
template <typename Launch>
int with_rows(int k, Launch launch) {
  if (k <= 4) return launch(std::integral_constant<int, 4>{});
  if (k <= 8) return launch(std::integral_constant<int, 8>{});
  if (k <= 16) return launch(std::integral_constant<int, 16>{});
  return launch(std::integral_constant<int, 32>{});
}

}  // namespace

extern "C" int tt_lanczos_orthonormalize(int64_t n, int k, double *V, int64_t ldv, double *w, double *coef,
                                         double *alpha, double *beta, void *stream) {
  if (n < 0 || k < 0 || k + 1 > kMultiMax) return fail("tt_lanczos_orthonormalize: k + 1 must lie in [1, 32]");
  if (!V || !w || !coef || !alpha || !beta) return fail("tt_lanczos_orthonormalize: null buffer");
  if (n == 0) return 0;
  double *partial = nullptr;
  if (int rc = multi_scratch(&partial)) return rc;
  hipStream_t s = as_stream(stream);
  const int rows = k + 1;
  double *first = coef, *second = coef + rows, *inverse = coef + 2 * rows;
  return with_rows(rows, [&](auto bound) -> int {
    constexpr int R = decltype(bound)::value;
    // first pass: the projections of w on rows 0..k
    hipLaunchKernelGGL(multidot_stage1<R>, dim3(kMultiBlocks), dim3(kDotThreads), 0, s, n, rows, V, ldv, w, partial);
    TT_LAUNCHED();
    hipLaunchKernelGGL(multidot_stage2, dim3(rows), dim3(kDotThreads), 0, s, partial, first);
    TT_LAUNCHED();
    // subtract them, and project the result again
    hipLaunchKernelGGL(project_kernel<R>, dim3(kMultiBlocks), dim3(kDotThreads), 0, s, n, rows, V, ldv, first, w, rows,
                       partial);
    TT_LAUNCHED();
    hipLaunchKernelGGL(multidot_stage2, dim3(rows), dim3(kDotThreads), 0, s, partial, second);
    TT_LAUNCHED();
    // subtract the second projections, with the norm of what remains
    hipLaunchKernelGGL(project_kernel<R>, dim3(kMultiBlocks), dim3(kDotThreads), 0, s, n, rows, V, ldv, second, w, 0,
                       partial);
    TT_LAUNCHED();
    hipLaunchKernelGGL(lanczos_scalars_kernel, dim3(1), dim3(kDotThreads), 0, s, partial, k, first, second, alpha, beta,
                       inverse);
    TT_LAUNCHED();
    // the next basis vector
    hipLaunchKernelGGL(axpby_kernel<double>, dim3(elementwise_grid(n)), dim3(256), 0, s, n, 0.0, inverse, w, 0.0,
                       nullptr, V + rows * ldv);
    TT_LAUNCHED();
    return 0;
  });
}

extern "C" int tt_multidot(int64_t n, int k, const double *V, int64_t ldv, const double *w, double *out, void *stream) {
  if (n < 0 || k < 0 || k > kMultiMax) return fail("tt_multidot: k must lie in [0, 32] and n be non-negative");
  if (k == 0) return 0;
  if (!V || !w || !out) return fail("tt_multidot: null buffer");
  double *partial = nullptr;
  if (int rc = multi_scratch(&partial)) return rc;
  if (int rc = with_rows(k, [&](auto bound) -> int {
        hipLaunchKernelGGL(multidot_stage1<decltype(bound)::value>, dim3(kMultiBlocks), dim3(kDotThreads), 0,
                           as_stream(stream), n, k, V, ldv, w, partial);
        TT_LAUNCHED();
        return 0;
      }))
    return rc;
  hipLaunchKernelGGL(multidot_stage2, dim3(k), dim3(kDotThreads), 0, as_stream(stream), partial, out);
  TT_LAUNCHED();
  return 0;
}

extern "C" int tt_multiaxpy(int64_t n, int k, const double *coef, double scale, const double *V, int64_t ldv, double *w,
                            void *stream) {
  return tt_multiaxpy_counted(n, k, nullptr, coef, scale, V, ldv, w, stream);
}

extern "C" int tt_multiaxpy_counted(int64_t n, int k, const int32_t *count, const double *coef, double scale,
                                    const double *V, int64_t ldv, double *w, void *stream) {
  if (n < 0 || k < 0) return fail("tt_multiaxpy: negative size");
  if (n == 0 || k == 0) return 0;
  if (!V || !w || !coef) return fail("tt_multiaxpy: null buffer");
  hipLaunchKernelGGL(multiaxpy_kernel, dim3(elementwise_grid(n)), dim3(256), 0, as_stream(stream), n, k, count, coef,
                     scale, V, ldv, w);
  TT_LAUNCHED();
  return 0;
}

extern "C" int tt_lanczos_ritz(int k, const double *alpha, const double *beta, double tol, double *coef, double *ritz,
                               int32_t *size, void *stream) {
  if (k < 1 || k > kRitzMax) return fail("tt_lanczos_ritz: k must lie in [1, 32]");
  if (!alpha || !beta || !coef || !ritz || !size) return fail("tt_lanczos_ritz: null buffer");
  hipLaunchKernelGGL(lanczos_ritz_kernel, dim3(1), dim3(64), 0, as_stream(stream), k, alpha, beta, tol, coef, ritz,
                     size);
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
  std::snprintf(buffer, size, "%s", ttdev::g_error.c_str());
  return 0;
}

extern "C" int tt_version(void) { return 1; }

namespace {
std::atomic<int64_t> g_launches{0};
}

void ttdev::note_launch() { g_launches.fetch_add(1, std::memory_order_relaxed); }

namespace {

// One arena chunk of plan tables: bump-allocated, freed when no live plan uses it and it no longer
// takes new tables.
struct ArenaChunk {
  char *base = nullptr;
  size_t capacity = 0, used = 0;
  int64_t live = 0;
};

// Per device: the upload stream, the event recorded after its latest copy, the pinned staging ring,
// the arena chunks (the last takes new tables), and how many uploads were issued and were last
// seen complete.
struct Uploader {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  unsigned char *staging = nullptr;
  size_t staging_capacity = 0, staging_used = 0;
  std::vector<ArenaChunk> chunks;
  uint64_t issued = 0, completed = 0;
  // Inside tt_upload_batch_begin/end: staged tables whose copy is not issued yet, one contiguous
  // range of staging and of the arena (consecutive uploads advance both by the same bytes).
  int batch_depth = 0;
  unsigned char *pending_stage = nullptr;
  char *pending_target = nullptr;
  size_t pending_bytes = 0;
};
std::mutex g_upload_mutex;
std::vector<Uploader> g_uploaders;  // one per device

constexpr size_t kChunkBytes = size_t(16) << 20;
constexpr size_t kStagingBytes = size_t(8) << 20;
constexpr size_t kTableAlign = 256;

size_t aligned(size_t bytes) { return (bytes + kTableAlign - 1) / kTableAlign * kTableAlign; }

// Issues the copy of the pending range and records the upload event after it.
int flush_pending(Uploader &u) {
  if (!u.pending_bytes) return 0;
  TT_HIP(hipMemcpyAsync(u.pending_target, u.pending_stage, u.pending_bytes, hipMemcpyHostToDevice, u.stream));
  TT_HIP(hipEventRecord(u.done, u.stream));
  u.issued += 1;
  u.pending_stage = nullptr;
  u.pending_target = nullptr;
  u.pending_bytes = 0;
  return 0;
}

Uploader *current_uploader(int *rc) {
  int device = 0;
  *rc = 0;
  if (hipGetDevice(&device) != hipSuccess) {
    *rc = fail("upload: no current device");
    return nullptr;
  }
  if (g_uploaders.size() <= static_cast<size_t>(device)) g_uploaders.resize(static_cast<size_t>(device) + 1);
  return &g_uploaders[static_cast<size_t>(device)];
}

}  // namespace

extern "C" int tt_upload_batch_begin(void) {
  std::lock_guard<std::mutex> lock(g_upload_mutex);
  int rc = 0;
  Uploader *u = current_uploader(&rc);
  if (!u) return rc;
  u->batch_depth += 1;
  return 0;
}

extern "C" int tt_upload_batch_end(void) {
  std::lock_guard<std::mutex> lock(g_upload_mutex);
  int rc = 0;
  Uploader *u = current_uploader(&rc);
  if (!u) return rc;
  if (u->batch_depth <= 0) return fail("tt_upload_batch_end: no batch is open");
  if (--u->batch_depth == 0) return flush_pending(*u);
  return 0;
}

int ttdev::upload_tables(void **block, std::initializer_list<Table> tables) {
  *block = nullptr;
  size_t total = 0;
  for (const Table &t : tables) {
    *t.where = nullptr;
    total = aligned(total) + t.bytes;
  }
  if (!total) return 0;
  int device = 0;
  TT_HIP(hipGetDevice(&device));
  std::lock_guard<std::mutex> lock(g_upload_mutex);
  if (g_uploaders.size() <= static_cast<size_t>(device)) g_uploaders.resize(static_cast<size_t>(device) + 1);
  Uploader &u = g_uploaders[static_cast<size_t>(device)];
  if (!u.stream) TT_HIP(hipStreamCreateWithFlags(&u.stream, hipStreamNonBlocking));
  if (!u.done) TT_HIP(hipEventCreateWithFlags(&u.done, hipEventDisableTiming));
  // device range: the current chunk, or a new one
  const size_t need = aligned(total);
  if (u.chunks.empty() || u.chunks.back().capacity - u.chunks.back().used < need) {
    if (int rc = flush_pending(u)) return rc;
    ArenaChunk chunk;
    chunk.capacity = std::max(kChunkBytes, need);
    TT_HIP(hipMalloc(reinterpret_cast<void **>(&chunk.base), chunk.capacity));
    // the chunk that stops taking tables goes now if no live plan uses it
    if (!u.chunks.empty() && u.chunks.back().live == 0) {
      TT_HIP(hipFree(u.chunks.back().base));
      u.chunks.pop_back();
    }
    u.chunks.push_back(chunk);
  }
  ArenaChunk &chunk = u.chunks.back();
  char *const target = chunk.base + chunk.used;
  // staging: the rest of the ring, or (after the copies in flight have finished) its front
  if (u.staging_capacity < total) {
    if (int rc = flush_pending(u)) return rc;
    TT_HIP(hipStreamSynchronize(u.stream));
    if (u.staging) TT_HIP(hipHostFree(u.staging));
    u.staging = nullptr;
    u.staging_capacity = u.staging_used = 0;
    const size_t grown = std::max(kStagingBytes, need);
    TT_HIP(hipHostMalloc(reinterpret_cast<void **>(&u.staging), grown, hipHostMallocDefault));
    u.staging_capacity = grown;
  } else if (u.staging_capacity - u.staging_used < total) {
    if (int rc = flush_pending(u)) return rc;
    TT_HIP(hipStreamSynchronize(u.stream));
    u.staging_used = 0;
  }
  unsigned char *const stage = u.staging + u.staging_used;
  std::vector<size_t> offsets;
  size_t at = 0;
  for (const Table &t : tables) {
    at = aligned(at);
    offsets.push_back(at);
    if (t.bytes) std::memcpy(stage + at, t.data, t.bytes);
    at += t.bytes;
  }
  // the copy: now, or (inside a batch) with the pending range it extends
  if (u.pending_bytes && (u.pending_stage + u.pending_bytes != stage || u.pending_target + u.pending_bytes != target))
    if (int rc = flush_pending(u)) return rc;
  if (!u.pending_bytes) {
    u.pending_stage = stage;
    u.pending_target = target;
  }
  u.pending_bytes += need;
  if (u.batch_depth == 0)
    if (int rc = flush_pending(u)) return rc;
  u.staging_used = std::min(u.staging_capacity, u.staging_used + need);
  chunk.used += need;
  chunk.live += 1;
  *block = target;
  size_t k = 0;
  for (const Table &t : tables) {
    if (t.bytes) *t.where = target + offsets[k];
    ++k;
  }
  return 0;
}

void ttdev::release_tables(void *block) {
  if (!block) return;
  const char *const p = static_cast<const char *>(block);
  std::lock_guard<std::mutex> lock(g_upload_mutex);
  for (Uploader &u : g_uploaders)
    for (size_t c = 0; c < u.chunks.size(); ++c) {
      ArenaChunk &chunk = u.chunks[c];
      if (p < chunk.base || p >= chunk.base + chunk.capacity) continue;
      if (--chunk.live == 0 && c + 1 < u.chunks.size()) {
        (void)hipFree(chunk.base);   // waits for the device, so no kernel still reads these tables
        u.chunks.erase(u.chunks.begin() + static_cast<std::ptrdiff_t>(c));
      }
      return;
    }
}

int ttdev::await_uploads(hipStream_t stream) {
  int device = 0;
  TT_HIP(hipGetDevice(&device));
  std::lock_guard<std::mutex> lock(g_upload_mutex);
  if (g_uploaders.size() <= static_cast<size_t>(device)) return 0;
  Uploader &u = g_uploaders[static_cast<size_t>(device)];
  // a launch inside a batch needs the tables staged so far
  if (int rc = flush_pending(u)) return rc;
  if (u.completed == u.issued) return 0;
  hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
  TT_HIP(hipStreamIsCapturing(stream, &capture));
  if (capture != hipStreamCaptureStatusNone) {
    // a graph cannot wait on an event recorded outside it: finish the uploads on the host
    TT_HIP(hipStreamSynchronize(u.stream));
    u.completed = u.issued;
    return 0;
  }
  const hipError_t query = hipEventQuery(u.done);
  if (query == hipSuccess) {
    u.completed = u.issued;
    return 0;
  }
  if (query != hipErrorNotReady) return fail(std::string("upload event query failed: ") + hipGetErrorString(query));
  TT_HIP(hipStreamWaitEvent(stream, u.done, 0));
  return 0;
}

extern "C" int64_t tt_launch_count(void) { return g_launches.load(std::memory_order_relaxed); }

namespace {
thread_local bool g_timing = false;
thread_local std::vector<ttdev::TimedPair> g_timed;
}  // namespace

int ttdev::timed_pair(TimedPair *pair) {
  if (!g_timing) return 0;
  TT_HIP(hipEventCreate(&pair->start));
  if (hipEventCreate(&pair->stop) != hipSuccess) {
    (void)hipEventDestroy(pair->start);
    return fail("hipEventCreate failed");
  }
  g_timed.push_back(*pair);
  return 0;
}

extern "C" int tt_timing_enable(int on) {
  g_timing = on != 0;
  return 0;
}

extern "C" int tt_timing_collect(double *milliseconds, int64_t capacity, int64_t *count) {
  if (count) *count = static_cast<int64_t>(g_timed.size());
  if (!milliseconds) return 0;
  if (capacity < static_cast<int64_t>(g_timed.size())) return fail("tt_timing_collect: the output is too small");
  int status = 0;
  for (size_t i = 0; i < g_timed.size(); ++i) {
    ttdev::TimedPair &p = g_timed[i];
    float ms = 0.0f;
    if (status == 0) {
      hipError_t st = hipEventSynchronize(p.stop);
      if (st == hipSuccess) st = hipEventElapsedTime(&ms, p.start, p.stop);
      if (st != hipSuccess) status = fail(std::string("kernel timing: ") + hipGetErrorString(st));
    }
    milliseconds[i] = ms;
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  g_timed.clear();
  return status;
}
