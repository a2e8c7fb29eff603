// The MI355X device library behind include/tentorch_amd.h (part two): the grouped block GEMM on
// FP64 and FP32 MFMA (complex128 as four FP64 products per step).  One launch computes every output block of a contraction step; each
// workgroup owns one BM x BN output tile of one block and sums that block's terms in list order.
// The tile shape is a property of the plan (chosen by the host from the block shapes, or by an
// autotuner through tt_ggemm_plan_create_config).  gfx950 only.
#include <algorithm>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "common.h"

using ttdev::as_stream;
using ttdev::cdouble;
using ttdev::fail;

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Result layout of one 16x16x4 MFMA (measured on gfx950): lane l holds column l % 16 in its four
// result registers r; the f64 form holds rows 4 r + l / 16, the f32 form rows 4 (l / 16) + r.
// Operands: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16].
template <typename TC>
struct Mfma;
template <>
struct Mfma<double> {
  using acc_t = d4;
  static __device__ acc_t zero() { return acc_t{0, 0, 0, 0}; }
  static __device__ acc_t step(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  static __device__ double get(const acc_t &c, int r) { return c[r]; }
  static __device__ int row(int lk, int r) { return 4 * r + lk; }
};
template <>
struct Mfma<float> {
  using acc_t = f4;
  static __device__ acc_t zero() { return acc_t{0, 0, 0, 0}; }
  static __device__ acc_t step(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ float get(const acc_t &c, int r) { return c[r]; }
  static __device__ int row(int lk, int r) { return 4 * lk + r; }
};
// Complex: (a_re + i a_im)(b_re + i b_im) as four FP64 MFMAs into a real and an imaginary
// accumulator, which share the f64 result layout.
struct cd4 {
  d4 re, im;
};
template <>
struct Mfma<cdouble> {
  using acc_t = cd4;
  static __device__ acc_t zero() { return acc_t{d4{0, 0, 0, 0}, d4{0, 0, 0, 0}}; }
  static __device__ acc_t step(cdouble a, cdouble b, acc_t c) {
    c.re = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.re, c.re, 0, 0, 0);
    c.re = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.im, b.im, c.re, 0, 0, 0);
    c.im = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.im, c.im, 0, 0, 0);
    c.im = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b.re, c.im, 0, 0, 0);
    return c;
  }
  static __device__ cdouble get(const acc_t &c, int r) { return cdouble{c.re[r], c.im[r]}; }
  static __device__ int row(int lk, int r) { return 4 * r + lk; }
};

// Element conversions between storage and compute types (real types convert, complex stays).
template <typename To>
struct Conv {
  template <typename From>
  static __device__ To from(From x) { return static_cast<To>(x); }
  static __device__ To zero() { return To(0); }
};
template <>
struct Conv<cdouble> {
  static __device__ cdouble from(cdouble x) { return x; }
  static __device__ cdouble zero() { return cdouble{0.0, 0.0}; }
};
template <typename T>
__device__ T add(T a, T b) { return a + b; }
__device__ cdouble add(cdouble a, cdouble b) { return cdouble{a.re + b.re, a.im + b.im}; }

// Raw buffer loads.  A buffer resource holds a base address and a byte range; a load at an offset
// outside the range returns zeros and touches no memory, so an operand slab can be read whole,
// without a branch or a select per element, and what lies past a block reads as zero.  Word 3 of
// the resource is the gfx9-family default (no format conversion, raw addressing).
constexpr int kBufferWord3 = 0x00020000;
// The resource of one operand slab.  Its fields are uniform over the workgroup; reading them
// through readfirstlane puts them in scalar registers, where a buffer load needs its resource (the
// compiler cannot always prove that a value derived from the tile record is uniform, and would
// otherwise wrap every load in a loop over distinct lane values).
__device__ inline __amdgpu_buffer_rsrc_t resource(const char *base, uint32_t bytes) {
  const uint64_t address = reinterpret_cast<uint64_t>(base);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(address));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(address >> 32));
  const uint32_t range = __builtin_amdgcn_readfirstlane(bytes);
  void *uniform = reinterpret_cast<void *>((static_cast<uint64_t>(hi) << 32) | lo);
  return __builtin_amdgcn_make_buffer_rsrc(uniform, static_cast<short>(0), static_cast<int>(range), kBufferWord3);
}
template <typename T>
struct BufferLoad;
template <>
struct BufferLoad<float> {
  static __device__ float at(__amdgpu_buffer_rsrc_t r, uint32_t byte) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte, 0, 0));
  }
};
template <>
struct BufferLoad<double> {
  static __device__ double at(__amdgpu_buffer_rsrc_t r, uint32_t byte) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, byte, 0, 0));
  }
};
template <>
struct BufferLoad<cdouble> {
  static __device__ cdouble at(__amdgpu_buffer_rsrc_t r, uint32_t byte) {
    return __builtin_bit_cast(cdouble, __builtin_amdgcn_raw_buffer_load_b128(r, byte, 0, 0));
  }
};

// Output tile of one workgroup: which block, and the tile's first row and column.
struct GemmTile {
  int32_t group;
  int32_t m0;
  int32_t n0;
  int32_t pad;
};

// Device form of a term: A's row stride travels with the depth, so that a piece of a term (a
// k-range of it, for a product whose depth is split) reads the rows of the whole term.
struct KTerm {
  int64_t a_offset;
  int64_t b_offset;
  int64_t k;
  int64_t lda;
};

// Device form of a group.  A part of a split product (partial != 0) writes its m x n partial sum,
// in the compute type, at c_offset of the plan's partial-sum buffer instead of into C.
struct KGroup {
  int64_t c_offset;
  int64_t m;
  int64_t n;
  int32_t first_term;
  int32_t num_terms;
  int32_t partial;
  int32_t pad;
};

// A split product: its parts' partial sums lie back to back from p_offset (part s at
// p_offset + s * elems) and are added in part order into the m x n block at c_offset of C.
struct RGroup {
  int64_t c_offset;
  int64_t elems;
  int64_t p_offset;
  int32_t parts;
  int32_t pad;
};

// One workgroup of the partial-sum launch: kReduceSpan consecutive elements of one split product.
struct ReduceTile {
  int32_t rgroup;
  int32_t pad;
  int64_t e0;
};
constexpr int kReduceThreads = 256;
constexpr int64_t kReduceSpan = kReduceThreads;  // one element per thread
constexpr int kMaxLoadedParts = 8;               // the most parts a split product has

constexpr int kGemmThreads = 256;  // four waves in a 2 x 2 grid over the tile
// Limits that keep a tile's byte offsets and the ranges of its operand blocks in 32 bits.
constexpr int64_t kMaxSide = int64_t(1) << 20;
constexpr int64_t kMaxBlockElements = int64_t(1) << 27;
constexpr int kPad = 1;

// The tile shapes the library is built with: (BM, BN, BK).
struct Shape {
  int bm, bn, bk;
};
constexpr Shape kShapes[] = {{64, 64, 16}, {32, 32, 16}, {128, 64, 16}, {64, 32, 16}, {32, 64, 16}, {128, 128, 16}, {64, 64, 32}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

// A BM x BN tile per workgroup: four waves in a 2 x 2 grid, each wave (BM/2) x (BN/2) as
// 16 x 16 x 4 MFMA tiles.  Operands are staged through LDS (A transposed to k-major) with the next
// k-slab (two for 32 x 32 tiles) in flight into registers while the current one is multiplied; one
// barrier per slab.
template <typename TA, typename TB, typename TO, typename TC, int BM, int BN, int BK>
__global__ __launch_bounds__(kGemmThreads) void ggemm_kernel(const GemmTile *__restrict__ tiles,
                                                             const KGroup *__restrict__ groups,
                                                             const KTerm *__restrict__ terms,
                                                             const TA *__restrict__ A, const TB *__restrict__ B,
                                                             TO *__restrict__ C, TC *__restrict__ P,
                                                             int32_t plain_group, int32_t plain_term,
                                                             int32_t accumulate) {
  using M = Mfma<TC>;
  constexpr int TM = BM / 2, TN = BN / 2, FM = TM / 16, FN = TN / 16;
  constexpr int A_PER = BM * BK / kGemmThreads, B_PER = BK * BN / kGemmThreads;
  constexpr int A_TPR = BK / A_PER, B_TPR = BN / B_PER;
  static_assert(FM >= 1 && FN >= 1 && A_PER >= 1 && B_PER >= 1, "tile too small for four waves");
  static_assert(BK % A_PER == 0 && BN % B_PER == 0 && BK % 4 == 0, "tile does not split over the threads");
  // FP32 tiles larger than 32 x 32 store their slabs depth-contiguous (As[row][k], Bs[col][k],
  // rows padded by four so that sixteen lanes reading sixteen rows hit distinct banks): a lane then
  // reads the four depths it feeds to four successive MFMAs with one 16-byte LDS read.  Other tiles
  // keep the depth-major layout (As[k][row], Bs[k][col]), which measured faster on 32 x 32.  Every
  // FP32 tile sums in one order: within each 16-deep chunk, MFMA kk takes depths 4 lk + kk from the
  // lane groups lk (A and B permuted alike), so the order depends on the chunk, not the tile shape.
  constexpr bool kFp32 = std::is_same<TC, float>::value;
  constexpr bool kDepthRows = kFp32 && FM * FN >= 2;
  static_assert(!kFp32 || BK % 16 == 0, "FP32 slabs are summed in 16-deep chunks");
  static_assert(!kDepthRows || (kGemmThreads % BN == 0 && (kGemmThreads / BN) * B_PER == BK),
                "depth-contiguous B slabs are loaded as whole columns of B_PER depths");
  constexpr int kRowPad = 4;
  constexpr int AR = kDepthRows ? BM : BK, AC = kDepthRows ? BK + kRowPad : BM + kPad;
  constexpr int BR = kDepthRows ? BN : BK, BC = kDepthRows ? BK + kRowPad : BN + kPad;
  __shared__ __align__(16) TC As[2][AR][AC];
  __shared__ __align__(16) TC Bs[2][BR][BC];

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
    const int32_t tiles_n = static_cast<int32_t>((groups[gi].n + BN - 1) / BN);
    m0 = static_cast<int32_t>(blockIdx.x / tiles_n) * BM;
    n0 = static_cast<int32_t>(blockIdx.x % tiles_n) * BN;
    t_begin = plain_term < 0 ? 0 : plain_term;
    t_end = plain_term < 0 ? 0 : plain_term + 1;
  }
  const KGroup group = groups[gi];
  const int64_t m = group.m, n = group.n;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int li = lane & 15, lk = lane >> 4;

  typename M::acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = M::zero();

  const int a_row = tid / A_TPR, a_k = (tid % A_TPR) * A_PER;
  // B: depth-major slabs take B_PER consecutive columns of one depth per thread; depth-contiguous
  // ones take B_PER consecutive depths of one column (lanes run along a row of B either way, and
  // the thread's store is then one contiguous LDS write)
  const int b_k = kDepthRows ? (tid / BN) * B_PER : tid / B_TPR;
  const int b_col = kDepthRows ? tid % BN : (tid % B_TPR) * B_PER;

  // The slabs of the sum: BK-deep k-ranges of each term in list order (terms of depth 0 skipped).
  while (t_begin < t_end && terms[t_begin].k <= 0) ++t_begin;
  int32_t slabs = 0;
  for (int32_t t = t_begin; t < t_end; ++t)
    if (terms[t].k > 0) slabs += static_cast<int32_t>((terms[t].k + BK - 1) / BK);

  // The load cursor: the slab whose loads are issued next.  Each slab is read through two buffer
  // resources based at its first element (A at row m0, column kb of the term's block; B at row kb,
  // column n0) whose ranges end where the term's blocks end (A's at the term's last element: a term
  // that is a depth range of a split product reads rows of the whole product).  A thread's elements
  // sit at fixed byte offsets from those bases (computed once per term), so every load is issued
  // unconditionally and the compiler counts the loads in flight.  Rows of A past m, rows of B past
  // the term's depth, and every load once the cursor has run past the last slab (range 0) read
  // zeros.  What a tile reads past n (B) or m (A) inside a range only reaches output rows and
  // columns that are never stored.  The one in-range overhang that reaches the output, A's columns
  // past the term's depth in its last slab, is masked when the slab is stored (`klim`).
  // The cursor's state is uniform: byte pointers to the next slab's first A and B elements, the
  // bytes from there to the end of each block, and the depth the term has left; a slab advances
  // them by constant steps.
  const char *a_ptr = nullptr, *b_ptr = nullptr;
  uint32_t a_left = 0, b_left = 0;
  int32_t k_left = 0;
  uint32_t a_off = 0;
  const uint32_t b_off = static_cast<uint32_t>((b_k * n + b_col) * sizeof(TB));
  const uint32_t b_step = static_cast<uint32_t>(BK * n * sizeof(TB));
  const uint32_t b_elem = kDepthRows ? static_cast<uint32_t>(n * sizeof(TB)) : static_cast<uint32_t>(sizeof(TB));
  constexpr uint32_t a_step = BK * sizeof(TA);
  int32_t lt = t_begin;
  auto enter = [&](const KTerm &tm) {
    a_ptr = reinterpret_cast<const char *>(A + tm.a_offset + m0 * tm.lda);
    a_left = static_cast<uint32_t>(((m - 1 - m0) * tm.lda + tm.k) * static_cast<int64_t>(sizeof(TA)));
    b_ptr = reinterpret_cast<const char *>(B + tm.b_offset + n0);
    b_left = static_cast<uint32_t>((tm.k * n - n0) * static_cast<int64_t>(sizeof(TB)));
    k_left = static_cast<int32_t>(tm.k);
    a_off = static_cast<uint32_t>((a_row * tm.lda + a_k) * sizeof(TA));
  };
  if (t_begin < t_end) enter(terms[t_begin]);

  auto issue = [&](TA (&xa)[A_PER], TB (&xb)[B_PER], int32_t &klim) {
    klim = k_left < BK ? k_left : BK;
    const __amdgpu_buffer_rsrc_t ra = resource(a_ptr, a_left);
    const __amdgpu_buffer_rsrc_t rb = resource(b_ptr, b_left);
#pragma unroll
    for (int j = 0; j < A_PER; ++j) xa[j] = BufferLoad<TA>::at(ra, a_off + j * sizeof(TA));
#pragma unroll
    for (int j = 0; j < B_PER; ++j) xb[j] = BufferLoad<TB>::at(rb, b_off + j * b_elem);
    a_ptr += a_step;
    a_left = a_left > a_step ? a_left - a_step : 0u;
    b_ptr += b_step;
    b_left = b_left > b_step ? b_left - b_step : 0u;
    k_left -= BK;
    if (k_left <= 0 && lt < t_end) {
      ++lt;
      while (lt < t_end && terms[lt].k <= 0) ++lt;
      if (lt < t_end) {
        enter(terms[lt]);
      } else {
        // past the last slab: every further load reads zeros, and the slabs are masked whole
        a_left = b_left = 0u;
        k_left = 0;
      }
    }
  };

  // kAhead slabs in flight, in as many register sets used in turn: two on FP32 tiles up to 64 x 64
  // (a set costs A_PER + B_PER registers) and on FP64 tiles of at most 32 x 32; one on larger FP64
  // tiles, where a second set costs occupancy, and on complex tiles.
  constexpr int kAhead = std::is_same<TC, cdouble>::value ? 1
                         : std::is_same<TC, float>::value ? (BM * BN <= 64 * 64 && BK == 16 ? 2 : 1)
                                                          : (BM * BN <= 32 * 32 ? 2 : 1);
  TA ra[kAhead][A_PER];
  TB rb[kAhead][B_PER];
  int32_t klim[kAhead] = {};
#pragma unroll
  for (int set = 0; set < kAhead; ++set)
    if (slabs > 0) issue(ra[set], rb[set], klim[set]);
  int buf = 0;
  // slot of depth k, row r of the A slab and depth k, column c of the B slab
  auto a_slot = [&](int k, int r) -> TC & {
    if constexpr (kDepthRows) return As[buf][r][k];
    else return As[buf][k][r];
  };
  auto b_slot = [&](int k, int c) -> TC & {
    if constexpr (kDepthRows) return Bs[buf][c][k];
    else return Bs[buf][k][c];
  };
  auto step = [&](TA (&xa)[A_PER], TB (&xb)[B_PER], int32_t &klim) {
    if (klim == BK) {
#pragma unroll
      for (int j = 0; j < A_PER; ++j) a_slot(a_k + j, a_row) = Conv<TC>::from(xa[j]);
    } else {
#pragma unroll
      for (int j = 0; j < A_PER; ++j) a_slot(a_k + j, a_row) = a_k + j < klim ? Conv<TC>::from(xa[j]) : Conv<TC>::zero();
    }
#pragma unroll
    for (int j = 0; j < B_PER; ++j) {
      if constexpr (kDepthRows) b_slot(b_k + j, b_col) = Conv<TC>::from(xb[j]);
      else b_slot(b_k, b_col + j) = Conv<TC>::from(xb[j]);
    }
    __syncthreads();
    // uniform; false only for a padding step over a slab of zeros, which one set never takes
    const bool live = kAhead == 1 || klim > 0;
    issue(xa, xb, klim);
    if (!live) {
      // a padding step past the last slab: nothing to add
    } else if constexpr (kDepthRows) {
#pragma unroll
      for (int chunk = 0; chunk < BK / 16; ++chunk) {
        f4 af[FM], bf[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) af[i] = *reinterpret_cast<const f4 *>(&As[buf][wm * TM + i * 16 + li][chunk * 16 + lk * 4]);
#pragma unroll
        for (int j = 0; j < FN; ++j) bf[j] = *reinterpret_cast<const f4 *>(&Bs[buf][wn * TN + j * 16 + li][chunk * 16 + lk * 4]);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = M::step(af[i][kk], bf[j][kk], acc[i][j]);
      }
    } else {
#pragma unroll
      for (int kk = 0; kk < BK / 4; ++kk) {
        const int k = kFp32 ? (kk / 4) * 16 + lk * 4 + kk % 4 : kk * 4 + lk;
        TC af[FM], bf[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) af[i] = As[buf][k][wm * TM + i * 16 + li];
#pragma unroll
        for (int j = 0; j < FN; ++j) bf[j] = Bs[buf][k][wn * TN + j * 16 + li];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) acc[i][j] = M::step(af[i], bf[j], acc[i][j]);
      }
    }
    buf ^= 1;
  };
  // Whole rounds of kAhead steps, with no exit between them; a count that is not a multiple of
  // kAhead ends with steps over slabs of zeros (loaded past the last slab), whose MFMAs are skipped.
  for (int32_t slab = 0; slab < slabs; slab += kAhead) {
#pragma unroll
    for (int set = 0; set < kAhead; ++set) step(ra[set], rb[set], klim[set]);
  }

#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm * TM + i * 16 + M::row(lk, r);
        const int64_t col = n0 + wn * TN + j * 16 + li;
        if (row < m && col < n) {
          const TC value = M::get(acc[i][j], r);
          if (group.partial) {
            P[group.c_offset + row * n + col] = value;
            continue;
          }
          TO *out = C + group.c_offset + row * n + col;
          if (accumulate)
            *out = Conv<TO>::from(add(Conv<TC>::from(*out), value));
          else
            *out = Conv<TO>::from(value);
        }
      }
}

// The second launch of a plan with split products: each element of a split block is the sum of
// its parts' partial sums, added in part order (so the result does not depend on the tile shape
// or on which part finished first), converted to the output type.  One element per thread, with
// the loads of all its parts issued before the adds (a part index past the last reads the last
// part again and is not added), so a thread waits for memory once.
template <typename TO, typename TC>
__global__ __launch_bounds__(kReduceThreads) void ggemm_reduce_kernel(const ReduceTile *__restrict__ rtiles,
                                                                      const RGroup *__restrict__ rgroups,
                                                                      const TC *__restrict__ P, TO *__restrict__ C) {
  const ReduceTile tile = rtiles[blockIdx.x];
  const RGroup g = rgroups[tile.rgroup];
  const int64_t e = tile.e0 + threadIdx.x;
  if (e >= g.elems) return;
  TC part[kMaxLoadedParts];
#pragma unroll
  for (int s = 0; s < kMaxLoadedParts; ++s) part[s] = P[g.p_offset + (s < g.parts ? s : g.parts - 1) * g.elems + e];
  TC sum = part[0];
#pragma unroll
  for (int s = 1; s < kMaxLoadedParts; ++s)
    if (s < g.parts) sum = add(sum, part[s]);
  C[g.c_offset + e] = Conv<TO>::from(sum);
}

// How the tiles of one launch are ordered.  Workgroup i of a launch runs on XCD i % 8, and each XCD
// has an L2 of its own.
enum Packing {
  kLongestFirst = 0,  // by decreasing K, so the last wave of the launch is made of short tiles
  kXcdBands = 1,      // each block product's tiles cut into row-major bands, one per XCD (a band's
                      // A rows and the product's B then share that XCD's L2)
};
constexpr int kXcds = 8;
constexpr int kNumPackings = 2;

// Host-side tile list and cost of a group list under one tile shape and packing.
// Groups flagged in `skip` (split products, whose parts are launched in their place) get no tile.
void tile_list(const std::vector<KGroup> &groups, const std::vector<KTerm> &terms, const std::vector<char> &skip,
               const Shape &shape, int packing, std::vector<GemmTile> &tiles, double *max_cost) {
  struct Costed {
    double cost;
    GemmTile tile;
  };
  // Tiles longest first: every tile of a product costs the same (its slab count), so the products
  // are ordered by cost (ties in list order) and each emits its tiles row-major, which is the
  // order a stable sort of all tiles by cost gives, without sorting the tiles.
  const int64_t num_groups = static_cast<int64_t>(groups.size());
  std::vector<double> group_cost(static_cast<size_t>(num_groups), 0.0);
  std::vector<int64_t> order;
  int64_t count_all = 0;
  *max_cost = 0.0;
  for (int64_t g = 0; g < num_groups; ++g) {
    if (skip[static_cast<size_t>(g)]) continue;
    const KGroup &grp = groups[static_cast<size_t>(g)];
    int64_t ksum = 0;
    for (int32_t t = grp.first_term; t < grp.first_term + grp.num_terms; ++t) ksum += terms[static_cast<size_t>(t)].k;
    const double cost = static_cast<double>((ksum + shape.bk - 1) / shape.bk) + 1.0;
    group_cost[static_cast<size_t>(g)] = cost;
    *max_cost = std::max(*max_cost, cost);
    order.push_back(g);
    count_all += ((grp.m + shape.bm - 1) / shape.bm) * ((grp.n + shape.bn - 1) / shape.bn);
  }
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    return group_cost[static_cast<size_t>(a)] > group_cost[static_cast<size_t>(b)];
  });
  std::vector<Costed> costed;
  costed.reserve(static_cast<size_t>(count_all));
  for (int64_t g : order) {
    const KGroup &grp = groups[static_cast<size_t>(g)];
    for (int64_t mt = 0; mt < grp.m; mt += shape.bm)
      for (int64_t nt = 0; nt < grp.n; nt += shape.bn)
        costed.push_back({group_cost[static_cast<size_t>(g)],
                          GemmTile{static_cast<int32_t>(g), static_cast<int32_t>(mt), static_cast<int32_t>(nt), 0}});
  }
  tiles.clear();
  tiles.reserve(costed.size());
  if (packing == kLongestFirst) {
    for (const Costed &c : costed) tiles.push_back(c.tile);
    return;
  }
  // XCD bands: the tiles of each block product, in row-major order, are cut into kXcds contiguous
  // ranges (bands of output rows when the product is taller than wide), range c on XCD
  // (c + rotation) % kXcds, the rotation advancing with the tiles of earlier products so that
  // products of few tiles spread over the XCDs.  Each XCD then reads the A rows of its band once
  // and B once, where round-robin placement has every XCD read all of A; and every XCD gets an
  // equal share of every product.  Each XCD keeps its tiles longest first, and position p of the
  // launch holds a tile of XCD p % 8; when an XCD runs out, the position takes a tile from the
  // fullest one.
  std::vector<int64_t> count(static_cast<size_t>(num_groups), 0), first(static_cast<size_t>(num_groups), 0);
  for (const Costed &c : costed) ++count[static_cast<size_t>(c.tile.group)];
  for (int64_t g = 1; g < num_groups; ++g)
    first[static_cast<size_t>(g)] = first[static_cast<size_t>(g - 1)] + count[static_cast<size_t>(g - 1)];
  std::vector<int64_t> seen(static_cast<size_t>(num_groups), 0);
  std::vector<int> xcd_of_tile(costed.size(), 0);
  for (size_t i = 0; i < costed.size(); ++i) {
    const size_t g = static_cast<size_t>(costed[i].tile.group);
    const int64_t index = seen[g]++;  // row-major: a product's tiles keep their order in `costed`
    xcd_of_tile[i] = static_cast<int>((first[g] + index * kXcds / count[g]) % kXcds);
  }
  std::vector<std::vector<GemmTile>> bins(kXcds);
  for (size_t i = 0; i < costed.size(); ++i) bins[static_cast<size_t>(xcd_of_tile[i])].push_back(costed[i].tile);
  std::vector<size_t> next(kXcds, 0);
  for (size_t p = 0; p < costed.size(); ++p) {
    int x = static_cast<int>(p % kXcds);
    if (next[static_cast<size_t>(x)] >= bins[static_cast<size_t>(x)].size()) {
      size_t most = 0;
      for (int y = 0; y < kXcds; ++y) {
        const size_t left = bins[static_cast<size_t>(y)].size() - next[static_cast<size_t>(y)];
        if (left > most) {
          most = left;
          x = y;
        }
      }
    }
    tiles.push_back(bins[static_cast<size_t>(x)][next[static_cast<size_t>(x)]++]);
  }
}

// The default configuration, from the number of 64 x 64 tiles of the launch (split products
// counted by their parts).  Measured on the DMRG matvec products (15-sector bonds, χ = 1024 and
// 2048, FP32 and FP64 MFMA): from 2048 tiles 64 x 64 tiles in XCD bands are best or within 3 % of
// the best; between 512 and 2048 tiles 64 x 32 tiles in XCD bands are best or within 8 %; below
// that the launch is short of parallel work and 32 x 32 tiles, longest first, win.
int choose_shape(const std::vector<KGroup> &groups, const std::vector<char> &skip) {
  int64_t tiles64 = 0;
  for (size_t g = 0; g < groups.size(); ++g)
    if (!skip[g]) tiles64 += ((groups[g].m + 63) / 64) * ((groups[g].n + 63) / 64);
  return tiles64 >= 2048 ? kNumShapes * kXcdBands + 0 : tiles64 >= 512 ? kNumShapes * kXcdBands + 3 : 1;
}

// Depth split.  A launch whose products are few and deep (the DMRG matvec's second step: 16-29
// blocks of at most ~350 x 350 with depths up to five times that) has too few output tiles to
// occupy the chip, and every tile walks its whole depth alone.  Such a product is cut along its
// depth into parts, each a product of its own that writes a partial sum; a second launch adds the
// parts in order.  The decision reads only the block shapes, never the tile shape, so every launch
// configuration of a plan sums in one order and gives the same values bit for bit.
constexpr int64_t kSplitTiles = 16 * 256;  // a plan with at least this many 32 x 32 tiles is not split
constexpr int64_t kSplitMinDepth = 512;    // nor is a product shallower than this
constexpr int64_t kSplitSide = 128;        // a part is about max(m, n, 128) deep
constexpr int64_t kMaxParts = 7;             // a split product has at most kMaxParts + 1 parts
constexpr int64_t kPieceAlign = 32;        // part boundaries at multiples of 32 within a term

// The device tables of a plan: the caller's groups and terms first (the per-block form indexes
// them), then the parts of split products and their term pieces; the tiles of one launch; the
// partial-sum records.
struct DeviceTables {
  std::vector<KGroup> groups;
  std::vector<KTerm> terms;
  std::vector<char> skip;  // split products (their parts are launched instead)
  std::vector<RGroup> rgroups;
  std::vector<ReduceTile> rtiles;
  int64_t partial_elems = 0;
};

void split_tables(const tt_gemm_group *groups, int64_t num_groups, const tt_gemm_term *terms, int64_t num_terms,
                  DeviceTables &d) {
  d.groups.clear();
  d.terms.clear();
  for (int64_t t = 0; t < num_terms; ++t) d.terms.push_back(KTerm{terms[t].a_offset, terms[t].b_offset, terms[t].k, terms[t].k});
  for (int64_t g = 0; g < num_groups; ++g)
    d.groups.push_back(KGroup{groups[g].c_offset, groups[g].m, groups[g].n, groups[g].first_term, groups[g].num_terms, 0, 0});
  d.skip.assign(static_cast<size_t>(num_groups), 0);
  int64_t tiles32 = 0;
  for (int64_t g = 0; g < num_groups; ++g) tiles32 += ((groups[g].m + 31) / 32) * ((groups[g].n + 31) / 32);
  if (tiles32 >= kSplitTiles) return;
  for (int64_t g = 0; g < num_groups; ++g) {
    const tt_gemm_group &grp = groups[g];
    int64_t depth = 0;
    for (int32_t t = grp.first_term; t < grp.first_term + grp.num_terms; ++t) depth += std::max<int64_t>(terms[t].k, 0);
    const int64_t side = std::max({grp.m, grp.n, kSplitSide});
    if (grp.m == 0 || grp.n == 0 || depth < kSplitMinDepth || depth < 2 * side) continue;
    const int64_t want = std::min(kMaxParts, depth / side);
    const int64_t chunk = ((depth + want - 1) / want + kPieceAlign - 1) / kPieceAlign * kPieceAlign;
    // Pieces of at most `chunk` along each term, in term order; a part collects consecutive pieces
    // until it is `chunk` deep.
    std::vector<std::vector<KTerm>> parts(1);
    int64_t part_depth = 0;
    for (int32_t t = grp.first_term; t < grp.first_term + grp.num_terms; ++t) {
      const tt_gemm_term &tm = terms[t];
      for (int64_t k0 = 0; k0 < tm.k; k0 += chunk) {
        if (part_depth >= chunk) {
          parts.emplace_back();
          part_depth = 0;
        }
        const int64_t len = std::min(chunk, tm.k - k0);
        parts.back().push_back(KTerm{tm.a_offset + k0, tm.b_offset + k0 * grp.n, len, tm.k});
        part_depth += len;
      }
    }
    // Every part but the last is at least `chunk` deep, so there are at most want + 1 parts; the
    // bound on the sum launch's loads is checked all the same.
    if (parts.size() < 2 || parts.size() > static_cast<size_t>(kMaxLoadedParts)) continue;
    const int64_t elems = grp.m * grp.n;
    d.skip[static_cast<size_t>(g)] = 1;
    const int32_t rg = static_cast<int32_t>(d.rgroups.size());
    d.rgroups.push_back(RGroup{grp.c_offset, elems, d.partial_elems, static_cast<int32_t>(parts.size()), 0});
    for (size_t s = 0; s < parts.size(); ++s) {
      d.groups.push_back(KGroup{d.partial_elems + static_cast<int64_t>(s) * elems, grp.m, grp.n,
                                static_cast<int32_t>(d.terms.size()), static_cast<int32_t>(parts[s].size()), 1, 0});
      d.terms.insert(d.terms.end(), parts[s].begin(), parts[s].end());
      d.skip.push_back(0);
    }
    for (int64_t e0 = 0; e0 < elems; e0 += kReduceSpan) d.rtiles.push_back(ReduceTile{rg, 0, e0});
    d.partial_elems += static_cast<int64_t>(parts.size()) * elems;
  }
}

// The partial-sum buffer: one per device, grown when a plan needs more and never freed, because a
// plan (or a graph captured from it) keeps the buffer that was current when it was made.  Split
// launches share it, so the library's contract of one stream per device at a time is what keeps
// two of them from overlapping.
constexpr int kMaxDevices = 64;
std::mutex g_partials_mutex;
void *g_partials[kMaxDevices] = {};
size_t g_partials_bytes[kMaxDevices] = {};

int partial_buffer(size_t bytes, void **out) {
  int device = 0;
  TT_HIP(hipGetDevice(&device));
  if (device < 0 || device >= kMaxDevices) return fail("grouped GEMM: device index out of range");
  std::lock_guard<std::mutex> lock(g_partials_mutex);
  if (g_partials_bytes[device] < bytes) {
    size_t want = std::max(bytes, 2 * g_partials_bytes[device]);
    want = (want + (size_t(1) << 20) - 1) & ~((size_t(1) << 20) - 1);
    void *fresh = nullptr;
    TT_HIP(hipMalloc(&fresh, want));
    g_partials[device] = fresh;
    g_partials_bytes[device] = want;
  }
  *out = g_partials[device];
  return 0;
}

}  // namespace

struct tt_ggemm_plan {
  void *block = nullptr;   // the plan's range of the table arena, holding its tables
  KGroup *groups = nullptr;
  KTerm *terms = nullptr;
  GemmTile *tiles = nullptr;
  RGroup *rgroups = nullptr;
  ReduceTile *rtiles = nullptr;
  void *partials = nullptr;  // partial sums of split products (16 bytes per element: any compute type)
  int64_t num_groups = 0, num_terms = 0, num_tiles = 0, num_rtiles = 0;
  int config = 0;
  double flops = 0.0;
  std::vector<tt_gemm_group> host_groups;
};

extern "C" int tt_ggemm_num_configs(void) { return kNumShapes * kNumPackings; }

extern "C" int tt_ggemm_config_shape(int config, int *rows, int *cols, int *depth) {
  if (config < 0 || config >= kNumShapes * kNumPackings) return fail("tt_ggemm_config_shape: no such configuration");
  const Shape &shape = kShapes[config % kNumShapes];
  if (rows) *rows = shape.bm;
  if (cols) *cols = shape.bn;
  if (depth) *depth = shape.bk;
  return 0;
}

extern "C" int tt_ggemm_plan_create_config(const tt_gemm_group *groups, int64_t num_groups, const tt_gemm_term *terms,
                                           int64_t num_terms, int config, tt_ggemm_plan **plan) {
  if (!plan) return fail("tt_ggemm_plan_create: null output handle");
  if (num_groups < 0 || num_terms < 0) return fail("tt_ggemm_plan_create: negative count");
  double flops = 0.0;
  for (int64_t g = 0; g < num_groups; ++g) {
    const tt_gemm_group &grp = groups[g];
    if (grp.m < 0 || grp.n < 0) return fail("tt_ggemm_plan_create: negative block shape");
    if (grp.num_terms < 0 || grp.first_term < 0 || grp.first_term + grp.num_terms > num_terms)
      return fail("tt_ggemm_plan_create: group terms out of range");
    // The kernel reads a tile's operands at 32-bit byte offsets from the tile's first element.
    if (grp.m >= kMaxSide || grp.n >= kMaxSide) return fail("tt_ggemm_plan_create: block too large (a side of 2^20 or more)");
    for (int32_t t = grp.first_term; t < grp.first_term + grp.num_terms; ++t) {
      if (terms[t].k < 0) return fail("tt_ggemm_plan_create: negative k");
      if (terms[t].k >= kMaxSide) return fail("tt_ggemm_plan_create: block too large (a depth of 2^20 or more)");
      if (grp.m * terms[t].k >= kMaxBlockElements || terms[t].k * grp.n >= kMaxBlockElements)
        return fail("tt_ggemm_plan_create: operand block too large (2^27 elements or more)");
      flops += 2.0 * static_cast<double>(grp.m) * static_cast<double>(grp.n) * static_cast<double>(terms[t].k);
    }
  }
  DeviceTables d;
  split_tables(groups, num_groups, terms, num_terms, d);
  if (config < 0) config = choose_shape(d.groups, d.skip);
  if (config >= kNumShapes * kNumPackings) return fail("tt_ggemm_plan_create: no such configuration");
  if (d.groups.size() >= (size_t(1) << 31) || d.terms.size() >= (size_t(1) << 31))
    return fail("tt_ggemm_plan_create: too many blocks");
  std::vector<GemmTile> tiles;
  double max_cost = 0.0;
  tile_list(d.groups, d.terms, d.skip, kShapes[config % kNumShapes], config / kNumShapes, tiles, &max_cost);
  void *partials = nullptr;
  if (d.partial_elems > 0)
    if (int rc = partial_buffer(static_cast<size_t>(d.partial_elems) * sizeof(cdouble), &partials)) return rc;
  auto *p = new tt_ggemm_plan();
  p->num_groups = num_groups;
  p->num_terms = num_terms;
  p->num_tiles = static_cast<int64_t>(tiles.size());
  p->num_rtiles = static_cast<int64_t>(d.rtiles.size());
  p->partials = partials;
  p->config = config;
  p->flops = flops;
  p->host_groups.assign(groups, groups + num_groups);
  if (ttdev::upload_tables(&p->block, {{reinterpret_cast<void **>(&p->groups), d.groups.data(), sizeof(KGroup) * d.groups.size()},
                                        {reinterpret_cast<void **>(&p->terms), d.terms.data(), sizeof(KTerm) * d.terms.size()},
                                        {reinterpret_cast<void **>(&p->tiles), tiles.data(), sizeof(GemmTile) * tiles.size()},
                                        {reinterpret_cast<void **>(&p->rgroups), d.rgroups.data(), sizeof(RGroup) * d.rgroups.size()},
                                        {reinterpret_cast<void **>(&p->rtiles), d.rtiles.data(), sizeof(ReduceTile) * d.rtiles.size()}})) {
    char keep[1024];
    tt_last_error(keep, sizeof(keep));
    tt_ggemm_plan_destroy(p);
    return fail(keep);
  }
  *plan = p;
  return 0;
}

extern "C" int tt_ggemm_plan_create(const tt_gemm_group *groups, int64_t num_groups, const tt_gemm_term *terms,
                                    int64_t num_terms, tt_ggemm_plan **plan) {
  return tt_ggemm_plan_create_config(groups, num_groups, terms, num_terms, -1, plan);
}

extern "C" int tt_ggemm_plan_destroy(tt_ggemm_plan *plan) {
  if (!plan) return 0;
  ttdev::release_tables(plan->block);
  delete plan;
  return 0;
}

extern "C" int tt_ggemm_plan_counts(const tt_ggemm_plan *plan, double *flops, int64_t *tiles) {
  if (!plan) return fail("tt_ggemm_plan_counts: null plan");
  if (flops) *flops = plan->flops;
  if (tiles) *tiles = plan->num_tiles;
  return 0;
}

extern "C" int tt_ggemm_plan_config(const tt_ggemm_plan *plan) { return plan ? plan->config : -1; }

namespace {

template <typename TA, typename TB, typename TO, typename TC, int BM, int BN, int BK>
int ggemm_launch(const tt_ggemm_plan *plan, const void *a, const void *b, void *c, hipStream_t stream, int64_t grid,
                 int32_t plain_group, int32_t plain_term, int32_t accumulate) {
  if (grid <= 0) return 0;
  if (int rc = ttdev::await_uploads(stream)) return rc;
  if (int rc = ttdev::launch(ggemm_kernel<TA, TB, TO, TC, BM, BN, BK>, dim3(static_cast<uint32_t>(grid)),
                             dim3(kGemmThreads), 0, stream, plan->tiles, plan->groups, plan->terms,
                             static_cast<const TA *>(a), static_cast<const TB *>(b), static_cast<TO *>(c),
                             static_cast<TC *>(plan->partials), plain_group, plain_term, accumulate))
    return rc;
  TT_LAUNCHED();
  ttdev::note_launch();
  return 0;
}

// The partial sums of a plan's split products, added into C (after the batched launch, in order on
// the same stream).
template <typename TO, typename TC>
int reduce_launch(const tt_ggemm_plan *plan, void *c, hipStream_t stream) {
  if (plan->num_rtiles == 0) return 0;
  if (int rc = ttdev::launch(ggemm_reduce_kernel<TO, TC>, dim3(static_cast<uint32_t>(plan->num_rtiles)),
                             dim3(kReduceThreads), 0, stream, plan->rtiles, plan->rgroups,
                             static_cast<const TC *>(plan->partials), static_cast<TO *>(c)))
    return rc;
  TT_LAUNCHED();
  ttdev::note_launch();
  return 0;
}

template <typename TA, typename TB, typename TO, typename TC>
int ggemm_tiles(const tt_ggemm_plan *plan, const void *a, const void *b, void *c, hipStream_t stream, int64_t grid,
                int32_t pg, int32_t pt, int32_t acc) {
  switch (plan->config % kNumShapes) {
    case 0: return ggemm_launch<TA, TB, TO, TC, 64, 64, 16>(plan, a, b, c, stream, grid, pg, pt, acc);
    case 1: return ggemm_launch<TA, TB, TO, TC, 32, 32, 16>(plan, a, b, c, stream, grid, pg, pt, acc);
    case 2: return ggemm_launch<TA, TB, TO, TC, 128, 64, 16>(plan, a, b, c, stream, grid, pg, pt, acc);
    case 3: return ggemm_launch<TA, TB, TO, TC, 64, 32, 16>(plan, a, b, c, stream, grid, pg, pt, acc);
    case 4: return ggemm_launch<TA, TB, TO, TC, 32, 64, 16>(plan, a, b, c, stream, grid, pg, pt, acc);
    case 5: return ggemm_launch<TA, TB, TO, TC, 128, 128, 16>(plan, a, b, c, stream, grid, pg, pt, acc);
    case 6: return ggemm_launch<TA, TB, TO, TC, 64, 64, 32>(plan, a, b, c, stream, grid, pg, pt, acc);
    default: return fail("grouped GEMM: no such tile shape");
  }
}

template <typename TA, typename TB, typename TO, typename TC>
int ggemm_shape(const tt_ggemm_plan *plan, const void *a, const void *b, void *c, hipStream_t stream, int64_t grid,
                int32_t pg, int32_t pt, int32_t acc) {
  if (int rc = ggemm_tiles<TA, TB, TO, TC>(plan, a, b, c, stream, grid, pg, pt, acc)) return rc;
  return pg < 0 ? reduce_launch<TO, TC>(plan, c, stream) : 0;
}

int ggemm_dispatch(const tt_ggemm_plan *plan, int at, int bt, int ct, int comp, const void *a, const void *b, void *c,
                   hipStream_t stream, int64_t grid, int32_t pg, int32_t pt, int32_t acc) {
#define TT_CASE(A_, B_, C_, K_, TA, TB, TO, TC) \
  if (at == A_ && bt == B_ && ct == C_ && comp == K_) return ggemm_shape<TA, TB, TO, TC>(plan, a, b, c, stream, grid, pg, pt, acc);
  TT_CASE(TT_F64, TT_F64, TT_F64, TT_F64, double, double, double, double)
  TT_CASE(TT_F32, TT_F32, TT_F32, TT_F32, float, float, float, float)
  TT_CASE(TT_F32, TT_F64, TT_F32, TT_F32, float, double, float, float)
  TT_CASE(TT_F32, TT_F32, TT_F64, TT_F32, float, float, double, float)
  TT_CASE(TT_F64, TT_F64, TT_F64, TT_F32, double, double, double, float)
  TT_CASE(TT_F64, TT_F64, TT_F32, TT_F32, double, double, float, float)
  TT_CASE(TT_F32, TT_F64, TT_F64, TT_F32, float, double, double, float)
  TT_CASE(TT_F64, TT_F32, TT_F32, TT_F32, double, float, float, float)
  TT_CASE(TT_F64, TT_F32, TT_F64, TT_F32, double, float, double, float)
  TT_CASE(TT_C128, TT_C128, TT_C128, TT_C128, cdouble, cdouble, cdouble, cdouble)
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
  if (dtype != TT_F64 && dtype != TT_F32 && dtype != TT_C128) return fail("tt_ggemm_execute_plain: unsupported dtype");
  const Shape &s = kShapes[plan->config % kNumShapes];
  for (int64_t g = 0; g < plan->num_groups; ++g) {
    const tt_gemm_group &grp = plan->host_groups[static_cast<size_t>(g)];
    const int64_t grid = ((grp.m + s.bm - 1) / s.bm) * ((grp.n + s.bn - 1) / s.bn);
    if (grp.num_terms == 0) {
      // A block with no term still owes its zeros: one launch over an empty term range.
      if (int rc = ggemm_dispatch(plan, dtype, dtype, dtype, dtype, a, b, c, as_stream(stream), grid,
                                  static_cast<int32_t>(g), -1, 0))
        return rc;
      continue;
    }
    for (int32_t t = 0; t < grp.num_terms; ++t)
      if (int rc = ggemm_dispatch(plan, dtype, dtype, dtype, dtype, a, b, c, as_stream(stream), grid,
                                  static_cast<int32_t>(g), grp.first_term + t, t > 0 ? 1 : 0))
        return rc;
  }
  return 0;
}
