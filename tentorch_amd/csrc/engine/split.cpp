#include "split.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>

namespace tt {

std::string device_error();

namespace {

constexpr int64_t kTile = 64;    // the bgemm kernel's tile side
constexpr int64_t kAlign = 256;  // staging parts start on 256-byte boundaries

int64_t aligned(int64_t bytes) { return (bytes + kAlign - 1) / kAlign * kAlign; }

int64_t tiles_of(int64_t m, int64_t n) { return ((m + kTile - 1) / kTile) * ((n + kTile - 1) / kTile); }

tt_bgemm_problem gemm(uintptr_t a, uintptr_t b, uintptr_t c, int64_t m, int64_t n, int64_t k, int64_t lda,
                      int64_t ldb, int64_t ldc) {
  tt_bgemm_problem p{};
  p.a = reinterpret_cast<const double *>(a);
  p.b = reinterpret_cast<const double *>(b);
  p.c = reinterpret_cast<double *>(c);
  p.m = static_cast<int32_t>(m);
  p.n = static_cast<int32_t>(n);
  p.k = static_cast<int32_t>(k);
  p.lda = static_cast<int32_t>(lda);
  p.ldb = static_cast<int32_t>(ldb);
  p.ldc = static_cast<int32_t>(ldc);
  p.alpha = 1.0;
  return p;
}

tt_lincomb_problem lincomb(uintptr_t x, uintptr_t dst, int64_t rows, int64_t cols, int64_t x_ld, int64_t dst_ld,
                           double alpha, double gamma) {
  tt_lincomb_problem p{};
  p.x = reinterpret_cast<const double *>(x);
  p.dst = reinterpret_cast<double *>(dst);
  p.rows = static_cast<int32_t>(rows);
  p.cols = static_cast<int32_t>(cols);
  p.x_ld = static_cast<int32_t>(x_ld);
  p.dst_ld = static_cast<int32_t>(dst_ld);
  p.alpha = alpha;
  p.gamma = gamma;
  return p;
}

// The tail of a pass starts at the widest gap between consecutive values (descending) among the
// values within (1, 10) times `level`, so that no near-degenerate pair straddles the cut; with no
// value there, at `first_below`.
int64_t tail_start(const double *values, int64_t size, int64_t first_below, double level) {
  int64_t lo_index = -1, hi_index = -1;
  for (int64_t i = 0; i < size; ++i)
    if (values[i] > level && values[i] < 10.0 * level) {
      if (lo_index < 0) lo_index = i;
      hi_index = i;
    }
  if (lo_index < 0) return first_below;
  const int64_t lo = std::max<int64_t>(1, lo_index), hi = std::min<int64_t>(size - 1, hi_index + 1);
  if (hi < lo) return first_below;
  int64_t best = lo;
  double widest = -1.0;
  for (int64_t i = lo; i <= hi; ++i) {
    const double ratio = values[i - 1] / std::max(values[i], std::numeric_limits<double>::min());
    if (ratio > widest) {
      widest = ratio;
      best = i;
    }
  }
  return best;
}

}  // namespace

std::vector<std::pair<int64_t, int64_t>> choose_tails(const double *norms, const int64_t *norms_at, int64_t blocks,
                                                      const int64_t *ranks, const int64_t *first, double floor,
                                                      double reliable, double resolved) {
  std::vector<std::pair<int64_t, int64_t>> tails;
  for (int64_t b = 0; b < blocks; ++b) {
    const double *v = norms + norms_at[b] + first[b];
    const int64_t rows = norms_at[b + 1] - norms_at[b] - first[b], size = std::max<int64_t>(ranks[b] - first[b], 0);
    if (size < 2) continue;
    double top = 0.0;
    for (int64_t i = 0; i < size; ++i) top = std::max(top, v[i]);
    if (top == 0.0 || floor >= 2.0 * resolved * top) continue;
    double unresolved = 0.0;
    for (int64_t i = 0; i < rows; ++i)
      if (v[i] < 2.0 * resolved * top) unresolved += v[i] * v[i];
    if (std::sqrt(unresolved) < 0.5 * floor) continue;
    int64_t below = -1;
    for (int64_t i = 0; i < size && below < 0; ++i)
      if (v[i] < reliable * top) below = i;
    if (below < 0) continue;
    const int64_t start = tail_start(v, size, below, reliable * top);
    if (start > 0) tails.emplace_back(b, first[b] + start);
  }
  return tails;
}

SplitRanking rank_split(const Structure &structure, const std::vector<int64_t> &blocks, const double *values,
                        const std::vector<int64_t> &at, int64_t max_bond, double tol) {
  std::vector<Sector> sector(blocks.size());
  for (size_t i = 0; i < blocks.size(); ++i) {
    const int64_t b = blocks[i];
    if (b < 0 || b >= static_cast<int64_t>(structure.blocks().size())) throw std::invalid_argument("rank_split: block");
    sector[i] = structure.blocks()[static_cast<size_t>(b)].key.domain;
  }
  return rank_split(*structure.cat(), sector, values, at, max_bond, tol);
}

SplitRanking rank_split(const Category &cat, const std::vector<Sector> &sector, const double *values,
                        const std::vector<int64_t> &at, int64_t max_bond, double tol) {
  const int64_t count = static_cast<int64_t>(sector.size());
  if (static_cast<int64_t>(at.size()) != count + 1 || at.front() != 0)
    throw std::invalid_argument("rank_split: value offsets");
  for (int64_t i = 0; i < count; ++i)
    if (at[static_cast<size_t>(i) + 1] < at[static_cast<size_t>(i)]) throw std::invalid_argument("rank_split: value offsets");
  const int64_t total = at.back();
  for (int64_t i = 0; i < total; ++i)
    if (std::isnan(values[i])) throw std::invalid_argument("a singular value is not a number");
  // candidates, largest first; a stable sort keeps ties in (block, position) order
  std::vector<int64_t> order(static_cast<size_t>(total));
  for (int64_t i = 0; i < total; ++i) order[static_cast<size_t>(i)] = i;
  std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return values[x] > values[y]; });
  int64_t live = 0;
  while (live < total && !(values[order[static_cast<size_t>(live)]] <= tol)) ++live;
  if (max_bond >= 0) live = std::min(live, max_bond);
  std::vector<int64_t> owner(static_cast<size_t>(total));
  for (int64_t i = 0; i < count; ++i)
    for (int64_t j = at[static_cast<size_t>(i)]; j < at[static_cast<size_t>(i) + 1]; ++j) owner[static_cast<size_t>(j)] = i;
  SplitRanking out;
  out.kept.assign(static_cast<size_t>(count), 0);
  for (int64_t c = 0; c < live; ++c) {
    const int64_t j = order[static_cast<size_t>(c)], i = owner[static_cast<size_t>(j)];
    out.kept[static_cast<size_t>(i)] = std::max(out.kept[static_cast<size_t>(i)], j - at[static_cast<size_t>(i)] + 1);
  }
  std::map<Sector, int64_t> mults;
  for (int64_t i = 0; i < count; ++i)
    if (out.kept[static_cast<size_t>(i)]) mults[sector[static_cast<size_t>(i)]] += out.kept[static_cast<size_t>(i)];
  std::map<Sector, int64_t> paired;
  for (const auto &[s, m] : mults) {
    const auto other = mults.find(cat.dual(s));
    if (other != mults.end()) paired[s] = std::min(m, other->second);
  }
  if (!paired.empty() && paired != mults) {
    // a sector keeps its first candidates, up to its paired count (none without a dual)
    std::map<Sector, int64_t> taken;
    std::fill(out.kept.begin(), out.kept.end(), 0);
    for (int64_t c = 0; c < live; ++c) {
      const int64_t j = order[static_cast<size_t>(c)], i = owner[static_cast<size_t>(j)];
      const Sector &s = sector[static_cast<size_t>(i)];
      const auto room = paired.find(s);
      const int64_t rank = taken[s]++;
      if (room != paired.end() && rank < room->second)
        out.kept[static_cast<size_t>(i)] = std::max(out.kept[static_cast<size_t>(i)], j - at[static_cast<size_t>(i)] + 1);
    }
    mults = paired;
  }
  std::map<Sector, int64_t> next, placed;
  out.offset.resize(static_cast<size_t>(count));
  for (int64_t i = 0; i < count; ++i) {
    int64_t &n = next[sector[static_cast<size_t>(i)]];
    out.offset[static_cast<size_t>(i)] = n;
    n += out.kept[static_cast<size_t>(i)];
  }
  int64_t run = 0;
  for (const auto &[s, m] : mults) {
    placed[s] = run;
    run += m;
    out.sectors.push_back(s);
    out.mults.push_back(m);
  }
  out.values.assign(static_cast<size_t>(run == 0 ? 1 : run), 0.0);
  for (int64_t i = 0; i < count; ++i) {
    const int64_t k = out.kept[static_cast<size_t>(i)];
    if (!k) continue;
    const int64_t to = placed[sector[static_cast<size_t>(i)]] + out.offset[static_cast<size_t>(i)];
    std::copy(values + at[static_cast<size_t>(i)], values + at[static_cast<size_t>(i)] + k, out.values.begin() + to);
  }
  if (mults.empty()) {
    out.sectors.push_back(cat.trivial());
    out.mults.push_back(1);
  }
  return out;
}

TailPass::TailPass(const std::vector<int64_t> &g, const std::vector<int64_t> &n, const std::vector<int64_t> &ld,
                   const std::vector<int64_t> &pq_at, const std::vector<int64_t> &block,
                   const std::vector<int64_t> &first, const std::vector<int64_t> &tail)
    : b_(block), f_(first), h_(tail) {
  const size_t count = b_.size();
  if (f_.size() != count || h_.size() != count) throw std::invalid_argument("tail pass: heads of unequal length");
  for (size_t i = 0; i < count; ++i) {
    const int64_t b = b_[i];
    if (b < 0 || b >= static_cast<int64_t>(g.size()) || n.size() != g.size() || ld.size() != g.size() ||
        pq_at.size() != g.size())
      throw std::invalid_argument("tail pass: a head names no block");
    if (!(0 <= f_[i] && f_[i] < h_[i] && h_[i] < g[static_cast<size_t>(b)]))
      throw std::invalid_argument("tail pass: a head's rows are not first < tail < rows");
    e_.push_back(h_[i] - f_[i]);
    q_.push_back(g[static_cast<size_t>(b)] - h_[i]);
    w_.push_back(ld[static_cast<size_t>(b)]);
    n_.push_back(n[static_cast<size_t>(b)]);
    pq_at_.push_back(pq_at[static_cast<size_t>(b)]);
  }
  // work per head: X (q x e), U, H', M (e x w), T' (q x w), K, A, B (e x e), heads one after another
  for (size_t i = 0; i < count; ++i) {
    const int64_t e = e_[i], q = q_[i], w = w_[i];
    for (int64_t size : {q * e, e * w, e * w, e * w, q * w, e * e, e * e, e * e}) {
      work_at_.push_back(work_total_);
      work_total_ += size;
    }
    head_at_.push_back(i == 0 ? 0 : head_at_.back() + e_[i - 1]);
    tail_at_.push_back(i == 0 ? 0 : tail_at_.back() + q_[i - 1]);
    sq_at_.push_back(gram_total_);
    gram_total_ += q * q;
    nr_at_.push_back(norm_total_);
    norm_total_ += q;
  }
  int64_t heads_total = 0;
  for (int64_t e : e_) heads_total += e;
  scale_at_ = 0;
  shift_at_ = aligned(8 * heads_total);
  stage_total_ = shift_at_ + aligned(8 * norm_total_);
  const int64_t c = static_cast<int64_t>(count);
  // kinds and problem counts of the ten tables in launch order
  const std::pair<int, int64_t> layout[] = {{0, c}, {0, 2 * c}, {1, 4 * c}, {0, 2 * c}, {0, c},
                                            {0, c}, {1, c},     {0, c},     {0, c},     {2, c}};
  for (const auto &[kind, problems] : layout) {
    Table t;
    t.kind = kind;
    t.count = problems;
    t.at = stage_total_;
    const int64_t bytes = problems * static_cast<int64_t>(kind == 0 ? sizeof(tt_bgemm_problem)
                                                          : kind == 1 ? sizeof(tt_lincomb_problem)
                                                                      : sizeof(tt_rows_problem));
    stage_total_ += aligned(std::max<int64_t>(bytes, 1));
    tables_.push_back(t);
  }
}

void TailPass::fill(uintptr_t pq, int64_t pq_bytes, uintptr_t work, uintptr_t grams, uintptr_t vectors,
                    uintptr_t norms, uintptr_t stage_device, void *stage_host, const double *norm_values,
                    const int64_t *norms_at) {
  const size_t count = b_.size();
  char *host = static_cast<char *>(stage_host);
  // squared row norms of the heads (the rotation's column scale) and of the tails (its row shift)
  double *scale = reinterpret_cast<double *>(host + scale_at_), *shift = reinterpret_cast<double *>(host + shift_at_);
  for (size_t i = 0; i < count; ++i) {
    const double *v = norm_values + norms_at[b_[i]];
    for (int64_t r = f_[i]; r < h_[i]; ++r) *scale++ = v[r] * v[r];
    for (int64_t r = h_[i]; r < h_[i] + q_[i]; ++r) *shift++ = v[r] * v[r];
  }
  auto at = [&](size_t i, int which) { return work + 8 * static_cast<uintptr_t>(work_at_[8 * i + which]); };
  auto row = [&](size_t i, int64_t r) { return pq + 8 * static_cast<uintptr_t>(pq_at_[i] + r * w_[i]); };
  std::vector<std::vector<tt_bgemm_problem>> gemms(10);
  std::vector<std::vector<tt_lincomb_problem>> lincombs(10);
  std::vector<tt_rows_problem> row_norms;
  // 0: X = diag-divided T_P H_P^T
  for (size_t i = 0; i < count; ++i) {
    tt_bgemm_problem p = gemm(row(i, h_[i]), row(i, f_[i]), at(i, 0), q_[i], e_[i], n_[i], w_[i], w_[i], e_[i]);
    p.trans_b = 1;
    p.col_scale = reinterpret_cast<const double *>(stage_device + scale_at_ + 8 * head_at_[i]);
    p.row_shift = reinterpret_cast<const double *>(stage_device + shift_at_ + 8 * tail_at_[i]);
    gemms[0].push_back(p);
  }
  // 1: K = X^T X, then U = X^T T
  for (size_t i = 0; i < count; ++i) {
    tt_bgemm_problem p = gemm(at(i, 0), at(i, 0), at(i, 5), e_[i], e_[i], q_[i], e_[i], e_[i], e_[i]);
    p.trans_a = 1;
    gemms[1].push_back(p);
  }
  for (size_t i = 0; i < count; ++i) {
    tt_bgemm_problem p = gemm(at(i, 0), row(i, h_[i]), at(i, 1), e_[i], w_[i], q_[i], e_[i], w_[i], w_[i]);
    p.trans_a = 1;
    gemms[1].push_back(p);
  }
  // 2: A = 1 - K/2, B = 1 - K/6, M = U/2, T' = T
  for (size_t i = 0; i < count; ++i) lincombs[2].push_back(lincomb(at(i, 5), at(i, 6), e_[i], e_[i], e_[i], e_[i], -0.5, 1.0));
  for (size_t i = 0; i < count; ++i)
    lincombs[2].push_back(lincomb(at(i, 5), at(i, 7), e_[i], e_[i], e_[i], e_[i], -1.0 / 6.0, 1.0));
  for (size_t i = 0; i < count; ++i) lincombs[2].push_back(lincomb(at(i, 1), at(i, 3), e_[i], w_[i], w_[i], w_[i], 0.5, 0.0));
  for (size_t i = 0; i < count; ++i)
    lincombs[2].push_back(lincomb(row(i, h_[i]), at(i, 4), q_[i], w_[i], w_[i], w_[i], 1.0, 0.0));
  // 3: H' = A H, then M = B H + M
  for (size_t i = 0; i < count; ++i) gemms[3].push_back(gemm(at(i, 6), row(i, f_[i]), at(i, 2), e_[i], w_[i], e_[i], e_[i], w_[i], w_[i]));
  for (size_t i = 0; i < count; ++i) {
    tt_bgemm_problem p = gemm(at(i, 7), row(i, f_[i]), at(i, 3), e_[i], w_[i], e_[i], e_[i], w_[i], w_[i]);
    p.beta = 1.0;
    gemms[3].push_back(p);
  }
  // 4: H' += B U
  for (size_t i = 0; i < count; ++i) {
    tt_bgemm_problem p = gemm(at(i, 7), at(i, 1), at(i, 2), e_[i], w_[i], e_[i], e_[i], w_[i], w_[i]);
    p.beta = 1.0;
    gemms[4].push_back(p);
  }
  // 5: T' -= X M
  for (size_t i = 0; i < count; ++i) {
    tt_bgemm_problem p = gemm(at(i, 0), at(i, 3), at(i, 4), q_[i], w_[i], e_[i], e_[i], w_[i], w_[i]);
    p.alpha = -1.0;
    p.beta = 1.0;
    gemms[5].push_back(p);
  }
  // 6: the heads back in place
  for (size_t i = 0; i < count; ++i) lincombs[6].push_back(lincomb(at(i, 2), row(i, f_[i]), e_[i], w_[i], w_[i], w_[i], 1.0, 0.0));
  // 7: the tails' negated Gram matrices -T'_P T'_P^T
  for (size_t i = 0; i < count; ++i) {
    tt_bgemm_problem p = gemm(at(i, 4), at(i, 4), grams + 8 * static_cast<uintptr_t>(sq_at_[i]), q_[i], q_[i], n_[i],
                              w_[i], w_[i], q_[i]);
    p.trans_b = 1;
    p.alpha = -1.0;
    gemms[7].push_back(p);
  }
  // 8: the tails' rows W T' in place of T
  for (size_t i = 0; i < count; ++i)
    gemms[8].push_back(gemm(vectors + 8 * static_cast<uintptr_t>(sq_at_[i]), at(i, 4), row(i, h_[i]), q_[i], w_[i],
                            q_[i], q_[i], w_[i], w_[i]));
  // 9: their row norms
  for (size_t i = 0; i < count; ++i) {
    tt_rows_problem p{};
    p.src = reinterpret_cast<const double *>(row(i, h_[i]));
    p.dst = reinterpret_cast<double *>(norms + 8 * static_cast<uintptr_t>(nr_at_[i]));
    p.rows = static_cast<int32_t>(q_[i]);
    p.cols = static_cast<int32_t>(n_[i]);
    p.src_ld = static_cast<int32_t>(w_[i]);
    p.dst_ld = 1;
    row_norms.push_back(p);
  }
  // number the work items and check every table against the buffers of the pass
  std::vector<std::pair<int64_t, int64_t>> spans = {
      {static_cast<int64_t>(pq), static_cast<int64_t>(pq) + pq_bytes},
      {static_cast<int64_t>(work), static_cast<int64_t>(work) + 8 * work_total_},
      {static_cast<int64_t>(grams), static_cast<int64_t>(grams) + 8 * gram_total_},
      {static_cast<int64_t>(vectors), static_cast<int64_t>(vectors) + 8 * gram_total_},
      {static_cast<int64_t>(norms), static_cast<int64_t>(norms) + 8 * norm_total_},
      {static_cast<int64_t>(stage_device), static_cast<int64_t>(stage_device) + stage_total_}};
  std::sort(spans.begin(), spans.end());
  std::vector<int64_t> starts, ends;
  for (const auto &[a, z] : spans) {
    if (!starts.empty() && a < ends.back()) {
      ends.back() = std::max(ends.back(), z);
    } else {
      starts.push_back(a);
      ends.push_back(z);
    }
  }
  const int64_t nspans = static_cast<int64_t>(starts.size());
  for (size_t t = 0; t < tables_.size(); ++t) {
    Table &table = tables_[t];
    int64_t items = 0;
    int status = 0;
    if (table.kind == 0) {
      std::vector<tt_bgemm_problem> &ps = gemms[t];
      for (tt_bgemm_problem &p : ps) {
        p.first = items;
        items += tiles_of(p.m, p.n);
      }
      status = tt_bgemm_check(ps.data(), static_cast<int64_t>(ps.size()), starts.data(), ends.data(), nspans, &table.items);
      std::memcpy(host + table.at, ps.data(), ps.size() * sizeof(tt_bgemm_problem));
    } else if (table.kind == 1) {
      std::vector<tt_lincomb_problem> &ps = lincombs[t];
      for (tt_lincomb_problem &p : ps) {
        p.first = items;
        items += p.rows;
      }
      status = tt_lincomb_check(ps.data(), static_cast<int64_t>(ps.size()), starts.data(), ends.data(), nspans,
                                &table.items);
      std::memcpy(host + table.at, ps.data(), ps.size() * sizeof(tt_lincomb_problem));
    } else {
      for (tt_rows_problem &p : row_norms) {
        p.first = items;
        items += p.rows;
      }
      status = tt_rows_check(row_norms.data(), static_cast<int64_t>(row_norms.size()), nullptr, 1, starts.data(),
                             ends.data(), nspans, &table.items);
      std::memcpy(host + table.at, row_norms.data(), row_norms.size() * sizeof(tt_rows_problem));
    }
    if (status != 0) throw std::runtime_error("tail pass table " + std::to_string(t) + ": " + device_error());
    if (table.items != items) throw std::runtime_error("tail pass table " + std::to_string(t) + ": work-item count");
  }
}

}  // namespace tt
