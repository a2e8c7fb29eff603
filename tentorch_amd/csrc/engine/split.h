// Host side of the split's further passes (the two-site split, tentorch/linalg/graded/svd.cpp:41-177):
// the launch tables of one tail pass over [P | Q] -- the decoupling rotation of each tail from its
// head rows, then the Gram product whose eigenvectors diagonalise the tail, the projection on them
// and its row norms -- laid out in one staging buffer, filled, and checked against the buffers they
// name; the caller copies the buffer to the device and launches them in order.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

#include "../../../include/tentorch_amd.h"
#include "tree.h"

namespace tt {

// The blocks whose current pass needs another, as (block, first row of the tail): the rule of
// tentorch_amd.gram.choose_tails (its docstring gives the reasoning) over every block's row norms
// `norms` (block b's at norms_at[b] .. norms_at[b + 1]), the pass of block b covering rows first[b]..
// and its values being the first ranks[b] rows.
std::vector<std::pair<int64_t, int64_t>> choose_tails(const double *norms, const int64_t *norms_at, int64_t blocks,
                                                      const int64_t *ranks, const int64_t *first, double floor,
                                                      double reliable, double resolved);

// The ranking of split_over (tentorch/linalg/graded/svd.cpp:72-136) over the singular values of
// blocks with domain sectors `sectors` (values[at[i] .. at[i+1]) descending for block i; a NaN is an
// error): candidates largest
// first (ties in block and position order), the ones at or below `tol` dropped, the first `max_bond`
// kept (all when negative), each block keeping the prefix up to its last candidate; where a bond
// sector's dual is on the bond too, both keep min of their counts, re-walking the candidates so a
// sector keeps its first ones.  Fills the kept count and the bond offset of every block, the bond's
// sectors and multiplicities (sorted; the trivial sector once when nothing is kept), and the kept
// values in bond order.
struct SplitRanking {
  std::vector<int64_t> kept, offset;
  std::vector<Sector> sectors;
  std::vector<int64_t> mults;
  std::vector<double> values;
};
SplitRanking rank_split(const Category &cat, const std::vector<Sector> &sectors, const double *values,
                        const std::vector<int64_t> &at, int64_t max_bond, double tol);
// The same over the blocks `blocks` (indices into structure.blocks()), each block's sector its key's
// domain sector.
SplitRanking rank_split(const Structure &structure, const std::vector<int64_t> &blocks, const double *values,
                        const std::vector<int64_t> &at, int64_t max_bond, double tol);

class TailPass {
 public:
  // Blocks: g (rows of [P | Q]), n (columns of P), ld (row length of [P | Q]) and the element offset
  // of each block's [P | Q]; heads: (block, first row of the pass, first row of the tail).
  TailPass(const std::vector<int64_t> &g, const std::vector<int64_t> &n, const std::vector<int64_t> &ld,
           const std::vector<int64_t> &pq_at, const std::vector<int64_t> &block, const std::vector<int64_t> &first,
           const std::vector<int64_t> &tail);

  // What the caller allocates: the decoupling's work doubles, the Gram matrices' doubles (their
  // eigenvectors take as many again), the norms' doubles, and the staging bytes of the tables.
  int64_t work_doubles() const { return work_total_; }
  int64_t gram_doubles() const { return gram_total_; }
  int64_t norm_doubles() const { return norm_total_; }
  int64_t stage_bytes() const { return stage_total_; }
  // Per head: the tail's order and the offset of its Gram matrix.
  const std::vector<int64_t> &orders() const { return q_; }
  const std::vector<int64_t> &gram_offsets() const { return sq_at_; }

  // Fill the staging buffer (host memory, later copied to `stage_device`) with the values the
  // rotation needs (squared row norms of heads and tails, from `norms` at `norms_at[block]`) and
  // every table, addressed at the given device buffers, and check each table against the byte
  // ranges of those buffers.  Throws std::runtime_error when a table reaches outside them.
  void fill(uintptr_t pq, int64_t pq_bytes, uintptr_t work, uintptr_t grams, uintptr_t vectors, uintptr_t norms,
            uintptr_t stage_device, void *stage_host, const double *norm_values, const int64_t *norms_at);

  // The tables in launch order: the decoupling's seven, the Gram products, then (after the
  // eigensolver) the projections and their row norms.
  struct Table {
    int kind;            // 0 bgemm, 1 lincomb, 2 row norms
    int64_t count = 0;   // problems
    int64_t items = 0;   // tiles or rows, as the library's check counted them
    int64_t at = 0;      // byte offset in the staging buffer
  };
  const std::vector<Table> &tables() const { return tables_; }

 private:
  std::vector<int64_t> b_, f_, h_, e_, q_, w_, n_, pq_at_;
  std::vector<int64_t> work_at_;  // 8 per head
  std::vector<int64_t> sq_at_, nr_at_, head_at_, tail_at_;
  int64_t work_total_ = 0, gram_total_ = 0, norm_total_ = 0, stage_total_ = 0;
  int64_t scale_at_ = 0, shift_at_ = 0;
  std::vector<Table> tables_;
};

}  // namespace tt
