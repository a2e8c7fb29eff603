#include "move.h"

#include <algorithm>
#include <map>
#include <numeric>
#include <stdexcept>
#include <tuple>

namespace tt {

std::string device_error();

std::vector<int64_t> complement(int64_t total, const std::vector<int64_t> &above) {
  std::vector<bool> named(static_cast<size_t>(std::max<int64_t>(total, 0)), false);
  for (int64_t leg : above)
    if (0 <= leg && leg < total) named[static_cast<size_t>(leg)] = true;
  std::vector<int64_t> below;
  for (int64_t leg = 0; leg < total; ++leg)
    if (!named[static_cast<size_t>(leg)]) below.push_back(leg);
  return below;
}

Chain build_chain(const Legs &codomain, const Legs &domain, const std::vector<int64_t> &above,
                  const std::vector<int64_t> &below) {
  const int64_t height = static_cast<int64_t>(codomain.size());
  const int64_t depth = static_cast<int64_t>(domain.size());
  const int64_t total = height + depth;
  std::vector<bool> named(static_cast<size_t>(total), false);
  for (const auto *side : {&above, &below})
    for (int64_t leg : *side) {
      if (leg < 0 || leg >= total)
        throw std::invalid_argument("the leg " + std::to_string(leg) + " does not exist, the tensor has " +
                                    std::to_string(total) + " legs");
      if (named[static_cast<size_t>(leg)]) throw std::invalid_argument("the leg " + std::to_string(leg) + " is named twice");
      named[static_cast<size_t>(leg)] = true;
    }
  if (static_cast<int64_t>(above.size() + below.size()) != total) {
    int64_t left_out = 0;
    while (named[static_cast<size_t>(left_out)]) ++left_out;
    throw std::invalid_argument("the leg " + std::to_string(left_out) + " is named on neither side");
  }
  Chain chain;
  Legs cod = codomain, dom = domain;
  auto add = [&](StepKind kind, int64_t position) {
    chain.steps.push_back(Step{kind, position});
    chain.sides.emplace_back(cod, dom);
  };
  auto exchange = [&](int64_t position) {
    std::swap(cod[static_cast<size_t>(position)], cod[static_cast<size_t>(position + 1)]);
    add(StepKind::exchange, position);
  };
  auto fold = [&](bool from_codomain) {
    if (from_codomain) {
      dom.push_back(cod.back());
      cod.pop_back();
    } else {
      cod.push_back(dom.back());
      dom.pop_back();
    }
    add(from_codomain ? StepKind::to_domain : StepKind::to_codomain, 0);
  };
  // The kept prefix of the domain, lifts, bubble sort and folds of place.cpp:224-246.
  bool descends = false;
  for (size_t at = 0; !descends && at < below.size(); ++at) descends = below[at] < height;
  int64_t kept = 0;
  while (!descends && kept < depth && kept < static_cast<int64_t>(below.size()) &&
         below[static_cast<size_t>(kept)] == height + kept)
    ++kept;
  for (int64_t at = kept; at < depth; ++at) fold(false);
  auto position_of = [height, depth](int64_t leg) { return leg < height ? leg : height + (depth - 1 - (leg - height)); };
  std::vector<int64_t> order;
  for (int64_t leg : above) order.push_back(position_of(leg));
  for (auto leg = below.rbegin(); leg != below.rend() - kept; ++leg) order.push_back(position_of(*leg));
  std::vector<int64_t> which(cod.size());
  std::iota(which.begin(), which.end(), 0);
  for (int64_t at = 0; at < static_cast<int64_t>(order.size()); ++at) {
    int64_t from = at;
    while (which[static_cast<size_t>(from)] != order[static_cast<size_t>(at)]) ++from;
    while (from > at) {
      exchange(from - 1);
      std::swap(which[static_cast<size_t>(from - 1)], which[static_cast<size_t>(from)]);
      --from;
    }
  }
  for (int64_t at = kept; at < static_cast<int64_t>(below.size()); ++at) fold(true);
  return chain;
}

namespace {

using Terms = std::vector<std::pair<int64_t, cplx>>;

// Tree pairs of a structure, numbered block by block, row tree major.
struct PairSpace {
  StructurePtr s;
  std::vector<int64_t> base, ncols;
  int64_t total = 0;
  explicit PairSpace(StructurePtr structure) : s(std::move(structure)) {
    for (const BlockInfo &b : s->blocks()) {
      base.push_back(total);
      ncols.push_back(static_cast<int64_t>(b.col_layout->trees.size()));
      total += static_cast<int64_t>(b.row_layout->trees.size()) * ncols.back();
    }
  }
  int64_t id(int64_t block, int64_t row, int64_t col) const {
    return base[static_cast<size_t>(block)] + row * ncols[static_cast<size_t>(block)] + col;
  }
  void decode(int64_t pair, int64_t *block, int64_t *row, int64_t *col) const {
    const auto it = std::upper_bound(base.begin(), base.end(), pair);
    const int64_t b = (it - base.begin()) - 1;
    const int64_t inside = pair - base[static_cast<size_t>(b)];
    *block = b;
    *row = inside / ncols[static_cast<size_t>(b)];
    *col = inside % ncols[static_cast<size_t>(b)];
  }
};

Sector chain_at(const Category &cat, const FusionTree &tree, int64_t last) {
  if (last < 0) return cat.trivial();
  if (last + 1 == tree.size()) return tree.coupled;
  if (last == 0) return tree.sectors.at(0);
  return tree.inner.at(static_cast<size_t>(last - 1));
}

// The recoupling of an exchange of codomain legs p, p + 1 at one root: for every source tree, the
// target trees it lands on with their coefficients, `braided . R . straight^dagger` restricted to
// trees that agree outside the pair (regroup.cpp:46-185).
std::vector<std::vector<std::pair<int64_t, cplx>>> exchange_factors(const CategoryPtr &cat, const Legs &codomain,
                                                                      const Legs &exchanged, const Sector &root,
                                                                      int64_t p) {
  const LayoutPtr source = layout_of(cat, codomain, root);
  const LayoutPtr target = layout_of(cat, exchanged, root);
  const int64_t count = static_cast<int64_t>(codomain.size());
  auto signature = [&](const FusionTree &tree, bool exchanged_tree) {
    std::vector<int64_t> out;
    for (int64_t axis = 0; axis < count; ++axis) {
      const int64_t from = exchanged_tree && (axis == p || axis == p + 1) ? (axis == p ? p + 1 : p) : axis;
      const Sector &s = tree.sectors[static_cast<size_t>(from)];
      out.insert(out.end(), s.v.begin(), s.v.begin() + s.n);
    }
    for (int64_t at = 0; at < static_cast<int64_t>(tree.inner.size()); ++at) {
      if (at == p - 1) continue;
      const Sector &s = tree.inner[static_cast<size_t>(at)];
      out.insert(out.end(), s.v.begin(), s.v.begin() + s.n);
    }
    for (int64_t at = 0; at < static_cast<int64_t>(tree.copies.size()); ++at) {
      if (at == p - 1 || at == p) continue;
      out.push_back(tree.copies[static_cast<size_t>(at)]);
    }
    return out;
  };
  std::map<std::vector<int64_t>, std::vector<int64_t>> groups, partners;
  for (int64_t at = 0; at < static_cast<int64_t>(source->trees.size()); ++at)
    groups[signature(source->trees[static_cast<size_t>(at)], false)].push_back(at);
  for (int64_t at = 0; at < static_cast<int64_t>(target->trees.size()); ++at)
    partners[signature(target->trees[static_cast<size_t>(at)], true)].push_back(at);
  std::vector<std::vector<std::pair<int64_t, cplx>>> out(source->trees.size());
  for (const auto &[marks, olds] : groups) {
    const auto found = partners.find(marks);
    if (found == partners.end()) throw std::logic_error("the exchange has no image for a tree of the codomain");
    const std::vector<int64_t> &news = found->second;
    const FusionTree &sample = source->trees[static_cast<size_t>(olds.front())];
    const Sector y = chain_at(*cat, sample, p - 1);
    const Sector d = chain_at(*cat, sample, p + 1);
    const Sector &b = sample.sectors[static_cast<size_t>(p)];
    const Sector &c = sample.sectors[static_cast<size_t>(p + 1)];
    const Matrix &straight = cat->F(y, b, c, d);
    const Matrix &braided = cat->F(y, c, b, d);
    if (straight.rows != static_cast<int64_t>(olds.size()) || braided.rows != static_cast<int64_t>(news.size()))
      throw std::logic_error("the exchange does not match the trees of the block");
    std::vector<std::pair<Sector, int64_t>> offset;
    int64_t at = 0;
    for (const Outcome &o : cat->outcomes(b, c)) {
      offset.emplace_back(o.sector, at);
      at += o.multiplicity;
    }
    const std::vector<FIndex> from = cat->f_columns(y, b, c, d);
    const std::vector<FIndex> to = cat->f_columns(y, c, b, d);
    const Matrix &braid = cat->R(b, c);
    auto braid_at = [&](const FIndex &index) -> int64_t {
      for (const auto &[sector, start] : offset)
        if (sector == index.sector) return start + index.inner;
      return -1;
    };
    struct Term {
      int64_t i, br, bc, j;
    };
    std::vector<Term> terms;
    for (int64_t i = 0; i < static_cast<int64_t>(to.size()); ++i) {
      const int64_t br = braid_at(to[static_cast<size_t>(i)]);
      if (br < 0) continue;
      for (int64_t j = 0; j < static_cast<int64_t>(from.size()); ++j) {
        const int64_t bc = braid_at(from[static_cast<size_t>(j)]);
        if (bc < 0) continue;
        if (to[static_cast<size_t>(i)].sector != from[static_cast<size_t>(j)].sector) continue;
        if (to[static_cast<size_t>(i)].outer != from[static_cast<size_t>(j)].outer) continue;
        terms.push_back({i, br, bc, j});
      }
    }
    for (size_t one = 0; one < news.size(); ++one)
      for (size_t other = 0; other < olds.size(); ++other) {
        cplx value(0.0, 0.0);
        for (const Term &t : terms)
          value += braided(static_cast<int64_t>(one), t.i) * braid(t.br, t.bc) *
                   std::conj(straight(static_cast<int64_t>(other), t.j));
        if (value.real() == 0.0 && value.imag() == 0.0) continue;
        out[static_cast<size_t>(olds[other])].emplace_back(news[one], value);
      }
  }
  for (auto &list : out) std::sort(list.begin(), list.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
  return out;
}

// One elementary step over tree pairs: for each pair of `from`, the pairs of `to` it lands on.
std::vector<Terms> step_map(const PairSpace &from, const PairSpace &to, const Step &step) {
  const Structure &S = *from.s;
  const Structure &T = *to.s;
  const CategoryPtr &cat = S.cat();
  const Sector &q = S.qtotal();
  std::vector<Terms> out(static_cast<size_t>(from.total));
  if (step.kind == StepKind::exchange) {
    const int64_t p = step.position;
    std::map<Sector, std::vector<std::vector<std::pair<int64_t, cplx>>>> by_root;
    for (int64_t b = 0; b < static_cast<int64_t>(S.blocks().size()); ++b) {
      const BlockInfo &blk = S.blocks()[static_cast<size_t>(b)];
      auto found = by_root.find(blk.key.codomain);
      if (found == by_root.end())
        found = by_root.emplace(blk.key.codomain, exchange_factors(cat, S.codomain(), T.codomain(), blk.key.codomain, p)).first;
      const int64_t tb = T.find(blk.key);
      if (tb < 0) throw std::logic_error("an exchange lost a block key");
      const int64_t nrows = static_cast<int64_t>(blk.row_layout->trees.size());
      const int64_t ncols = from.ncols[static_cast<size_t>(b)];
      for (int64_t row = 0; row < nrows; ++row)
        for (const auto &[nrow, coef] : found->second[static_cast<size_t>(row)])
          for (int64_t col = 0; col < ncols; ++col)
            out[static_cast<size_t>(from.id(b, row, col))].emplace_back(to.id(tb, nrow, col), coef);
    }
    return out;
  }
  // A bend: the last leg of the giving side onto the end of the taking side (regroup.cpp:260-417).
  const bool from_codomain = step.kind == StepKind::to_domain;
  const LegPtr moved_leg = from_codomain ? S.codomain().back() : S.domain().back();
  for (int64_t b = 0; b < static_cast<int64_t>(S.blocks().size()); ++b) {
    const BlockInfo &blk = S.blocks()[static_cast<size_t>(b)];
    const SideLayout &moved_side = from_codomain ? *blk.row_layout : *blk.col_layout;
    const SideLayout &gains_side = from_codomain ? *blk.col_layout : *blk.row_layout;
    const Sector &moved_root = from_codomain ? blk.key.codomain : blk.key.domain;
    const Sector &gains_root = from_codomain ? blk.key.domain : blk.key.codomain;
    for (int64_t mi = 0; mi < static_cast<int64_t>(moved_side.trees.size()); ++mi) {
      const FusionTree &moved_tree = moved_side.trees[static_cast<size_t>(mi)];
      const Sector charge = moved_tree.sectors.back();
      (void)moved_leg;
      const FusionTree kept_tree = trimmed(*cat, moved_tree);
      const Sector prefix = kept_tree.coupled;
      const int64_t copy = moved_tree.size() >= 2 ? moved_tree.copies.back() : 0;
      const Sector first = from_codomain ? prefix : gains_root;
      const Sector middle = from_codomain ? charge : cat->dual(charge);
      const Sector last = from_codomain ? cat->dual(gains_root) : cat->dual(prefix);
      const Matrix &coefficients = cat->F(first, middle, last, q);
      const std::vector<FIndex> across = cat->f_rows(first, middle, last, q);
      const std::vector<FIndex> along = cat->f_columns(first, middle, last, q);
      const std::vector<FIndex> &named = from_codomain ? across : along;
      const std::vector<FIndex> &free = from_codomain ? along : across;
      const Sector named_root = from_codomain ? moved_root : cat->dual(moved_root);
      std::vector<int64_t> taken;
      for (size_t at = 0; at < named.size(); ++at)
        if (named[at].sector == named_root && named[at].inner == copy && named[at].outer == blk.key.copy)
          taken.push_back(static_cast<int64_t>(at));
      if (taken.empty()) continue;
      for (int64_t gi = 0; gi < static_cast<int64_t>(gains_side.trees.size()); ++gi) {
        const FusionTree &gains_tree = gains_side.trees[static_cast<size_t>(gi)];
        const int64_t source_pair = from_codomain ? from.id(b, mi, gi) : from.id(b, gi, mi);
        for (int64_t index : taken)
          for (size_t column = 0; column < free.size(); ++column) {
            const FIndex &taken_side = free[column];
            // The reference reads entry [index * width + column] in both directions (regroup.cpp:346).
            const cplx value = coefficients(index, static_cast<int64_t>(column));
            if (value.real() == 0.0 && value.imag() == 0.0) continue;
            const Sector extended_root = from_codomain ? cat->dual(taken_side.sector) : taken_side.sector;
            FusionTree extended = gains_tree;
            extended.sectors.push_back(cat->dual(charge));
            if (gains_tree.size() >= 2) extended.inner.push_back(gains_root);
            if (gains_tree.size() >= 1) extended.copies.push_back(taken_side.inner);
            extended.coupled = extended_root;
            const BlockKey key{from_codomain ? prefix : extended_root, from_codomain ? extended_root : prefix,
                               taken_side.outer};
            const int64_t tb = T.find(key);
            if (tb < 0) throw std::logic_error("a bend lands outside the blocks of the result");
            const BlockInfo &tblk = T.blocks()[static_cast<size_t>(tb)];
            const int64_t row = tblk.row_layout->find(from_codomain ? kept_tree : extended);
            const int64_t col = tblk.col_layout->find(from_codomain ? extended : kept_tree);
            if (row < 0 || col < 0) continue;  // regroup.cpp:383
            out[static_cast<size_t>(source_pair)].emplace_back(to.id(tb, row, col), value);
          }
      }
    }
  }
  return out;
}

// Multiplicities of the legs at a tree pair, codomain legs then domain legs.
void pair_extents(const Structure &S, int64_t block, int64_t row, int64_t col, std::vector<int64_t> &ext) {
  const BlockInfo &b = S.blocks()[static_cast<size_t>(block)];
  const SideLayout &rl = *b.row_layout, &cl = *b.col_layout;
  ext.assign(rl.mults.begin() + row * rl.nlegs, rl.mults.begin() + (row + 1) * rl.nlegs);
  ext.insert(ext.end(), cl.mults.begin() + col * cl.nlegs, cl.mults.begin() + (col + 1) * cl.nlegs);
}

// Element offset and per-axis strides of one tree pair's box inside the flat buffer.
int64_t pair_strides(const Structure &S, int64_t block, int64_t row, int64_t col, const std::vector<int64_t> &ext,
                     std::vector<int64_t> &stride) {
  const BlockInfo &b = S.blocks()[static_cast<size_t>(block)];
  const int64_t ncod = static_cast<int64_t>(S.codomain().size());
  const int64_t nall = static_cast<int64_t>(ext.size());
  stride.assign(static_cast<size_t>(nall), 0);
  int64_t run = 1;
  for (int64_t a = nall - 1; a >= ncod; --a) {
    stride[static_cast<size_t>(a)] = run;
    run *= ext[static_cast<size_t>(a)];
  }
  run = b.cols;
  for (int64_t a = ncod - 1; a >= 0; --a) {
    stride[static_cast<size_t>(a)] = run;
    run *= ext[static_cast<size_t>(a)];
  }
  return b.offset + b.row_layout->offsets[static_cast<size_t>(row)] * b.cols + b.col_layout->offsets[static_cast<size_t>(col)];
}

// The leg permutation of a chain: target axis j carries source leg perm[j].
std::vector<int64_t> chain_permutation(int64_t ncod0, int64_t nlegs, const std::vector<Step> &steps) {
  std::vector<int64_t> perm(static_cast<size_t>(nlegs));
  std::iota(perm.begin(), perm.end(), 0);
  int64_t ncod = ncod0;
  for (const Step &step : steps) {
    if (step.kind == StepKind::exchange) {
      std::swap(perm[static_cast<size_t>(step.position)], perm[static_cast<size_t>(step.position + 1)]);
    } else if (step.kind == StepKind::to_domain) {
      const int64_t leg = perm[static_cast<size_t>(ncod - 1)];
      perm.erase(perm.begin() + (ncod - 1));
      perm.push_back(leg);
      --ncod;
    } else {
      const int64_t leg = perm.back();
      perm.pop_back();
      perm.insert(perm.begin() + ncod, leg);
      ++ncod;
    }
  }
  return perm;
}

// For abelian categories every vertex is the number 1, so each elementary step maps one tree pair to
// exactly one tree pair: an exchange swaps two sectors with the factor R(b, c) (a sign for fermion
// parity, 1 for the groups), a bend moves the last sector of one side to the end of the other as its
// dual with F = 1 (regroup.cpp:25-417 reduce to this).  The composed map is therefore the chain
// replayed on the sector lists of each source pair; only the final structure is built.  Where every R
// is 1 the replay collapses further: target axis j carries the sector of source leg perm[j], taken
// dual when that leg ends on the other side (the dual is an involution, so two bends cancel).
// For each source pair: the target pair it lands on (-1 when a dual sector is not on its leg) and the
// coefficient.
void compose_abelian(const StructurePtr &source, const std::vector<Step> &steps, const std::vector<int64_t> &perm,
                     const PairSpace &target, std::vector<int64_t> &to, std::vector<cplx> &coefs) {
  const Category &cat = *source->cat();
  const Structure &T = *target.s;
  const int64_t ncod0 = static_cast<int64_t>(source->codomain().size());
  const int64_t ncod1 = static_cast<int64_t>(T.codomain().size());
  const int64_t nlegs = source->nlegs();
  const bool replay = !cat.trivial_braiding();
  std::vector<char> flips(static_cast<size_t>(nlegs));
  for (int64_t j = 0; j < nlegs; ++j) flips[static_cast<size_t>(j)] = (perm[static_cast<size_t>(j)] < ncod0) != (j < ncod1);
  PairSpace space(source);
  to.assign(static_cast<size_t>(space.total), -1);
  coefs.assign(static_cast<size_t>(space.total), cplx(1.0, 0.0));
  auto root_of = [&](const std::vector<Sector> &sectors) {
    Sector r = cat.trivial();
    for (const Sector &s : sectors) r = cat.outcomes(r, s).front().sector;
    return r;
  };
  std::vector<Sector> cod, dom;
  for (int64_t b = 0; b < static_cast<int64_t>(source->blocks().size()); ++b) {
    const BlockInfo &blk = source->blocks()[static_cast<size_t>(b)];
    for (int64_t row = 0; row < static_cast<int64_t>(blk.row_layout->trees.size()); ++row)
      for (int64_t col = 0; col < static_cast<int64_t>(blk.col_layout->trees.size()); ++col) {
        const std::vector<Sector> &rs = blk.row_layout->trees[static_cast<size_t>(row)].sectors;
        const std::vector<Sector> &cs = blk.col_layout->trees[static_cast<size_t>(col)].sectors;
        cplx coef(1.0, 0.0);
        if (replay) {
          cod = rs;
          dom = cs;
          for (const Step &step : steps) {
            if (step.kind == StepKind::exchange) {
              const size_t p = static_cast<size_t>(step.position);
              coef *= cat.R(cod[p], cod[p + 1])(0, 0);
              std::swap(cod[p], cod[p + 1]);
            } else if (step.kind == StepKind::to_domain) {
              dom.push_back(cat.dual(cod.back()));
              cod.pop_back();
            } else {
              cod.push_back(cat.dual(dom.back()));
              dom.pop_back();
            }
          }
        } else {
          cod.clear();
          dom.clear();
          for (int64_t j = 0; j < nlegs; ++j) {
            const int64_t leg = perm[static_cast<size_t>(j)];
            const Sector &s = leg < ncod0 ? rs[static_cast<size_t>(leg)] : cs[static_cast<size_t>(leg - ncod0)];
            (j < ncod1 ? cod : dom).push_back(flips[static_cast<size_t>(j)] ? cat.dual(s) : s);
          }
        }
        const int64_t tb = T.find(root_of(cod), root_of(dom), blk.key.copy);
        if (tb < 0) continue;
        const BlockInfo &tblk = T.blocks()[static_cast<size_t>(tb)];
        const int64_t trow = tblk.row_layout->find_sectors(cod);
        const int64_t tcol = tblk.col_layout->find_sectors(dom);
        if (trow < 0 || tcol < 0) continue;  // the dual sector is not on the leg (regroup.cpp:383)
        const size_t at = static_cast<size_t>(space.id(b, row, col));
        to[at] = target.id(tb, trow, tcol);
        coefs[at] = coef;
      }
  }
}

// Compose the steps from `source` and emit one tile per target tree pair.
void compose(const StructurePtr &source, const std::vector<Step> &steps, const std::vector<StructurePtr> &after,
             MovePlan &plan, bool per_block_groups) {
  const int64_t nlegs = source->nlegs();
  if (nlegs > TT_MAX_DIMS) throw std::invalid_argument("a leg move handles at most " + std::to_string(TT_MAX_DIMS) + " legs");
  const std::vector<int64_t> perm = chain_permutation(static_cast<int64_t>(source->codomain().size()), nlegs, steps);
  const StructurePtr &target = after.back();
  const PairSpace source_space(source), target_space(target);
  // Target-major contributions, each in source-pair order: the sources of target pair t are
  // from_pair/from_coef[start[t], start[t + 1]).
  std::vector<int64_t> start(static_cast<size_t>(target_space.total + 1), 0), from_pair;
  std::vector<cplx> from_coef;
  auto gather = [&](auto &&for_each_term) {
    for_each_term([&](int64_t, int64_t t, const cplx &) { ++start[static_cast<size_t>(t + 1)]; });
    for (size_t t = 1; t < start.size(); ++t) start[t] += start[t - 1];
    from_pair.resize(static_cast<size_t>(start.back()));
    from_coef.resize(static_cast<size_t>(start.back()));
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for_each_term([&](int64_t i, int64_t t, const cplx &c) {
      const size_t at = static_cast<size_t>(fill[static_cast<size_t>(t)]++);
      from_pair[at] = i;
      from_coef[at] = c;
    });
  };
  auto nonzero = [](const cplx &c) { return c.real() != 0.0 || c.imag() != 0.0; };
  if (source->cat()->abelian()) {
    std::vector<int64_t> to;
    std::vector<cplx> coefs;
    compose_abelian(source, steps, perm, target_space, to, coefs);
    gather([&](auto &&visit) {
      for (int64_t i = 0; i < source_space.total; ++i)
        if (to[static_cast<size_t>(i)] >= 0 && nonzero(coefs[static_cast<size_t>(i)]))
          visit(i, to[static_cast<size_t>(i)], coefs[static_cast<size_t>(i)]);
    });
  } else {
    PairSpace space(source);
    std::vector<Terms> comp(static_cast<size_t>(space.total));
    for (int64_t i = 0; i < space.total; ++i) comp[static_cast<size_t>(i)].emplace_back(i, cplx(1.0, 0.0));
    for (size_t k = 0; k < steps.size(); ++k) {
      PairSpace next(after[k]);
      const std::vector<Terms> map = step_map(space, next, steps[k]);
      for (Terms &terms : comp) {
        Terms out;
        for (const auto &[pair, coef] : terms)
          for (const auto &[target_pair, c2] : map[static_cast<size_t>(pair)]) {
            auto it = std::find_if(out.begin(), out.end(), [&](const auto &e) { return e.first == target_pair; });
            if (it == out.end())
              out.emplace_back(target_pair, coef * c2);
            else
              it->second += coef * c2;
          }
        terms = std::move(out);
      }
      space = std::move(next);
    }
    gather([&](auto &&visit) {
      for (int64_t i = 0; i < static_cast<int64_t>(comp.size()); ++i)
        for (const auto &[pair, coef] : comp[static_cast<size_t>(i)])
          if (nonzero(coef)) visit(i, pair, coef);
    });
  }
  // Buffers reused across tiles: per source, its strides mapped onto the target axes (nlegs each),
  // and after squeezing and merging, its strides per tile axis (nlegs each).
  std::vector<int64_t> text, tstride, sext, sstride, soff, sstr, ext, dstr, src;
  plan.tasks.clear();
  plan.sources.clear();
  plan.tasks.reserve(static_cast<size_t>(target_space.total));
  plan.sources.reserve(from_pair.size());
  const size_t L = static_cast<size_t>(nlegs);
  for (int64_t b = 0; b < static_cast<int64_t>(target->blocks().size()); ++b) {
    const BlockInfo &blk = target->blocks()[static_cast<size_t>(b)];
    const int64_t nrows = static_cast<int64_t>(blk.row_layout->trees.size());
    const int64_t ncols = static_cast<int64_t>(blk.col_layout->trees.size());
    for (int64_t row = 0; row < nrows; ++row)
      for (int64_t col = 0; col < ncols; ++col) {
        const int64_t pair = target_space.id(b, row, col);
        pair_extents(*target, b, row, col, text);
        const int64_t doff = pair_strides(*target, b, row, col, text, tstride);
        const size_t first = static_cast<size_t>(start[static_cast<size_t>(pair)]);
        const size_t count = static_cast<size_t>(start[static_cast<size_t>(pair + 1)]) - first;
        soff.resize(count);
        sstr.resize(count * L);
        src.resize(count * L);
        for (size_t s = 0; s < count; ++s) {
          int64_t sb, sr, sc;
          source_space.decode(from_pair[first + s], &sb, &sr, &sc);
          pair_extents(*source, sb, sr, sc, sext);
          soff[s] = pair_strides(*source, sb, sr, sc, sext, sstride);
          for (size_t j = 0; j < L; ++j) {
            const size_t from_axis = static_cast<size_t>(perm[j]);
            if (sext[from_axis] != text[j])
              throw std::logic_error("a leg moved across sides must carry the same multiplicity in a sector and its dual");
            sstr[s * L + j] = sstride[from_axis];
          }
        }
        // Squeeze unit axes and merge axes that stay contiguous for the target and every source.
        ext.clear();
        dstr.clear();
        for (size_t j = 0; j < L; ++j) {
          const int64_t e = text[j];
          if (e == 1) continue;
          bool merge = !ext.empty() && dstr.back() == tstride[j] * e;
          const size_t d = merge ? ext.size() - 1 : ext.size();
          for (size_t s = 0; merge && s < count; ++s) merge = src[s * L + d] == sstr[s * L + j] * e;
          if (merge) {
            ext.back() *= e;
            dstr.back() = tstride[j];
            for (size_t s = 0; s < count; ++s) src[s * L + d] = sstr[s * L + j];
          } else {
            const size_t fresh = ext.size();
            ext.push_back(e);
            dstr.push_back(tstride[j]);
            for (size_t s = 0; s < count; ++s) src[s * L + fresh] = sstr[s * L + j];
          }
        }
        tt_gss_task task{};
        task.dst_offset = doff;
        task.ndim = static_cast<int32_t>(ext.size());
        task.first_source = static_cast<int32_t>(plan.sources.size());
        task.num_sources = static_cast<int32_t>(count);
        task.group = per_block_groups ? static_cast<int32_t>(b) : 0;
        for (size_t d = 0; d < ext.size(); ++d) {
          task.extent[d] = ext[d];
          task.dst_stride[d] = dstr[d];
        }
        // Every address the tile touches stays inside its buffer (checked here, on the host, so a
        // planning mistake raises instead of reaching the device).
        auto span = [&](int64_t base, const int64_t *strides, int64_t size, const char *what) {
          int64_t hi = base;
          for (size_t d = 0; d < ext.size(); ++d) hi += (ext[d] - 1) * strides[d];
          if (base < 0 || hi >= size) throw std::logic_error(std::string("a leg move tile leaves its ") + what + " buffer");
        };
        span(doff, dstr.data(), target->size(), "target");
        for (size_t s = 0; s < count; ++s) span(soff[s], src.data() + s * L, source->size(), "source");
        plan.tasks.push_back(task);
        for (size_t s = 0; s < count; ++s) {
          tt_gss_source source_box{};
          source_box.src_offset = soff[s];
          for (size_t d = 0; d < ext.size(); ++d) source_box.src_stride[d] = src[s * L + d];
          source_box.coef_re = from_coef[first + s].real();
          source_box.coef_im = from_coef[first + s].imag();
          plan.sources.push_back(source_box);
        }
      }
  }
}

struct MoveKey {
  uint64_t structure;
  std::vector<int64_t> above, below;
  bool plain;
  bool operator<(const MoveKey &o) const {
    return std::tie(structure, above, below, plain) < std::tie(o.structure, o.above, o.below, o.plain);
  }
};

std::map<MoveKey, MovePlanPtr> &move_cache() {
  static std::map<MoveKey, MovePlanPtr> cache;
  return cache;
}

MovePlanPtr build(const StructurePtr &source, const std::vector<int64_t> &above, const std::vector<int64_t> &below,
                  bool plain) {
  MoveKey key{source->id(), above, below, plain};
  auto &cache = move_cache();
  auto found = cache.find(key);
  if (found != cache.end()) return found->second;
  if (cache.size() > 8192) cache.clear();
  const Chain chain = build_chain(source->codomain(), source->domain(), above, below);
  auto plan = std::make_shared<MovePlan>();
  plan->source = source;
  std::vector<StructurePtr> after;
  // An abelian chain composes without its intermediate structures; the per-step form needs them.
  const bool direct = source->cat()->abelian() && !plain;
  for (size_t k = direct && !chain.sides.empty() ? chain.sides.size() - 1 : 0; k < chain.sides.size(); ++k)
    after.push_back(Structure::make(source->cat(), chain.sides[k].first, chain.sides[k].second, source->qtotal()));
  plan->target = after.empty() ? source : after.back();
  plan->identity = chain.steps.empty();
  plan->above = above;
  plan->below = below;
  plan->chained = true;
  if (direct)
    plan->max_intermediate = std::max(source->size(), plan->target->size());
  else
    for (size_t k = 0; k + 1 < after.size(); ++k) plan->max_intermediate = std::max(plan->max_intermediate, after[k]->size());
  if (!plan->identity) {
    compose(source, chain.steps, after, *plan, false);
    if (plain) {
      StructurePtr from = source;
      for (size_t k = 0; k < chain.steps.size(); ++k) {
        auto one = std::make_shared<MovePlan>();
        one->source = from;
        one->target = after[k];
        compose(from, {chain.steps[k]}, {after[k]}, *one, true);
        plan->elementary.push_back(one);
        from = after[k];
      }
    }
  }
  cache.emplace(std::move(key), plan);
  return plan;
}

}  // namespace

MovePlan::~MovePlan() {
  if (device_) tt_gss_plan_destroy(device_);
}

tt_gss_plan *MovePlan::device() {
  if (!device_) {
    if (tt_gss_plan_create(tasks.data(), static_cast<int64_t>(tasks.size()), sources.data(),
                           static_cast<int64_t>(sources.size()), &device_) != 0)
      throw std::runtime_error("leg move upload: " + device_error());
  }
  return device_;
}

MovePlanPtr plan_move(const StructurePtr &source, const std::vector<int64_t> &above, const std::vector<int64_t> &below) {
  return build(source, above, below, false);
}

MovePlanPtr plan_move_plain(const StructurePtr &source, const std::vector<int64_t> &above,
                            const std::vector<int64_t> &below) {
  return build(source, above, below, true);
}

void clear_move_cache() { move_cache().clear(); }

std::string device_error() {
  char buffer[1024];
  tt_last_error(buffer, sizeof(buffer));
  return buffer;
}

}  // namespace tt
