#include "program.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <functional>
#include <limits>
#include <map>
#include <stdexcept>
#include <tuple>

namespace tt {

std::string device_error();

namespace {

constexpr int64_t kMaxCharges = 4096;  // einsum.cpp:25

int64_t element_size(int dtype) { return dtype == TT_F32 ? 4 : dtype == TT_C128 ? 16 : 8; }

void check(int status, const char *what) {
  if (status != 0) throw std::runtime_error(std::string(what) + ": " + device_error());
}

// contract.cpp:95-106
void require_single_closing(const Category &cat, const Sector &above, const Sector &below, const Sector &charge,
                            const std::string &subject = "the roots", const std::string &what = "a contraction") {
  if (cat.n_symbol(above, cat.dual(below), charge) > 1)
    throw std::invalid_argument(subject + " " + render(above) + " and " + render(below) + " close to " + render(charge) +
                                " in more than one way, which " + what + " does not recouple yet");
}

std::vector<int64_t> without(int64_t total, const std::vector<int64_t> &taken) {
  std::vector<bool> named(static_cast<size_t>(total), false);
  for (int64_t leg : taken) named[static_cast<size_t>(leg)] = true;
  std::vector<int64_t> out;
  for (int64_t leg = 0; leg < total; ++leg)
    if (!named[static_cast<size_t>(leg)]) out.push_back(leg);
  return out;
}

// contract.cpp:47-77
void check_pairs(const Structure &left, const Structure &right, const std::vector<int64_t> &left_legs,
                 const std::vector<int64_t> &right_legs) {
  if (left.cat() != right.cat()) throw std::invalid_argument("a contraction needs both tensors over one category");
  if (left_legs.size() != right_legs.size())
    throw std::invalid_argument("the two legs of a contraction must be given in pairs, got " +
                                std::to_string(left_legs.size()) + " and " + std::to_string(right_legs.size()));
  auto check_side = [](const Structure &t, const std::vector<int64_t> &legs, const char *side) {
    std::vector<bool> named(static_cast<size_t>(t.nlegs()), false);
    for (int64_t leg : legs) {
      if (leg < 0 || leg >= t.nlegs())
        throw std::invalid_argument("the leg " + std::to_string(leg) + " of the " + side +
                                    " tensor is out of range of its " + std::to_string(t.nlegs()) + " legs");
      if (named[static_cast<size_t>(leg)])
        throw std::invalid_argument("the leg " + std::to_string(leg) + " of the " + side + " tensor is summed twice");
      named[static_cast<size_t>(leg)] = true;
    }
  };
  check_side(left, left_legs, "first");
  check_side(right, right_legs, "second");
  const Legs la = left.legs(), ra = right.legs();
  for (size_t at = 0; at < left_legs.size(); ++at)
    if (la[static_cast<size_t>(left_legs[at])] != ra[static_cast<size_t>(right_legs[at])])
      throw std::invalid_argument("the pair at position " + std::to_string(at) + " joins the leg " +
                                  std::to_string(left_legs[at]) + " of the first tensor to the leg " +
                                  std::to_string(right_legs[at]) + " of the second, and the two do not name one space");
}

// The block products of two aligned operands (contract.cpp:128-163).
GemmPlanPtr plan_gemm(const StructurePtr &above, const StructurePtr &under, const Sector &charge) {
  const Category &cat = *above->cat();
  auto g = std::make_shared<GemmPlan>();
  g->above = above;
  g->under = under;
  g->out = Structure::make(above->cat(), above->codomain(), under->domain(), charge);
  for (const BlockInfo &b : above->blocks()) require_single_closing(cat, b.key.codomain, b.key.domain, above->qtotal());
  for (const BlockInfo &b : under->blocks()) require_single_closing(cat, b.key.codomain, b.key.domain, under->qtotal());
  for (const BlockInfo &b : g->out->blocks()) require_single_closing(cat, b.key.codomain, b.key.domain, charge);
  std::map<Sector, std::vector<int64_t>> rows;
  for (int64_t j = 0; j < static_cast<int64_t>(under->blocks().size()); ++j)
    rows[under->blocks()[static_cast<size_t>(j)].key.codomain].push_back(j);
  std::vector<std::vector<tt_gemm_term>> per_block(g->out->blocks().size());
  for (const BlockInfo &lb : above->blocks()) {
    const auto found = rows.find(lb.key.domain);
    if (found == rows.end()) continue;
    for (int64_t j : found->second) {
      const BlockInfo &rb = under->blocks()[static_cast<size_t>(j)];
      if (cat.n_symbol(lb.key.codomain, cat.dual(rb.key.domain), charge) == 0) continue;
      if (lb.cols != rb.rows)
        throw std::logic_error("the two blocks over the roots " + render(lb.key.domain) + " do not line up: " +
                               std::to_string(lb.cols) + " against " + std::to_string(rb.rows));
      const int64_t ob = g->out->find(lb.key.codomain, rb.key.domain, 0);
      if (ob < 0)
        throw std::out_of_range("the tensor has no block at " + render(lb.key.codomain) + " x " + render(rb.key.domain) +
                                " of copy 0");
      const BlockInfo &obk = g->out->blocks()[static_cast<size_t>(ob)];
      if (obk.rows != lb.rows || obk.cols != rb.cols) throw std::logic_error("a block product does not fit its result block");
      per_block[static_cast<size_t>(ob)].push_back(tt_gemm_term{lb.offset, rb.offset, lb.cols});
      g->flops += 2.0 * static_cast<double>(lb.rows) * static_cast<double>(rb.cols) * static_cast<double>(lb.cols);
    }
  }
  for (size_t ob = 0; ob < per_block.size(); ++ob) {
    const BlockInfo &b = g->out->blocks()[ob];
    tt_gemm_group grp{};
    grp.c_offset = b.offset;
    grp.m = b.rows;
    grp.n = b.cols;
    grp.first_term = static_cast<int32_t>(g->terms.size());
    grp.num_terms = static_cast<int32_t>(per_block[ob].size());
    g->terms.insert(g->terms.end(), per_block[ob].begin(), per_block[ob].end());
    g->groups.push_back(grp);
  }
  return g;
}

struct Builder {
  ProgramPtr p = std::make_shared<Program>();
  std::vector<uint32_t> deps;

  int32_t operand(int32_t index, const StructurePtr &s) {
    p->slots.push_back(Slot{Slot::operand, index, 0, s});
    deps.push_back(1u << (index < 32 ? index : 31));
    return static_cast<int32_t>(p->slots.size() - 1);
  }
  int32_t fresh(const StructurePtr &s, uint32_t dep) {
    p->slots.push_back(Slot{Slot::workspace, 0, 0, s});
    deps.push_back(dep);
    return static_cast<int32_t>(p->slots.size() - 1);
  }
  const StructurePtr &structure(int32_t v) const { return p->slots[static_cast<size_t>(v)].structure; }
  int32_t move(int32_t v, const std::vector<int64_t> &above, const std::vector<int64_t> &below) {
    MovePlanPtr plan = plan_move(structure(v), above, below);
    if (plan->identity) return v;
    const int32_t d = fresh(plan->target, deps[static_cast<size_t>(v)]);
    Op op{OpKind::move, plan, nullptr, v, -1, d, deps[static_cast<size_t>(v)]};
    p->ops.push_back(op);
    p->move_elements += plan->elements();
    return d;
  }
  // A leg move planned elsewhere (a partial trace): same op, same hoisting rules.
  int32_t apply(int32_t v, const MovePlanPtr &plan) {
    const int32_t d = fresh(plan->target, deps[static_cast<size_t>(v)]);
    p->ops.push_back(Op{OpKind::move, plan, nullptr, v, -1, d, deps[static_cast<size_t>(v)]});
    p->move_elements += plan->elements();
    return d;
  }
  int32_t gemm(int32_t a, int32_t b, const GemmPlanPtr &g) {
    const uint32_t dep = deps[static_cast<size_t>(a)] | deps[static_cast<size_t>(b)];
    const int32_t d = fresh(g->out, dep);
    p->ops.push_back(Op{OpKind::gemm, nullptr, g, a, b, d, dep});
    p->flops += g->flops;
    return d;
  }
  // Route the given result values into the output buffer: the first one is written there by its
  // producer (or copied when it is an operand), the rest are added.
  void finish(const std::vector<int32_t> &results) {
    const int32_t first = results.front();
    p->out = structure(first);
    Slot &slot = p->slots[static_cast<size_t>(first)];
    if (slot.kind == Slot::operand) {
      const int32_t d = fresh(p->out, deps[static_cast<size_t>(first)]);
      p->slots[static_cast<size_t>(d)].kind = Slot::output;
      p->ops.push_back(Op{OpKind::copy, nullptr, nullptr, first, -1, d, deps[static_cast<size_t>(first)]});
    } else {
      slot.kind = Slot::output;
    }
    int32_t out_value = -1;
    for (size_t i = 0; i < p->slots.size(); ++i)
      if (p->slots[i].kind == Slot::output) out_value = static_cast<int32_t>(i);
    for (size_t r = 1; r < results.size(); ++r)
      p->ops.push_back(Op{OpKind::add, nullptr, nullptr, results[r], -1, out_value, ~0u});
    // Workspace: every value that is neither an operand nor the output gets its own region.
    int64_t offset = 0;
    for (Slot &s : p->slots)
      if (s.kind == Slot::workspace) {
        s.offset = offset;
        offset += s.structure->size();
      }
    p->workspace = offset;
    int64_t widest = 0;
    for (const Op &op : p->ops)
      if (op.kind == OpKind::move) widest = std::max(widest, op.move->max_intermediate);
    p->plain_workspace = 2 * widest;
  }
};

struct Side {
  std::vector<char> labels;
  std::vector<bool> above;
};

Side side_of(const std::string &text) {
  Side out;
  for (char letter : text) {
    if (letter == ' ') continue;
    out.labels.push_back(static_cast<char>(std::tolower(static_cast<unsigned char>(letter))));
    out.above.push_back(letter >= 'A' && letter <= 'Z');
  }
  return out;
}

bool holds(const std::vector<char> &labels, char label) {
  return std::find(labels.begin(), labels.end(), label) != labels.end();
}

int64_t position_of(const std::vector<char> &labels, char label) {
  auto it = std::find(labels.begin(), labels.end(), label);
  if (it == labels.end()) throw std::logic_error(std::string("the label '") + label + "' left the equation");
  return it - labels.begin();
}

std::vector<char> union_of(const std::vector<char> &l, const std::vector<char> &r) {
  std::vector<char> out = l;
  for (char c : r)
    if (!holds(out, c)) out.push_back(c);
  return out;
}

std::vector<char> shared_of(const std::vector<char> &l, const std::vector<char> &r) {
  std::vector<char> out;
  for (char c : l)
    if (holds(r, c)) out.push_back(c);
  return out;
}

std::vector<char> product_labels(const std::vector<char> &l, const std::vector<char> &r) {
  const std::vector<char> shared = shared_of(l, r);
  std::vector<char> out;
  for (char c : union_of(l, r))
    if (!holds(shared, c)) out.push_back(c);
  return out;
}

double tensor_size(const std::vector<char> &labels, const std::map<char, double> &dims) {
  double out = 1;
  for (char c : labels) out *= dims.at(c);
  return out;
}

// einsum.cpp:216-301
std::vector<int64_t> fold_order(const std::vector<std::vector<char>> &inputs, const std::map<char, double> &dims) {
  const int64_t count = static_cast<int64_t>(inputs.size());
  std::vector<int64_t> order;
  if (count <= 12) {
    const int64_t masks = int64_t(1) << (count - 1);
    std::vector<std::vector<char>> labels(static_cast<size_t>(masks), inputs.front());
    for (int64_t mask = 1; mask < masks; ++mask) {
      const int64_t bit = __builtin_ctzll(static_cast<unsigned long long>(mask));
      labels[static_cast<size_t>(mask)] = union_of(labels[static_cast<size_t>(mask ^ (int64_t(1) << bit))], inputs[static_cast<size_t>(bit + 1)]);
    }
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<double> cost(static_cast<size_t>(masks), std::numeric_limits<double>::infinity());
      std::vector<int64_t> summed(static_cast<size_t>(masks), 0), last(static_cast<size_t>(masks), 0);
      cost[0] = 0;
      for (int64_t mask = 1; mask < masks; ++mask)
        for (int64_t bit = 0; bit < count - 1; ++bit)
          if (mask & (int64_t(1) << bit)) {
            const auto shared = shared_of(labels[static_cast<size_t>(mask ^ (int64_t(1) << bit))], inputs[static_cast<size_t>(bit + 1)]);
            if (pass == 0 && shared.empty()) continue;
            const double candidate = cost[static_cast<size_t>(mask ^ (int64_t(1) << bit))] + tensor_size(labels[static_cast<size_t>(mask)], dims);
            const size_t m = static_cast<size_t>(mask);
            const int64_t ns = static_cast<int64_t>(shared.size());
            const bool better = candidate < cost[m] || (candidate == cost[m] && ns > summed[m]) ||
                                (candidate == cost[m] && ns == summed[m] && bit > last[m]);
            if (better) {
              cost[m] = candidate;
              summed[m] = ns;
              last[m] = bit;
            }
          }
      if (cost[static_cast<size_t>(masks - 1)] == std::numeric_limits<double>::infinity()) continue;
      for (int64_t mask = masks - 1; mask; mask ^= int64_t(1) << last[static_cast<size_t>(mask)])
        order.push_back(last[static_cast<size_t>(mask)] + 1);
      order.push_back(0);
      std::reverse(order.begin(), order.end());
      return order;
    }
  }
  std::vector<bool> taken(inputs.size(), false);
  std::vector<char> live = inputs.front();
  order.push_back(0);
  taken[0] = true;
  for (int64_t step = 1; step < count; ++step) {
    int64_t best = -1;
    double best_size = 0;
    bool found = false;
    for (int pass = 0; pass < 2 && !found; ++pass)
      for (int64_t at = 0; at < count; ++at) {
        if (taken[static_cast<size_t>(at)]) continue;
        if (shared_of(live, inputs[static_cast<size_t>(at)]).empty() != (pass == 1)) continue;
        const double size = tensor_size(union_of(live, inputs[static_cast<size_t>(at)]), dims);
        if (!found || size < best_size) {
          found = true;
          best = at;
          best_size = size;
        }
      }
    order.push_back(best);
    taken[static_cast<size_t>(best)] = true;
    live = union_of(live, inputs[static_cast<size_t>(best)]);
  }
  return order;
}

double leg_size(const Category &cat, const Leg &leg) {
  double out = 0;
  for (int64_t b = 0; b < leg.nblocks(); ++b) out += static_cast<double>(leg.multiplicities()[static_cast<size_t>(b)]) * cat.q_dim(leg.sectors()[static_cast<size_t>(b)]);
  return out;
}

// align + contract of two values, one result per closing charge (einsum.cpp:306-345).
std::vector<int32_t> combine(Builder &bld, const std::vector<int32_t> &lefts, const std::vector<char> &left_labels,
                             const std::vector<int32_t> &rights, const std::vector<char> &right_labels) {
  const auto shared = shared_of(left_labels, right_labels);
  std::vector<int64_t> left_legs, right_legs;
  for (char c : shared) {
    left_legs.push_back(position_of(left_labels, c));
    right_legs.push_back(position_of(right_labels, c));
  }
  std::vector<int32_t> folded;
  for (int32_t l : lefts)
    for (int32_t r : rights) {
      const StructurePtr ls = bld.structure(l), rs = bld.structure(r);
      check_pairs(*ls, *rs, left_legs, right_legs);
      const int32_t above = bld.move(l, without(ls->nlegs(), left_legs), left_legs);
      const int32_t under = bld.move(r, right_legs, without(rs->nlegs(), right_legs));
      const Category &cat = *ls->cat();
      std::vector<Sector> closings;
      for (const Outcome &o : cat.outcomes(ls->qtotal(), rs->qtotal()))
        if (std::find(closings.begin(), closings.end(), o.sector) == closings.end()) closings.push_back(o.sector);
      for (const Sector &closed : closings) {
        require_single_closing(cat, ls->qtotal(), rs->qtotal(), closed, "the charges", "a contraction");
        folded.push_back(bld.gemm(above, under, plan_gemm(bld.structure(above), bld.structure(under), closed)));
      }
    }
  return folded;
}

// The partial trace of a tensor over its last codomain leg and its last domain leg, which name one
// space (einsum.cpp:95-159), as gather-scale-scatter tables: one per charge the kept roots close to.
// A source tree pair whose two traced sectors agree feeds the tile of its two trimmed trees; the
// traced copy digit is the fastest radix of both trees, so the sum over the copies is `copies`
// source boxes, each the diagonal offset (c, c) with stride `copies` along both axes.  Tiles nothing
// feeds are tasks with no source, which the kernel writes with zeros.
std::vector<MovePlanPtr> plan_trace(const StructurePtr &source) {
  const CategoryPtr &catp = source->cat();
  const Category &cat = *catp;
  const Legs out_codomain(source->codomain().begin(), source->codomain().end() - 1);
  const Legs out_domain(source->domain().begin(), source->domain().end() - 1);
  struct Box {
    int64_t offset, row_stride, col_stride;
  };
  // per closing charge: its structure and, per target tile (block, row tree, column tree), the boxes
  std::map<Sector, StructurePtr> targets;
  std::map<Sector, std::map<std::tuple<int64_t, int64_t, int64_t>, std::vector<Box>>> feeds;
  for (const BlockInfo &b : source->blocks()) {
    require_single_closing(cat, b.key.codomain, b.key.domain, source->qtotal(), "the roots", "a trace");
    const SideLayout &rows = *b.row_layout, &cols = *b.col_layout;
    for (size_t ra = 0; ra < rows.trees.size(); ++ra) {
      const FusionTree &above_tree = rows.trees[ra];
      const int64_t row_span = rows.offsets[ra + 1] - rows.offsets[ra];
      const FusionTree kept_above = trimmed(cat, above_tree);
      for (size_t ca = 0; ca < cols.trees.size(); ++ca) {
        const FusionTree &under_tree = cols.trees[ca];
        if (!(above_tree.sectors.back() == under_tree.sectors.back())) continue;
        const int64_t col_span = cols.offsets[ca + 1] - cols.offsets[ca];
        const FusionTree kept_under = trimmed(cat, under_tree);
        const LayoutPtr out_rows = layout_of(catp, out_codomain, kept_above.coupled);
        const LayoutPtr out_cols = layout_of(catp, out_domain, kept_under.coupled);
        const int64_t new_row = out_rows->find(kept_above), new_col = out_cols->find(kept_under);
        if (new_row < 0 || new_col < 0)
          throw std::logic_error("a kept tree of a trace is not in the layout it was trimmed from");
        const int64_t inside_rows = out_rows->offsets[static_cast<size_t>(new_row) + 1] - out_rows->offsets[static_cast<size_t>(new_row)];
        const int64_t inside_cols = out_cols->offsets[static_cast<size_t>(new_col) + 1] - out_cols->offsets[static_cast<size_t>(new_col)];
        const int64_t copies = row_span / inside_rows;
        if (copies * inside_rows != row_span || copies * inside_cols != col_span)
          throw std::logic_error("the traced legs of a tree pair carry different multiplicities");
        for (const Outcome &closing : cat.outcomes(kept_above.coupled, cat.dual(kept_under.coupled))) {
          require_single_closing(cat, kept_above.coupled, kept_under.coupled, closing.sector, "the kept roots", "a trace");
          auto slot = targets.find(closing.sector);
          if (slot == targets.end())
            slot = targets.emplace(closing.sector, Structure::make(catp, out_codomain, out_domain, closing.sector)).first;
          const int64_t tb = slot->second->find(kept_above.coupled, kept_under.coupled, 0);
          if (tb < 0) throw std::logic_error("a trace has no result block for its kept roots");
          auto &boxes = feeds[closing.sector][{tb, new_row, new_col}];
          for (int64_t c = 0; c < copies; ++c)
            boxes.push_back(Box{b.offset + (rows.offsets[ra] + c) * b.cols + cols.offsets[ca] + c, copies * b.cols, copies});
        }
      }
    }
  }
  std::vector<MovePlanPtr> out;
  for (const auto &[charge, target] : targets) {
    auto plan = std::make_shared<MovePlan>();
    plan->source = source;
    plan->target = target;
    const auto &fed = feeds[charge];
    for (int64_t tb = 0; tb < static_cast<int64_t>(target->blocks().size()); ++tb) {
      const BlockInfo &t = target->blocks()[static_cast<size_t>(tb)];
      const SideLayout &rows = *t.row_layout, &cols = *t.col_layout;
      for (size_t ra = 0; ra < rows.trees.size(); ++ra)
        for (size_t ca = 0; ca < cols.trees.size(); ++ca) {
          const int64_t nr = rows.offsets[ra + 1] - rows.offsets[ra], nc = cols.offsets[ca + 1] - cols.offsets[ca];
          if (nr == 0 || nc == 0) continue;
          tt_gss_task task{};
          task.dst_offset = t.offset + rows.offsets[ra] * t.cols + cols.offsets[ca];
          task.ndim = 2;
          task.first_source = static_cast<int32_t>(plan->sources.size());
          task.extent[0] = nr;
          task.extent[1] = nc;
          task.dst_stride[0] = t.cols;
          task.dst_stride[1] = 1;
          const auto found = fed.find({tb, static_cast<int64_t>(ra), static_cast<int64_t>(ca)});
          if (found != fed.end())
            for (const Box &box : found->second) {
              // the box's last element lies inside the source
              if (box.offset < 0 || box.offset + (nr - 1) * box.row_stride + (nc - 1) * box.col_stride >= source->size())
                throw std::logic_error("a trace reads past its source");
              tt_gss_source s{};
              s.src_offset = box.offset;
              s.src_stride[0] = box.row_stride;
              s.src_stride[1] = box.col_stride;
              s.coef_re = 1.0;
              plan->sources.push_back(s);
            }
          task.num_sources = static_cast<int32_t>(plan->sources.size()) - task.first_source;
          plan->tasks.push_back(task);
        }
    }
    out.push_back(plan);
  }
  return out;
}

struct ProgramKey {
  int kind;
  std::string spec;
  std::vector<uint64_t> structures;
  std::vector<int64_t> a, b;
  bool has_charge;
  std::vector<int64_t> charge;
  bool operator<(const ProgramKey &o) const {
    return std::tie(kind, spec, structures, a, b, has_charge, charge) <
           std::tie(o.kind, o.spec, o.structures, o.a, o.b, o.has_charge, o.charge);
  }
};

std::map<ProgramKey, ProgramPtr> &program_cache() {
  static std::map<ProgramKey, ProgramPtr> cache;
  return cache;
}

ProgramPtr cached(ProgramKey key, const std::function<ProgramPtr()> &make) {
  auto &cache = program_cache();
  auto found = cache.find(key);
  if (found != cache.end()) return found->second;
  if (cache.size() > 4096) cache.clear();
  ProgramPtr p = make();
  cache.emplace(std::move(key), p);
  return p;
}

ProgramPtr build_einsum(const std::string &spec, const std::vector<StructurePtr> &operands, const Sector *outcharge) {
  if (operands.empty()) throw std::invalid_argument("an einsum needs at least one operand");
  const size_t arrow = spec.find("->");
  if (arrow == std::string::npos) throw std::invalid_argument("the equation '" + spec + "' does not say what it leaves");
  std::vector<Side> groups;
  const std::string left = spec.substr(0, arrow);
  size_t start = 0;
  while (true) {
    const size_t comma = left.find(',', start);
    groups.push_back(side_of(left.substr(start, comma == std::string::npos ? std::string::npos : comma - start)));
    if (comma == std::string::npos) break;
    start = comma + 1;
  }
  const Side output = side_of(spec.substr(arrow + 2));
  if (groups.size() != operands.size())
    throw std::invalid_argument("the equation '" + spec + "' names " + std::to_string(groups.size()) +
                                " operands but got " + std::to_string(operands.size()));
  const CategoryPtr &catp = operands.front()->cat();
  const Category &cat = *catp;
  const Sector charge = outcharge ? *outcharge : cat.trivial();
  for (size_t at = 0; at < operands.size(); ++at) {
    if (operands[at]->cat() != catp) throw std::invalid_argument("an einsum needs every operand over one category");
    if (groups[at].labels.size() != static_cast<size_t>(operands[at]->nlegs()))
      throw std::invalid_argument("the group " + std::to_string(at) + " of '" + spec + "' names " +
                                  std::to_string(groups[at].labels.size()) + " legs but the tensor has " +
                                  std::to_string(operands[at]->nlegs()));
  }
  std::vector<char> seen;
  for (const Side &g : groups)
    for (char c : g.labels) {
      if (std::count(seen.begin(), seen.end(), c) >= 2)
        throw std::invalid_argument(std::string("the label '") + c + "' of '" + spec + "' names more than two legs");
      seen.push_back(c);
    }
  for (char c : seen) {
    const auto count = std::count(seen.begin(), seen.end(), c);
    if (count == 1 && !holds(output.labels, c))
      throw std::invalid_argument(std::string("the label '") + c + "' of '" + spec +
                                  "' names one leg but is not kept in the output, which would leave it open");
    if (count == 2 && holds(output.labels, c))
      throw std::invalid_argument(std::string("the label '") + c + "' of '" + spec + "' is summed over and still kept");
  }
  for (char c : output.labels)
    if (std::count(seen.begin(), seen.end(), c) != 1)
      throw std::invalid_argument(std::string("the output label '") + c + "' of '" + spec +
                                  "' does not name a leg of the operands");
  std::vector<char> summed;
  for (char c : seen)
    if (std::count(seen.begin(), seen.end(), c) == 2 && !holds(summed, c)) summed.push_back(c);

  Builder bld;
  std::vector<std::vector<int32_t>> prepared;
  std::vector<std::vector<char>> labels;
  for (size_t at = 0; at < operands.size(); ++at) {
    const int32_t v = bld.operand(static_cast<int32_t>(at), operands[at]);
    const int64_t ncod = static_cast<int64_t>(operands[at]->codomain().size());
    std::vector<int64_t> above, below;
    for (int64_t leg = 0; leg < operands[at]->nlegs(); ++leg) {
      const char label = groups[at].labels[static_cast<size_t>(leg)];
      const bool stays = holds(summed, label) ? leg < ncod : groups[at].above[static_cast<size_t>(leg)];
      (stays ? above : below).push_back(leg);
    }
    std::vector<char> lab;
    for (int64_t leg : above) lab.push_back(groups[at].labels[static_cast<size_t>(leg)]);
    for (int64_t leg : below) lab.push_back(groups[at].labels[static_cast<size_t>(leg)]);
    // Labels named twice within the operand are traced away, one pair at a time, the first repeated
    // label first (einsum.cpp:386-439); a trace can leave one tensor per charge its kept roots close
    // to, and every later step runs on each of them.
    std::vector<int32_t> alternatives{bld.move(v, above, below)};
    while (true) {
      int64_t first = -1, second = -1;
      for (size_t one = 0; one < lab.size() && first < 0; ++one)
        for (size_t other = one + 1; other < lab.size(); ++other)
          if (lab[one] == lab[other]) {
            first = static_cast<int64_t>(one);
            second = static_cast<int64_t>(other);
            break;
          }
      if (first < 0) break;
      const StructurePtr &shape = bld.structure(alternatives.front());
      const Legs legs = shape->legs();
      if (legs[static_cast<size_t>(first)] != legs[static_cast<size_t>(second)])
        throw std::invalid_argument(std::string("the label '") + lab[static_cast<size_t>(first)] + "' of '" + spec +
                                    "' names two different spaces in the operand " + std::to_string(at) +
                                    ", and a trace sums a space with itself");
      const int64_t ncodomain = static_cast<int64_t>(shape->codomain().size());
      const int64_t upper = first < ncodomain ? first : second, lower = first < ncodomain ? second : first;
      std::vector<int64_t> keep_above, keep_below;
      for (int64_t leg = 0; leg < shape->nlegs(); ++leg)
        if (leg != upper && leg != lower) (leg < ncodomain ? keep_above : keep_below).push_back(leg);
      keep_above.push_back(upper);
      keep_below.push_back(lower);
      std::vector<int32_t> traced;
      for (int32_t alternative : alternatives) {
        const int32_t placed = bld.move(alternative, keep_above, keep_below);
        for (const MovePlanPtr &plan : plan_trace(bld.structure(placed))) traced.push_back(bld.apply(placed, plan));
      }
      alternatives = std::move(traced);
      lab.erase(lab.begin() + second);
      lab.erase(lab.begin() + first);
    }
    if (alternatives.empty())
      throw std::invalid_argument("the traces of '" + spec + "' leave nothing to sum");
    prepared.push_back(std::move(alternatives));
    labels.push_back(lab);
  }
  size_t combinations = 1;
  for (const auto &alternatives : prepared) {
    if (static_cast<int64_t>(alternatives.size()) > kMaxCharges / static_cast<int64_t>(combinations))
      throw std::invalid_argument("the traces of '" + spec + "' leave more than " + std::to_string(kMaxCharges) +
                                  " combinations of charges, which is more than this sums");
    combinations *= alternatives.size();
  }
  // The output legs have to close to the charge (einsum.cpp:523-544).
  Legs above_legs, below_legs;
  for (size_t one = 0; one < output.labels.size(); ++one)
    for (size_t at = 0; at < labels.size(); ++at)
      if (holds(labels[at], output.labels[one])) {
        const LegPtr leg = bld.structure(prepared[at].front())->legs()[static_cast<size_t>(position_of(labels[at], output.labels[one]))];
        (output.above[one] ? above_legs : below_legs).push_back(leg);
        break;
      }
  bool closes = output.labels.empty();
  for (const Sector &a : reachable_charges(cat, above_legs))
    for (const Sector &b : reachable_charges(cat, below_legs))
      if (cat.n_symbol(a, cat.dual(b), charge) > 0) closes = true;
  if (!closes) throw std::invalid_argument("the output legs of '" + spec + "' cannot close to " + render(charge));
  std::map<char, double> dims;
  for (size_t at = 0; at < labels.size(); ++at) {
    const Legs legs = bld.structure(prepared[at].front())->legs();
    for (size_t leg = 0; leg < labels[at].size(); ++leg) dims[labels[at][leg]] = leg_size(cat, *legs[leg]);
  }
  const std::vector<int64_t> order = fold_order(labels, dims);
  std::vector<char> live = labels[static_cast<size_t>(order[0])];
  std::vector<int32_t> states = prepared[static_cast<size_t>(order[0])];
  for (size_t step = 1; step < order.size(); ++step) {
    const size_t at = static_cast<size_t>(order[step]);
    std::vector<int32_t> folded = combine(bld, states, live, prepared[at], labels[at]);
    if (static_cast<int64_t>(folded.size()) > kMaxCharges)
      throw std::invalid_argument("the products of the equation leave more than " + std::to_string(kMaxCharges) +
                                  " charges to sum, which is more than this takes");
    live = product_labels(live, labels[at]);
    states = std::move(folded);
  }
  std::vector<int64_t> above, below;
  for (size_t at = 0; at < output.labels.size(); ++at)
    (output.above[at] ? above : below).push_back(position_of(live, output.labels[at]));
  std::vector<int32_t> results;
  for (int32_t s : states) {
    const int32_t placed = bld.move(s, above, below);
    const StructurePtr &st = bld.structure(placed);
    if (!(st->qtotal() == charge) || st->blocks().empty()) continue;
    results.push_back(placed);
  }
  if (results.empty())
    throw std::invalid_argument("no way for the charges of '" + spec + "' to fuse to " + render(charge));
  bld.p->operands = operands;
  bld.finish(results);
  return bld.p;
}

}  // namespace

GemmPlan::~GemmPlan() {
  if (device_) tt_ggemm_plan_destroy(device_);
}

tt_ggemm_plan *GemmPlan::device() {
  if (!device_)
    check(tt_ggemm_plan_create_config(groups.data(), static_cast<int64_t>(groups.size()), terms.data(),
                                      static_cast<int64_t>(terms.size()), config_, &device_),
          "block product upload");
  return device_;
}

void GemmPlan::set_config(int config) {
  if (config < -1 || config >= tt_ggemm_num_configs())
    throw std::invalid_argument("no block-product configuration " + std::to_string(config));
  if (config == config_) return;
  if (device_) tt_ggemm_plan_destroy(device_);
  device_ = nullptr;
  config_ = config;
}

std::vector<float> GemmPlan::features(int element_size) const {
  // Plan-level inputs: sizes and counts on log2 scales, and flop-weighted histograms of log2 m, n
  // and k over six bins of width 2 (the last open).  Configuration inputs: the tile shape (one-hot
  // and BK), the packing, and what the shape does to this plan: tile count, padding, the longest
  // tile's k-slab count and the work per CU.
  auto lg = [](double v) { return static_cast<float>(std::log2(1.0 + v)); };
  double bytes = 0.0, mmax = 0, nmax = 0, kmax = 0;
  float hist[3][6] = {{0}};
  for (const tt_gemm_group &g : groups)
    for (int32_t t = g.first_term; t < g.first_term + g.num_terms; ++t) {
      const double m = static_cast<double>(g.m), n = static_cast<double>(g.n), k = static_cast<double>(terms[t].k);
      bytes += (m * k + k * n + m * n) * element_size;
      mmax = std::max(mmax, m), nmax = std::max(nmax, n), kmax = std::max(kmax, k);
      const double share = flops > 0 ? 2.0 * m * n * k / flops : 0.0;
      const double dims[3] = {m, n, k};
      for (int d = 0; d < 3; ++d) {
        const int bin = std::min(5, static_cast<int>(std::log2(std::max(1.0, dims[d])) / 2.0));
        hist[d][bin] += static_cast<float>(share);
      }
    }
  std::vector<float> base = {lg(static_cast<double>(groups.size())), lg(static_cast<double>(terms.size())), lg(flops),
                             lg(bytes), static_cast<float>(element_size) / 8.0f};
  for (int d = 0; d < 3; ++d) base.insert(base.end(), hist[d], hist[d] + 6);
  base.push_back(lg(mmax));
  base.push_back(lg(nmax));
  base.push_back(lg(kmax));
  const int configs = tt_ggemm_num_configs();
  int shapes = 0;
  for (int c = 0; c < configs; ++c) {
    int bm0 = 0, bn0 = 0, bk0 = 0, bm = 0, bn = 0, bk = 0;
    tt_ggemm_config_shape(0, &bm0, &bn0, &bk0);
    tt_ggemm_config_shape(c, &bm, &bn, &bk);
    if (c > 0 && bm == bm0 && bn == bn0 && bk == bk0) break;
    ++shapes;
  }
  std::vector<float> out;
  out.reserve(static_cast<size_t>(configs) * kGemmFeatures);
  for (int c = 0; c < configs; ++c) {
    int bm = 0, bn = 0, bk = 0;
    tt_ggemm_config_shape(c, &bm, &bn, &bk);
    double tiles = 0.0, padded = 0.0, area = 0.0, longest = 0.0, work = 0.0;
    for (const tt_gemm_group &g : groups) {
      int64_t ksum = 0;
      for (int32_t t = g.first_term; t < g.first_term + g.num_terms; ++t) ksum += terms[t].k;
      const double slabs = static_cast<double>((ksum + bk - 1) / bk) + 1.0;
      const double count = static_cast<double>(((g.m + bm - 1) / bm) * ((g.n + bn - 1) / bn));
      tiles += count;
      padded += count * bm * bn;
      area += static_cast<double>(g.m) * static_cast<double>(g.n);
      longest = std::max(longest, slabs);
      work += count * slabs;
    }
    std::vector<float> row = base;
    for (int k = 0; k < 7; ++k) row.push_back(c % shapes == k ? 1.0f : 0.0f);
    row.push_back(static_cast<float>(c / shapes));
    row.push_back(lg(tiles));
    row.push_back(area > 0 ? static_cast<float>(padded / area - 1.0) : 0.0f);
    row.push_back(lg(longest));
    row.push_back(lg(work / 256.0));
    row.push_back(lg(tiles / 256.0));
    row.push_back(static_cast<float>(bk) / 16.0f);
    if (row.size() != static_cast<size_t>(kGemmFeatures)) throw std::logic_error("cost-model feature count");
    out.insert(out.end(), row.begin(), row.end());
  }
  return out;
}

ProgramPtr plan_einsum(const std::string &spec, const std::vector<StructurePtr> &operands, const Sector *charge) {
  ProgramKey key{0, spec, {}, {}, {}, charge != nullptr, charge ? charge->to_vector() : std::vector<int64_t>{}};
  for (const auto &s : operands) key.structures.push_back(s->id());
  return cached(std::move(key), [&] { return build_einsum(spec, operands, charge); });
}

ProgramPtr plan_contract(const StructurePtr &left, const StructurePtr &right, const std::vector<int64_t> &left_legs,
                         const std::vector<int64_t> &right_legs, const Sector *charge) {
  ProgramKey key{1, "", {left->id(), right->id()}, left_legs, right_legs, charge != nullptr,
                 charge ? charge->to_vector() : std::vector<int64_t>{}};
  return cached(std::move(key), [&] {
    const Category &cat = *left->cat();
    Sector q;
    if (charge) {
      q = *charge;
    } else {
      if (!(right->qtotal() == cat.trivial()))
        throw std::invalid_argument("the second tensor of a contraction must carry the trivial charge");
      q = left->qtotal();
    }
    check_pairs(*left, *right, left_legs, right_legs);
    if (cat.n_symbol(left->qtotal(), right->qtotal(), q) == 0)
      throw std::invalid_argument("the charge " + render(q) + " is not an outcome of " + render(left->qtotal()) + " and " +
                                  render(right->qtotal()));
    require_single_closing(cat, left->qtotal(), right->qtotal(), q, "the charges", "a contraction");
    Builder bld;
    const int32_t l = bld.operand(0, left), r = bld.operand(1, right);
    const int32_t above = bld.move(l, without(left->nlegs(), left_legs), left_legs);
    const int32_t under = bld.move(r, right_legs, without(right->nlegs(), right_legs));
    const int32_t out = bld.gemm(above, under, plan_gemm(bld.structure(above), bld.structure(under), q));
    bld.p->operands = {left, right};
    bld.finish({out});
    return bld.p;
  });
}

ProgramPtr plan_place(const StructurePtr &source, const std::vector<int64_t> &above, const std::vector<int64_t> &below) {
  ProgramKey key{2, "", {source->id()}, above, below, false, {}};
  return cached(std::move(key), [&] {
    Builder bld;
    const int32_t v = bld.operand(0, source);
    const int32_t placed = bld.move(v, above, below);
    bld.p->operands = {source};
    bld.finish({placed});
    return bld.p;
  });
}

// The conjugate's blocks are the transposed blocks under swapped roots (tensor.cpp:105-113).
ProgramPtr plan_conj(const StructurePtr &source) {
  ProgramKey key{3, "", {source->id()}, {}, {}, false, {}};
  return cached(std::move(key), [&] {
    const Category &cat = *source->cat();
    auto move = std::make_shared<MovePlan>();
    move->source = source;
    move->target = Structure::make(source->cat(), source->domain(), source->codomain(), cat.dual(source->qtotal()));
    for (const BlockInfo &b : source->blocks()) {
      const int64_t tb = move->target->find(b.key.domain, b.key.codomain, b.key.copy);
      if (tb < 0) throw std::logic_error("the conjugate has no block for a key of the tensor");
      const BlockInfo &t = move->target->blocks()[static_cast<size_t>(tb)];
      if (t.rows != b.cols || t.cols != b.rows) throw std::logic_error("the conjugate's block does not transpose");
      if (b.rows == 0 || b.cols == 0) continue;
      tt_gss_task task{};
      task.dst_offset = t.offset;
      task.ndim = 2;
      task.first_source = static_cast<int32_t>(move->sources.size());
      task.num_sources = 1;
      task.extent[0] = t.rows;
      task.extent[1] = t.cols;
      task.dst_stride[0] = t.cols;
      task.dst_stride[1] = 1;
      move->tasks.push_back(task);
      tt_gss_source s{};
      s.src_offset = b.offset;
      s.src_stride[0] = 1;
      s.src_stride[1] = b.cols;
      s.coef_re = 1.0;
      move->sources.push_back(s);
    }
    Builder bld;
    const int32_t v = bld.operand(0, source);
    const int32_t d = bld.fresh(move->target, 1u);
    bld.p->ops.push_back(Op{OpKind::move, move, nullptr, v, -1, d, 1u});
    bld.p->move_elements += move->elements();
    bld.p->operands = {source};
    bld.finish({d});
    return bld.p;
  });
}

uint32_t Program::hoistable() const {
  uint32_t mask = 0;
  for (const Op &op : ops)
    if (op.kind == OpKind::move && op.inputs && (op.inputs & (op.inputs - 1)) == 0 &&
        slots[static_cast<size_t>(op.dst)].kind == Slot::workspace)
      mask |= op.inputs;
  return mask;
}

void Program::prepare() const {
  // every table of the program in one upload
  struct Batch {
    Batch() { check(tt_upload_batch_begin(), "table upload"); }
    ~Batch() { (void)tt_upload_batch_end(); }
  } batch;
  for (const Op &op : ops) {
    if (op.kind == OpKind::move) op.move->device();
    if (op.kind == OpKind::gemm) op.gemm->device();
  }
  prepared_ = true;
}

void Program::run(int dtype, const std::vector<const void *> &operand_ptrs, void *out_ptr, void *workspace_ptr,
                  void *stream, Mode mode, uint32_t skip, int64_t first, int64_t last, int ws_dtype, int compute) const {
  if (operand_ptrs.size() != operands.size()) throw std::invalid_argument("a program got the wrong number of operands");
  if (ws_dtype < 0) ws_dtype = dtype;
  if (compute < 0) compute = dtype;
  if (mode == Mode::plain && (ws_dtype != dtype || compute != dtype))
    throw std::invalid_argument("the per-block form runs in one precision");
  if (dtype != TT_C128 && first == 0)
    for (const Op &op : ops)
      if (op.kind == OpKind::move && op.move->complex_coefficients())
        throw std::invalid_argument("a leg move of this program recouples with complex coefficients (" +
                                    op.move->source->cat()->name() + " R or F symbols), which a real tensor "
                                    "cannot hold: use complex128");
  const int64_t es = element_size(ws_dtype);
  char *ws = static_cast<char *>(workspace_ptr);
  auto ptr = [&](int32_t v) -> void * {
    const Slot &s = slots[static_cast<size_t>(v)];
    switch (s.kind) {
      case Slot::operand: return const_cast<void *>(operand_ptrs[static_cast<size_t>(s.index)]);
      case Slot::output: return out_ptr;
      default: return ws + s.offset * es;
    }
  };
  auto type_of = [&](int32_t v) { return slots[static_cast<size_t>(v)].kind == Slot::workspace ? ws_dtype : dtype; };
  if (workspace > 0 && !workspace_ptr) throw std::invalid_argument("a program needs its workspace");
  if (mode == Mode::batched && !prepared_) prepare();
  const int64_t end = last < 0 ? static_cast<int64_t>(ops.size()) : std::min<int64_t>(last, static_cast<int64_t>(ops.size()));
  for (int64_t at = std::max<int64_t>(first, 0); at < end; ++at) {
    const Op &op = ops[static_cast<size_t>(at)];
    const bool hoisted = op.kind == OpKind::move && op.inputs && (op.inputs & (op.inputs - 1)) == 0 &&
                         (op.inputs & skip) && slots[static_cast<size_t>(op.dst)].kind == Slot::workspace;
    if (hoisted && mode == Mode::batched) continue;
    switch (op.kind) {
      case OpKind::move:
        if (mode == Mode::plain && op.move->chained) {
          // Per-step form: intermediates ping-pong through the extra workspace after the main one.
          char *extra = ws + workspace * es;
          const MovePlanPtr plain = plan_move_plain(op.move->source, op.move->above, op.move->below);
          const auto &steps = plain->elementary;
          if (dtype != TT_C128)
            for (const auto &step : steps)
              if (step->complex_coefficients())
                throw std::invalid_argument("a step of this leg move recouples with complex coefficients, which a real "
                                            "tensor cannot hold in the per-block form: use complex128");
          const void *from = ptr(op.a);
          const int64_t half = plain_workspace / 2;
          for (size_t k = 0; k < steps.size(); ++k) {
            void *to = k + 1 == steps.size() ? ptr(op.dst) : extra + static_cast<int64_t>(k % 2) * half * es;
            check(tt_gss_execute_plain(steps[k]->device(), dtype, from, to, stream), "leg move");
            from = to;
          }
        } else {
          check(tt_gss_execute_typed(op.move->device(), type_of(op.a), type_of(op.dst), ptr(op.a), ptr(op.dst), stream),
                "leg move");
        }
        break;
      case OpKind::gemm:
        if (mode == Mode::plain)
          check(tt_ggemm_execute_plain(op.gemm->device(), dtype, ptr(op.a), ptr(op.b), ptr(op.dst), stream), "block products");
        else
          check(tt_ggemm_execute(op.gemm->device(), type_of(op.a), type_of(op.b), type_of(op.dst), compute, ptr(op.a),
                                 ptr(op.b), ptr(op.dst), stream),
                "block products");
        break;
      case OpKind::copy:
      case OpKind::add: {
        const int from = type_of(op.a), to = type_of(op.dst);
        const double keep = op.kind == OpKind::add ? 1.0 : 0.0;
        if (from != to) {
          if (op.kind == OpKind::add) throw std::invalid_argument("a sum over charges runs in one precision");
          check(tt_convert(from, to, out->size(), ptr(op.a), ptr(op.dst), stream), "copy");
        } else if (dtype == TT_C128) {
          check(tt_axpby(TT_F64, 2 * out->size(), 1.0, nullptr, ptr(op.a), keep, nullptr, ptr(op.dst), stream), "copy");
        } else {
          check(tt_axpby(from, out->size(), 1.0, nullptr, ptr(op.a), keep, nullptr, ptr(op.dst), stream), "copy");
        }
        break;
      }
    }
  }
}

void clear_program_cache() { program_cache().clear(); }

}  // namespace tt
