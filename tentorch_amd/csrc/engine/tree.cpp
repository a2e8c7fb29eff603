#include "tree.h"

#include <algorithm>
#include <functional>
#include <map>
#include <numeric>
#include <set>
#include <unordered_map>
#include <unordered_set>

namespace tt {

namespace {

void mix(size_t &seed, size_t value) { seed ^= value + 0x9e3779b97f4a7c15ULL + (seed << 6) + (seed >> 2); }

struct LegKey {
  const Category *cat;
  std::vector<int64_t> data;
  bool operator<(const LegKey &o) const {
    if (cat != o.cat) return cat < o.cat;
    return data < o.data;
  }
};

std::map<LegKey, LegPtr> &leg_table() {
  static std::map<LegKey, LegPtr> table;
  return table;
}

size_t hash_legs(size_t seed, const std::vector<const Leg *> &legs) {
  for (const Leg *l : legs) mix(seed, std::hash<const void *>()(l));
  mix(seed, legs.size());
  return seed;
}

struct LayoutKey {
  const Category *cat;
  std::vector<const Leg *> legs;
  Sector root;
  bool operator==(const LayoutKey &o) const { return cat == o.cat && legs == o.legs && root == o.root; }
};

struct LayoutKeyHash {
  size_t operator()(const LayoutKey &k) const {
    size_t seed = hash_legs(std::hash<const void *>()(k.cat), k.legs);
    mix(seed, SectorHash()(k.root));
    return seed;
  }
};

std::unordered_map<LayoutKey, LayoutPtr, LayoutKeyHash> &layout_table() {
  static std::unordered_map<LayoutKey, LayoutPtr, LayoutKeyHash> table;
  return table;
}

// Tree sets by (category, the legs' sector lists, root): two keys are equal when every leg has the
// same sectors, whatever its multiplicities.
struct TreeSetKeyEq {
  bool operator()(const LayoutKey &a, const LayoutKey &b) const {
    if (a.cat != b.cat || a.legs.size() != b.legs.size() || !(a.root == b.root)) return false;
    for (size_t i = 0; i < a.legs.size(); ++i)
      if (a.legs[i] != b.legs[i] && a.legs[i]->sectors() != b.legs[i]->sectors()) return false;
    return true;
  }
};

struct TreeSetKeyHash {
  size_t operator()(const LayoutKey &k) const {
    size_t seed = std::hash<const void *>()(k.cat);
    for (const Leg *l : k.legs) mix(seed, l->sectors_hash());
    mix(seed, k.legs.size());
    mix(seed, SectorHash()(k.root));
    return seed;
  }
};

std::unordered_map<LayoutKey, std::shared_ptr<const TreeSet>, TreeSetKeyHash, TreeSetKeyEq> &tree_set_table() {
  static std::unordered_map<LayoutKey, std::shared_ptr<const TreeSet>, TreeSetKeyHash, TreeSetKeyEq> table;
  return table;
}

struct StructureKey {
  const Category *cat;
  std::vector<const Leg *> codomain, domain;
  Sector q;
  bool operator==(const StructureKey &o) const {
    return cat == o.cat && codomain == o.codomain && domain == o.domain && q == o.q;
  }
};

struct StructureKeyHash {
  size_t operator()(const StructureKey &k) const {
    size_t seed = hash_legs(hash_legs(std::hash<const void *>()(k.cat), k.codomain), k.domain);
    mix(seed, SectorHash()(k.q));
    return seed;
  }
};

std::unordered_map<StructureKey, StructurePtr, StructureKeyHash> &structure_table() {
  static std::unordered_map<StructureKey, StructurePtr, StructureKeyHash> table;
  return table;
}

uint64_t next_structure_id() {
  static uint64_t id = 0;
  return ++id;
}

}  // namespace

// Sorted, unique, positive multiplicities (tree.cpp:90-140).
LegPtr Leg::make(CategoryPtr cat, std::vector<Sector> sectors, std::vector<int64_t> mults) {
  if (!cat) throw std::invalid_argument("A leg needs a category.");
  if (sectors.size() != mults.size())
    throw std::invalid_argument("A leg needs one multiplicity per sector, got " + std::to_string(mults.size()) +
                                " for " + std::to_string(sectors.size()) + " sectors.");
  for (int64_t m : mults)
    if (m <= 0) throw std::invalid_argument("A block multiplicity must be positive, got " + std::to_string(m) + ".");
  std::vector<size_t> order(sectors.size());
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return sectors[a] < sectors[b]; });
  std::vector<Sector> ss;
  std::vector<int64_t> ms;
  for (size_t i : order) {
    ss.push_back(sectors[i]);
    ms.push_back(mults[i]);
  }
  for (size_t i = 1; i < ss.size(); ++i)
    if (ss[i] == ss[i - 1]) throw std::invalid_argument("A leg cannot hold the sector " + render(ss[i]) + " twice.");
  for (const Sector &s : ss)
    if (cat->q_dim(s) <= 0.0)
      throw std::invalid_argument("The sector " + render(s) + " has no components in " + cat->name() + ".");
  LegKey key{cat.get(), {}};
  size_t hash = std::hash<std::string>{}(cat->name()), sectors_hash = hash;
  for (size_t i = 0; i < ss.size(); ++i) {
    key.data.push_back(ss[i].n);
    for (int j = 0; j < ss[i].n; ++j) {
      key.data.push_back(ss[i].v[j]);
      mix(hash, std::hash<int64_t>{}(ss[i].v[j]));
      mix(sectors_hash, std::hash<int64_t>{}(ss[i].v[j]));
    }
    mix(sectors_hash, static_cast<size_t>(ss[i].n));
    key.data.push_back(ms[i]);
    mix(hash, std::hash<int64_t>{}(ms[i]));
  }
  auto &table = leg_table();
  auto found = table.find(key);
  if (found != table.end()) return found->second;
  LegPtr leg = std::make_shared<Leg>(cat, std::move(ss), std::move(ms), hash, sectors_hash);
  table.emplace(std::move(key), leg);
  return leg;
}

int64_t Leg::multiplets() const { return std::accumulate(mults_.begin(), mults_.end(), int64_t{0}); }

int64_t Leg::find(const Sector &s) const {
  auto it = std::lower_bound(sectors_.begin(), sectors_.end(), s);
  if (it == sectors_.end() || *it != s) return -1;
  return it - sectors_.begin();
}

int64_t Leg::block_of(const Sector &s) const {
  const int64_t at = find(s);
  if (at < 0)
    throw std::invalid_argument("a leg of " + std::to_string(nblocks()) + " blocks carries none in sector " + render(s));
  return at;
}

LegPtr Leg::fuse(const Leg &other) const {
  if (cat_ != other.cat_)
    throw std::invalid_argument("Cannot fuse a leg of " + cat_->name() + " with one of " + other.cat_->name() + ".");
  std::map<Sector, int64_t> fused;
  for (int64_t a = 0; a < nblocks(); ++a)
    for (int64_t b = 0; b < other.nblocks(); ++b)
      for (const Outcome &o : cat_->outcomes(sectors_[a], other.sectors_[b]))
        fused[o.sector] += mults_[a] * other.mults_[b] * o.multiplicity;
  std::vector<Sector> ss;
  std::vector<int64_t> ms;
  for (const auto &[s, m] : fused) {
    ss.push_back(s);
    ms.push_back(m);
  }
  return make(cat_, ss, ms);
}

LegPtr Leg::conj() const {
  std::vector<Sector> ss;
  for (const Sector &s : sectors_) ss.push_back(cat_->dual(s));
  return make(cat_, ss, mults_);
}

LegPtr Leg::capped(int64_t max_multiplets) const {
  if (max_multiplets <= 0)
    throw std::invalid_argument("A cap must keep at least one multiplet, got " + std::to_string(max_multiplets) + ".");
  std::vector<Sector> ss;
  std::vector<int64_t> ms;
  int64_t kept = 0;
  for (int64_t b = 0; b < nblocks() && kept < max_multiplets; ++b) {
    const int64_t take = std::min(mults_[b], max_multiplets - kept);
    ss.push_back(sectors_[b]);
    ms.push_back(take);
    kept += take;
  }
  return make(cat_, ss, ms);
}

bool FusionTree::operator<(const FusionTree &o) const {
  if (sectors != o.sectors) return sectors < o.sectors;
  if (inner != o.inner) return inner < o.inner;
  if (copies != o.copies) return copies < o.copies;
  return coupled < o.coupled;
}

size_t SectorListHash::operator()(const std::vector<Sector> &sectors) const {
  size_t seed = 0;
  for (const Sector &s : sectors) mix(seed, SectorHash{}(s));
  return seed;
}

size_t TreeHash::operator()(const FusionTree &t) const {
  size_t seed = 0;
  for (const Sector &s : t.sectors) mix(seed, SectorHash{}(s));
  for (const Sector &s : t.inner) mix(seed, SectorHash{}(s) * 31u);
  for (int64_t c : t.copies) mix(seed, std::hash<int64_t>{}(c));
  mix(seed, SectorHash{}(t.coupled));
  return seed;
}

namespace {
// Reachable charges depend on the legs' sectors only: keys are equal when every leg has the same
// sectors, whatever its multiplicities.
using ReachableKey = std::pair<const Category *, std::vector<const Leg *>>;

struct ReachableKeyHash {
  size_t operator()(const ReachableKey &k) const {
    size_t seed = std::hash<const void *>()(k.first);
    for (const Leg *l : k.second) mix(seed, l->sectors_hash());
    mix(seed, k.second.size());
    return seed;
  }
};

struct ReachableKeyEq {
  bool operator()(const ReachableKey &a, const ReachableKey &b) const {
    if (a.first != b.first || a.second.size() != b.second.size()) return false;
    for (size_t i = 0; i < a.second.size(); ++i)
      if (a.second[i] != b.second[i] && a.second[i]->sectors() != b.second[i]->sectors()) return false;
    return true;
  }
};

std::unordered_map<ReachableKey, std::vector<Sector>, ReachableKeyHash, ReachableKeyEq> &reachable_table() {
  static std::unordered_map<ReachableKey, std::vector<Sector>, ReachableKeyHash, ReachableKeyEq> table;
  return table;
}
std::vector<Sector> compute_reachable(const Category &cat, const Legs &legs);
}  // namespace

// Insertion order of first appearance (tree.cpp:375-398), kept per (category, legs).
const std::vector<Sector> &reachable_charges(const Category &cat, const Legs &legs) {
  std::pair<const Category *, std::vector<const Leg *>> key{&cat, {}};
  key.second.reserve(legs.size());
  for (const LegPtr &l : legs) key.second.push_back(l.get());
  auto &table = reachable_table();
  auto found = table.find(key);
  if (found != table.end()) return found->second;
  return table.emplace(std::move(key), compute_reachable(cat, legs)).first->second;
}

namespace {
std::vector<Sector> compute_reachable(const Category &cat, const Legs &legs) {
  std::vector<Sector> charges{cat.trivial()};
  for (const LegPtr &leg : legs) {
    std::vector<Sector> next;
    std::set<Sector> seen;
    for (const Sector &charge : charges)
      for (const Sector &s : leg->sectors())
        for (const Outcome &o : cat.outcomes(charge, s))
          if (seen.insert(o.sector).second) next.push_back(o.sector);
    charges = std::move(next);
  }
  return charges;
}
}  // namespace

namespace {
// Every fusion tree of `legs`, depth first over leg blocks in sector order and outcomes in category
// order (tree.cpp:400-471), handed to `emit` as it is completed; `coupled`, when given, keeps only
// the trees of that root (the last fusion is filtered, the walk above it is shared).
template <class Emit>
void walk_trees(const Category &cat, const Legs &legs, const Sector *coupled, Emit &&emit) {
  const int64_t count = static_cast<int64_t>(legs.size());
  if (count == 0) {
    if (!coupled || *coupled == cat.trivial()) emit(FusionTree{{}, {}, {}, cat.trivial()});
    return;
  }
  FusionTree tree;
  tree.sectors.resize(static_cast<size_t>(count));
  tree.inner.resize(static_cast<size_t>(count >= 2 ? count - 2 : 0));
  tree.copies.resize(static_cast<size_t>(count >= 2 ? count - 1 : 0));
  std::function<void(int64_t, const Sector &)> walk = [&](int64_t leg, const Sector &previous) {
    if (leg == 0) {
      for (const Sector &s : legs[0]->sectors()) {
        tree.sectors[0] = s;
        if (count == 1) {
          if (!coupled || s == *coupled) {
            tree.coupled = s;
            emit(tree);
          }
        } else {
          walk(1, s);
        }
      }
      return;
    }
    for (const Sector &s : legs[static_cast<size_t>(leg)]->sectors()) {
      tree.sectors[static_cast<size_t>(leg)] = s;
      for (const Outcome &o : cat.outcomes(previous, s)) {
        if (leg + 1 == count && coupled && o.sector != *coupled) continue;
        if (leg + 1 < count) tree.inner[static_cast<size_t>(leg - 1)] = o.sector;
        for (int64_t copy = 0; copy < o.multiplicity; ++copy) {
          tree.copies[static_cast<size_t>(leg - 1)] = copy;
          if (leg + 1 == count) {
            tree.coupled = o.sector;
            emit(tree);
          } else {
            walk(leg + 1, o.sector);
          }
        }
      }
    }
  };
  walk(0, cat.trivial());
}
}  // namespace

std::vector<FusionTree> fusion_trees(const Category &cat, const Legs &legs, const Sector &coupled) {
  std::vector<FusionTree> out;
  walk_trees(cat, legs, &coupled, [&](const FusionTree &t) { out.push_back(t); });
  return out;
}

FusionTree trimmed(const Category &cat, const FusionTree &tree) {
  FusionTree out = tree;
  out.sectors.pop_back();
  if (!out.inner.empty()) out.inner.pop_back();
  if (!out.copies.empty()) out.copies.pop_back();
  const int64_t n = tree.size();
  out.coupled = n <= 1 ? cat.trivial() : n == 2 ? tree.sectors.at(0) : tree.inner.at(static_cast<size_t>(n - 3));
  return out;
}

LayoutPtr layout_of(const CategoryPtr &cat, const Legs &legs, const Sector &root) {
  LayoutKey key{cat.get(), {}, root};
  key.legs.reserve(legs.size());
  for (const LegPtr &l : legs) key.legs.push_back(l.get());
  auto &table = layout_table();
  auto found = table.find(key);
  if (found != table.end()) return found->second;
  // the trees and their lookup, shared with every layout over legs of the same sectors
  auto &sets = tree_set_table();
  auto known = sets.find(key);
  if (known == sets.end()) {
    auto set = std::make_shared<TreeSet>();
    set->trees = fusion_trees(*cat, legs, root);
    set->abelian = cat->abelian();
    // an abelian tree is fixed by its sector list, so one index serves both lookups
    if (set->abelian) {
      set->by_sectors.reserve(set->trees.size());
      for (size_t t = 0; t < set->trees.size(); ++t) set->by_sectors.emplace(set->trees[t].sectors, static_cast<int64_t>(t));
    } else {
      set->index.reserve(set->trees.size());
      for (size_t t = 0; t < set->trees.size(); ++t) set->index.emplace(set->trees[t], static_cast<int64_t>(t));
    }
    known = sets.emplace(key, std::move(set)).first;
  }
  auto layout = std::make_shared<SideLayout>(known->second);
  layout->nlegs = static_cast<int64_t>(legs.size());
  layout->offsets.reserve(layout->trees.size() + 1);
  layout->offsets.push_back(0);
  layout->mults.reserve(layout->trees.size() * legs.size());
  for (size_t t = 0; t < layout->trees.size(); ++t) {
    int64_t rows = 1;
    for (size_t axis = 0; axis < legs.size(); ++axis) {
      const int64_t m = legs[axis]->mult_of(layout->trees[t].sectors[axis]);
      layout->mults.push_back(m);
      rows *= m;
    }
    layout->offsets.push_back(layout->offsets.back() + rows);
  }
  table.emplace(std::move(key), layout);
  return layout;
}

std::vector<BlockKey> block_keys(const Category &cat, const Legs &codomain, const Legs &domain, const Sector &q) {
  std::vector<BlockKey> keys;
  const std::vector<Sector> &rows = reachable_charges(cat, codomain), &cols = reachable_charges(cat, domain);
  if (cat.abelian()) {
    // r x dual(s) = q has the one solution s = r x dual(q), with no copies
    std::unordered_set<Sector, SectorHash> reached(cols.begin(), cols.end());
    const Sector back = cat.dual(q);
    for (const Sector &r : rows)
      for (const Outcome &o : cat.outcomes(r, back))
        if (reached.count(o.sector)) keys.push_back(BlockKey{r, o.sector, 0});
    return keys;
  }
  for (const Sector &r : rows)
    for (const Sector &s : cols)
      for (int64_t copy = 0; copy < cat.n_symbol(r, cat.dual(s), q); ++copy) keys.push_back(BlockKey{r, s, copy});
  return keys;
}

Structure::Structure(CategoryPtr cat, Legs codomain, Legs domain, Sector qtotal, uint64_t id)
    : cat_(std::move(cat)), codomain_(std::move(codomain)), domain_(std::move(domain)), qtotal_(qtotal), id_(id) {
  std::vector<BlockKey> keys = block_keys(*cat_, codomain_, domain_, qtotal_);
  std::sort(keys.begin(), keys.end());
  int64_t offset = 0;
  // a root's layout is shared by every block over it
  std::unordered_map<Sector, LayoutPtr, SectorHash> rows_at, cols_at;
  for (const BlockKey &k : keys) {
    auto r = rows_at.find(k.codomain);
    if (r == rows_at.end()) r = rows_at.emplace(k.codomain, layout_of(cat_, codomain_, k.codomain)).first;
    auto c = cols_at.find(k.domain);
    if (c == cols_at.end()) c = cols_at.emplace(k.domain, layout_of(cat_, domain_, k.domain)).first;
    const LayoutPtr &rl = r->second, &cl = c->second;
    BlockInfo info{k, rl->rows(), cl->rows(), offset, rl, cl};
    index_.emplace(k, static_cast<int64_t>(blocks_.size()));
    blocks_.push_back(info);
    offset += info.rows * info.cols;
  }
  size_ = offset;
}

StructurePtr Structure::make(CategoryPtr cat, Legs codomain, Legs domain, Sector qtotal) {
  if (!cat) throw std::invalid_argument("a tensor needs a category");
  for (const LegPtr &l : codomain)
    if (l->cat() != cat) throw std::invalid_argument("a codomain leg is over a different category");
  for (const LegPtr &l : domain)
    if (l->cat() != cat) throw std::invalid_argument("a domain leg is over a different category");
  StructureKey key{cat.get(), {}, {}, qtotal};
  for (const LegPtr &l : codomain) key.codomain.push_back(l.get());
  for (const LegPtr &l : domain) key.domain.push_back(l.get());
  auto &table = structure_table();
  auto found = table.find(key);
  if (found != table.end()) return found->second;
  auto s = std::make_shared<Structure>(cat, std::move(codomain), std::move(domain), qtotal, next_structure_id());
  table.emplace(std::move(key), s);
  return s;
}

Legs Structure::legs() const {
  Legs out = codomain_;
  out.insert(out.end(), domain_.begin(), domain_.end());
  return out;
}

int64_t Structure::find(const BlockKey &k) const {
  auto f = index_.find(k);
  return f == index_.end() ? -1 : f->second;
}

void clear_engine_caches() {
  reachable_table().clear();
  structure_table().clear();
  layout_table().clear();
  tree_set_table().clear();
}

}  // namespace tt
