// Legs, fusion trees, side layouts and tensor structures of the host engine.
//
// Conventions restate the reference bit for bit, because they fix where every entry of a block lives:
// - a leg keeps its sectors sorted and unique (tentorch/symmetry/tree/tree.cpp:90-140);
// - trees are left-nested and enumerated depth first over the leg blocks in sector order and the
//   fusion outcomes in category order (tree.cpp:400-471);
// - a side's rows are its trees in that order, each spanning the product of its legs'
//   multiplicities, mixed radix with the last leg fastest (tentorch/tensors/layout.cpp:18-32);
// - block keys are root pairs (r, s) with N(r, dual(s), q) > 0, one key per closing copy
//   (layout.cpp:34-47), stored in key order (the reference's std::map<BlockKey, ...>).
//
// Legs and structures are interned: two equal legs are one object, so caches key on addresses.
#pragma once

#include <memory>
#include <unordered_map>
#include <vector>

#include "category.h"

namespace tt {

class Leg;
using LegPtr = std::shared_ptr<const Leg>;

class Leg {
 public:
  // Interned constructor: sorts, validates, returns the one object for these blocks.
  static LegPtr make(CategoryPtr cat, std::vector<Sector> sectors, std::vector<int64_t> multiplicities);

  const CategoryPtr &cat() const { return cat_; }
  const std::vector<Sector> &sectors() const { return sectors_; }
  const std::vector<int64_t> &multiplicities() const { return mults_; }
  int64_t nblocks() const { return static_cast<int64_t>(sectors_.size()); }
  int64_t multiplets() const;
  // Block index of a sector, or -1.
  int64_t find(const Sector &s) const;
  // Block index of a sector (throws std::invalid_argument when absent; layout.cpp:8-16).
  int64_t block_of(const Sector &s) const;
  int64_t mult_of(const Sector &s) const { return mults_[static_cast<size_t>(block_of(s))]; }
  size_t hash() const { return hash_; }
  // A hash of the category and the sectors alone (legs that differ only in multiplicities share it).
  size_t sectors_hash() const { return sectors_hash_; }

  LegPtr fuse(const Leg &other) const;
  LegPtr conj() const;
  LegPtr capped(int64_t max_multiplets) const;

  Leg(CategoryPtr cat, std::vector<Sector> sectors, std::vector<int64_t> mults, size_t hash, size_t sectors_hash)
      : cat_(std::move(cat)), sectors_(std::move(sectors)), mults_(std::move(mults)), hash_(hash),
        sectors_hash_(sectors_hash) {}

 private:
  CategoryPtr cat_;
  std::vector<Sector> sectors_;
  std::vector<int64_t> mults_;
  size_t hash_, sectors_hash_;
};

using Legs = std::vector<LegPtr>;

struct FusionTree {
  std::vector<Sector> sectors;
  std::vector<Sector> inner;
  std::vector<int64_t> copies;
  Sector coupled;

  int64_t size() const { return static_cast<int64_t>(sectors.size()); }
  bool operator==(const FusionTree &o) const {
    return sectors == o.sectors && inner == o.inner && copies == o.copies && coupled == o.coupled;
  }
  bool operator<(const FusionTree &o) const;
};

struct TreeHash {
  size_t operator()(const FusionTree &t) const;
};

const std::vector<Sector> &reachable_charges(const Category &cat, const Legs &legs);
std::vector<FusionTree> fusion_trees(const Category &cat, const Legs &legs, const Sector &coupled);
FusionTree trimmed(const Category &cat, const FusionTree &tree);

struct SectorListHash {
  size_t operator()(const std::vector<Sector> &sectors) const;
};

// The fusion trees of one side at one root, in order, with their lookup.  They depend on the legs'
// sectors only, so the layouts of legs that differ in multiplicities share one set.
struct TreeSet {
  std::vector<FusionTree> trees;
  bool abelian = false;
  // Non-abelian categories: the position of each tree.
  std::unordered_map<FusionTree, int64_t, TreeHash> index;
  // Abelian categories: a tree is fixed by its sector list (one outcome per fusion, no copies), so
  // this index alone serves both lookups.
  std::unordered_map<std::vector<Sector>, int64_t, SectorListHash> by_sectors;
};

struct SideLayout {
  explicit SideLayout(std::shared_ptr<const TreeSet> s) : set(std::move(s)), trees(set->trees), abelian(set->abelian) {}
  std::shared_ptr<const TreeSet> set;
  const std::vector<FusionTree> &trees;   // the set's trees
  std::vector<int64_t> offsets;   // size trees + 1
  std::vector<int64_t> mults;     // per tree, the multiplicity each leg carries (trees x legs)
  int64_t nlegs = 0;
  bool abelian = false;
  int64_t rows() const { return offsets.back(); }
  int64_t find(const FusionTree &t) const {
    if (abelian) {
      const int64_t at = find_sectors(t.sectors);
      return at >= 0 && trees[static_cast<size_t>(at)] == t ? at : -1;
    }
    auto f = set->index.find(t);
    return f == set->index.end() ? -1 : f->second;
  }
  int64_t find_sectors(const std::vector<Sector> &sectors) const {
    auto f = set->by_sectors.find(sectors);
    return f == set->by_sectors.end() ? -1 : f->second;
  }
};

using LayoutPtr = std::shared_ptr<const SideLayout>;

// The layout of one side at one root, built once per (legs, root) and kept.
LayoutPtr layout_of(const CategoryPtr &cat, const Legs &legs, const Sector &root);

struct BlockKey {
  Sector codomain, domain;
  int64_t copy = 0;
  bool operator<(const BlockKey &o) const {
    if (codomain != o.codomain) return codomain < o.codomain;
    if (domain != o.domain) return domain < o.domain;
    return copy < o.copy;
  }
  bool operator==(const BlockKey &o) const { return codomain == o.codomain && domain == o.domain && copy == o.copy; }
};

std::vector<BlockKey> block_keys(const Category &cat, const Legs &codomain, const Legs &domain, const Sector &q);

struct BlockInfo {
  BlockKey key;
  int64_t rows, cols, offset;
  LayoutPtr row_layout;
  LayoutPtr col_layout;
};

class Structure;
using StructurePtr = std::shared_ptr<const Structure>;

// The metadata of one block-sparse tensor: its legs, charge and blocks laid out back to back in one
// flat buffer, in key order, each block row-major.
class Structure {
 public:
  static StructurePtr make(CategoryPtr cat, Legs codomain, Legs domain, Sector qtotal);

  const CategoryPtr &cat() const { return cat_; }
  const Legs &codomain() const { return codomain_; }
  const Legs &domain() const { return domain_; }
  Legs legs() const;
  int64_t nlegs() const { return static_cast<int64_t>(codomain_.size() + domain_.size()); }
  const Sector &qtotal() const { return qtotal_; }
  const std::vector<BlockInfo> &blocks() const { return blocks_; }
  int64_t size() const { return size_; }
  // Index of a block key, or -1.
  int64_t find(const BlockKey &k) const;
  int64_t find(const Sector &r, const Sector &s, int64_t copy = 0) const { return find(BlockKey{r, s, copy}); }
  uint64_t id() const { return id_; }

  Structure(CategoryPtr cat, Legs codomain, Legs domain, Sector qtotal, uint64_t id);

 private:
  CategoryPtr cat_;
  Legs codomain_, domain_;
  Sector qtotal_;
  std::vector<BlockInfo> blocks_;
  std::map<BlockKey, int64_t> index_;
  int64_t size_ = 0;
  uint64_t id_;
};

// Drop every interned leg, layout and structure that nothing else holds.
void clear_engine_caches();

}  // namespace tt
