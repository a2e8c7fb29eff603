// Leg moves: a `place(above, below)` compiled into one gather-scale-scatter table.
//
// The reference runs a move as a chain of elementary steps (exchange two codomain legs, bend the
// last leg of one side onto the other; tentorch/tensors/place.cpp:41-107,199-247), each expanded
// into per-entry strided runs (tentorch/tensors/regroup.cpp:25-417) and applied block by block.
// Here the same chain is composed once, at the granularity of tree pairs (one codomain tree and one
// domain tree of one block): every step maps a tree pair to a combination of tree pairs with the
// F/R coefficients the reference's regroupings carry, and moves the multiplicity digits of the legs
// with it.  The composition is a sparse matrix over tree pairs plus one permutation of the legs, so
// the device table has one tile per target tree pair, each a strided box gathered from its source
// boxes and scaled.
#pragma once

#include <memory>
#include <vector>

#include "../../../include/tentorch_amd.h"
#include "tree.h"

namespace tt {

enum class StepKind { exchange, to_codomain, to_domain };

struct Step {
  StepKind kind;
  int64_t position;  // exchange: the codomain pair (position, position + 1)
};

// The chain of elementary steps of place(above, below) (place.cpp:199-247), with the legs after each.
struct Chain {
  std::vector<Step> steps;
  std::vector<std::pair<Legs, Legs>> sides;
};

// Validate a placement and build its chain; throws std::invalid_argument with the reference's
// messages for out-of-range, repeated or missing legs (place.cpp:203-221).
Chain build_chain(const Legs &codomain, const Legs &domain, const std::vector<int64_t> &above,
                  const std::vector<int64_t> &below);

// The legs named below when only `above` is given (place.cpp:183-197).
std::vector<int64_t> complement(int64_t total, const std::vector<int64_t> &above);

class MovePlan {
 public:
  StructurePtr source, target;
  std::vector<int64_t> above, below;
  bool chained = false;           // built from a place() chain, so it has a per-step form
  bool identity = false;          // the placement is already the layout: nothing to do
  int64_t max_intermediate = 0;   // largest structure between two elementary steps
  std::vector<tt_gss_task> tasks;  // one per target tree pair
  std::vector<tt_gss_source> sources;
  // The per-step form (one table per elementary step, launched per target block), built on request.
  std::vector<std::shared_ptr<MovePlan>> elementary;

  ~MovePlan();
  // The device table, uploaded on first use.
  tt_gss_plan *device();
  int64_t elements() const { return target ? target->size() : 0; }
  // Whether any coefficient has an imaginary part (R or F symbols with phases): such a move cannot
  // act on a real tensor.
  bool complex_coefficients() const {
    for (const tt_gss_source &s : sources)
      if (s.coef_im != 0.0) return true;
    return false;
  }

 private:
  tt_gss_plan *device_ = nullptr;
};

using MovePlanPtr = std::shared_ptr<MovePlan>;

// Plan `place(above, below)` of a tensor of structure `source`; cached by (structure, above, below).
MovePlanPtr plan_move(const StructurePtr &source, const std::vector<int64_t> &above,
                      const std::vector<int64_t> &below);

// The same plan with its per-step form filled in.
MovePlanPtr plan_move_plain(const StructurePtr &source, const std::vector<int64_t> &above,
                            const std::vector<int64_t> &below);

void clear_move_cache();

}  // namespace tt
