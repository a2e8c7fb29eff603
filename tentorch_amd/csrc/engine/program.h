// Contractions and einsums compiled into programs of device operations.
//
// `contract` (tentorch/linalg/graded/contract.cpp:117-175) becomes: two composed leg moves that
// align the operands, then one grouped GEMM whose groups are the result blocks and whose terms are
// the block products `addmm_` runs one by one (contract.cpp:144-163).  `einsum`
// (tentorch/linalg/graded/einsum.cpp:444-598) becomes the sequence of such steps along the fold
// order the reference picks (einsum.cpp:216-301), with the final placement and the sum over
// intermediate charges.  A program is planned once per (equation, operand structures, charge) and
// cached; running it is a handful of launches with no host synchronisation.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "move.h"

namespace tt {

class GemmPlan {
 public:
  StructurePtr above, under, out;
  std::vector<tt_gemm_group> groups;
  std::vector<tt_gemm_term> terms;
  double flops = 0.0;
  ~GemmPlan();
  tt_ggemm_plan *device();
  // The launch configuration (tile shape and packing, tt_ggemm_num_configs of them); -1 lets the
  // device library pick from the block shapes.  Changing it drops the uploaded tables.
  int config() const { return config_; }
  void set_config(int config);
  // Cost-model inputs for each configuration: one row of kGemmFeatures values per configuration.
  std::vector<float> features(int element_size) const;

 private:
  tt_ggemm_plan *device_ = nullptr;
  int config_ = -1;
};
constexpr int kGemmFeatures = 40;
using GemmPlanPtr = std::shared_ptr<GemmPlan>;

enum class OpKind { move, gemm, copy, add };

struct Op {
  OpKind kind;
  MovePlanPtr move;
  GemmPlanPtr gemm;
  int32_t a = -1, b = -1, dst = -1;  // value ids
  uint32_t inputs = 0;               // operands the result depends on (bit mask)
};

// Where a value of a program lives.
struct Slot {
  enum Kind { operand, workspace, output } kind;
  int32_t index = 0;    // operand index
  int64_t offset = 0;   // element offset into the workspace
  StructurePtr structure;
};

// How a program runs.
enum class Mode {
  batched = 0,  // composed moves, one grouped GEMM launch per contraction step
  plain = 1,    // the reference's per-step, per-block launches, nothing hoisted
};

struct Precision {
  int storage = TT_F64;   // element type of every buffer
  int compute = TT_F64;   // MFMA type of the block products
};

class Program {
 public:
  std::vector<StructurePtr> operands;
  StructurePtr out;
  std::vector<Slot> slots;
  std::vector<Op> ops;
  int64_t workspace = 0;        // elements
  int64_t plain_workspace = 0;  // extra elements for the per-step form's intermediates
  double flops = 0.0;           // real multiply-adds x 2 of all block products
  int64_t move_elements = 0;    // target elements written by all leg moves

  // Run with device pointers.  `skip` is a mask of operands whose hoistable moves are still valid in
  // the workspace from the previous run of this program on the same workspace.
  // Ops [first, last) only when `last` >= 0 (for instrumented runs that time each launch).
  // `dtype` is the type of the operands and of the output; `ws_dtype` the type of every workspace
  // value and `compute` the MFMA type of the block products (both `dtype` when negative).  The FP32
  // form runs FP64 operands with FP32 workspaces and FP32 block products.
  void run(int dtype, const std::vector<const void *> &operand_ptrs, void *out_ptr, void *workspace_ptr,
           void *stream, Mode mode, uint32_t skip, int64_t first = 0, int64_t last = -1, int ws_dtype = -1,
           int compute = -1) const;
  // Operands whose moves can be hoisted (moves that read one operand and write the workspace).
  uint32_t hoistable() const;
  // Uploads the device tables of every grouped op now (they are otherwise uploaded on first run):
  // before stream capture, and before timing single launches.
  void prepare() const;

 private:
  mutable bool prepared_ = false;   // every device table uploaded (prepare ran)
};
using ProgramPtr = std::shared_ptr<Program>;

ProgramPtr plan_einsum(const std::string &spec, const std::vector<StructurePtr> &operands, const Sector *charge);
ProgramPtr plan_contract(const StructurePtr &left, const StructurePtr &right, const std::vector<int64_t> &left_legs,
                         const std::vector<int64_t> &right_legs, const Sector *charge);
ProgramPtr plan_place(const StructurePtr &source, const std::vector<int64_t> &above, const std::vector<int64_t> &below);
// The conjugate transpose's layout move (tensor.cpp:105-113); complex entries also need conj().
ProgramPtr plan_conj(const StructurePtr &source);

void clear_program_cache();

}  // namespace tt
