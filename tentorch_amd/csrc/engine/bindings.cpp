// Python binding of the host engine: categories, legs, fusion trees, tensor structures, and the
// planned programs (einsum, contract, place, conj) that run on the device through the C ABI.
// Mirrors the reference's module surface (tentorch/_ext/bindings.cpp:166-471) for the objects the
// contraction path needs; data pointers and streams cross as integers.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "program.h"
#include "split.h"

namespace py = pybind11;
using namespace tt;

namespace tt {
std::string device_error();
}

namespace {

Sector to_sector(const py::handle &h) {
  std::vector<int64_t> v;
  for (const py::handle item : h) v.push_back(item.cast<int64_t>());
  return Sector(v);
}

py::tuple from_sector(const Sector &s) {
  py::tuple t(s.n);
  for (int i = 0; i < s.n; ++i) t[static_cast<size_t>(i)] = py::int_(s.v[i]);
  return t;
}

py::list sectors_list(const std::vector<Sector> &ss) {
  py::list out;
  for (const Sector &s : ss) out.append(from_sector(s));
  return out;
}

py::array_t<std::complex<double>> matrix_array(const Matrix &m) {
  py::array_t<std::complex<double>> a({m.rows, m.cols});
  auto r = a.mutable_unchecked<2>();
  for (int64_t i = 0; i < m.rows; ++i)
    for (int64_t j = 0; j < m.cols; ++j) r(i, j) = m(i, j);
  return a;
}

struct PyCategory {
  CategoryPtr cat;
};

struct PyLeg {
  LegPtr leg;
};

struct PyStructure {
  StructurePtr s;
};

struct PyProgram {
  ProgramPtr p;
};

// The block-product op at index `op` of a program.
const Op &gemm_op(const PyProgram &p, int64_t op) {
  if (op < 0 || op >= static_cast<int64_t>(p.p->ops.size())) throw std::out_of_range("the program has no op " + std::to_string(op));
  const Op &o = p.p->ops[static_cast<size_t>(op)];
  if (o.kind != OpKind::gemm) throw std::invalid_argument("op " + std::to_string(op) + " is not a block product");
  return o;
}

Legs to_legs(const std::vector<PyLeg> &legs) {
  Legs out;
  for (const PyLeg &l : legs) out.push_back(l.leg);
  return out;
}

py::object tree_object(const FusionTree &t) {
  py::dict d;
  return py::make_tuple(sectors_list(t.sectors), sectors_list(t.inner), py::cast(t.copies), from_sector(t.coupled));
}

}  // namespace

// (kept per block, bond offset per block, bond sectors, bond multiplicities, kept values) of a split ranking.
py::tuple ranking_tuple(const SplitRanking &r) {
  py::list sectors;
  for (const Sector &x : r.sectors) sectors.append(from_sector(x));
  py::array_t<double> kept_values(static_cast<py::ssize_t>(r.values.size()));
  std::copy(r.values.begin(), r.values.end(), kept_values.mutable_data());
  return py::make_tuple(r.kept, r.offset, sectors, r.mults, kept_values);
}

PYBIND11_MODULE(_engine, m) {
  m.doc() = "TenTorch MI355X host engine";

  py::register_exception_translator([](std::exception_ptr p) {
    try {
      if (p) std::rethrow_exception(p);
    } catch (const std::out_of_range &e) {
      PyErr_SetString(PyExc_IndexError, e.what());
    }
  });

  py::class_<PyCategory>(m, "FusionCategory")
      .def_property_readonly("name", [](const PyCategory &c) { return c.cat->name(); })
      .def_property_readonly("abelian", [](const PyCategory &c) { return c.cat->abelian(); })
      .def_property_readonly("real_symbols", [](const PyCategory &c) { return c.cat->real_symbols(); })
      .def("outcomes",
           [](const PyCategory &c, py::handle a, py::handle b) {
             py::list out;
             for (const Outcome &o : c.cat->outcomes(to_sector(a), to_sector(b)))
               out.append(py::make_tuple(from_sector(o.sector), o.multiplicity));
             return out;
           })
      .def("n_symbol", [](const PyCategory &c, py::handle a, py::handle b, py::handle x) {
        return c.cat->n_symbol(to_sector(a), to_sector(b), to_sector(x));
      })
      .def("can_fuse_to", [](const PyCategory &c, py::handle a, py::handle b, py::handle x) {
        return c.cat->n_symbol(to_sector(a), to_sector(b), to_sector(x)) > 0;
      })
      .def("F_symbol", [](const PyCategory &c, py::handle a, py::handle b, py::handle x, py::handle d) {
        return matrix_array(c.cat->F(to_sector(a), to_sector(b), to_sector(x), to_sector(d)));
      })
      .def("R_symbol", [](const PyCategory &c, py::handle b, py::handle x) {
        return matrix_array(c.cat->R(to_sector(b), to_sector(x)));
      })
      .def("q_dim", [](const PyCategory &c, py::handle a) { return c.cat->q_dim(to_sector(a)); })
      .def("dual", [](const PyCategory &c, py::handle a) { return from_sector(c.cat->dual(to_sector(a))); })
      .def("frobenius_schur_sign", [](const PyCategory &c, py::handle a) { return c.cat->frobenius_schur_sign(to_sector(a)); })
      .def("trivial", [](const PyCategory &c) { return from_sector(c.cat->trivial()); })
      .def("clear_cache", [](const PyCategory &c) { c.cat->clear_cache(); })
      .def("__eq__", [](const PyCategory &a, const PyCategory &b) { return a.cat == b.cat; })
      .def("__hash__", [](const PyCategory &c) { return reinterpret_cast<uintptr_t>(c.cat.get()); })
      .def("__repr__", [](const PyCategory &c) { return c.cat->name(); });

  m.def("no_symmetry", [] { return PyCategory{no_symmetry()}; });
  m.def("u1", [] { return PyCategory{u1_symmetry()}; });
  m.def("z_n", [](int64_t n) { return PyCategory{z_n_symmetry(n)}; });
  m.def("fermion_parity", [] { return PyCategory{fermion_parity()}; });
  m.def("su2", [] { return PyCategory{su2_symmetry()}; });
  m.def("fibonacci", [] { return PyCategory{fibonacci_category()}; });

  py::class_<PyLeg>(m, "Leg")
      .def(py::init([](const PyCategory &cat, py::iterable sectors, py::iterable mults) {
             std::vector<Sector> ss;
             for (const py::handle s : sectors) ss.push_back(to_sector(s));
             std::vector<int64_t> ms;
             for (const py::handle x : mults) ms.push_back(x.cast<int64_t>());
             return PyLeg{Leg::make(cat.cat, ss, ms)};
           }),
           py::arg("cat"), py::arg("sectors"), py::arg("multiplicities"))
      .def_property_readonly("cat", [](const PyLeg &l) { return PyCategory{l.leg->cat()}; })
      .def_property_readonly("sectors", [](const PyLeg &l) { return sectors_list(l.leg->sectors()); })
      .def_property_readonly("multiplicities", [](const PyLeg &l) { return l.leg->multiplicities(); })
      .def("nblocks", [](const PyLeg &l) { return l.leg->nblocks(); })
      .def("sector", [](const PyLeg &l, int64_t b) {
        if (b < 0 || b >= l.leg->nblocks()) throw std::out_of_range("block index out of range");
        return from_sector(l.leg->sectors()[static_cast<size_t>(b)]);
      })
      .def("multiplicity", [](const PyLeg &l, int64_t b) {
        if (b < 0 || b >= l.leg->nblocks()) throw std::out_of_range("block index out of range");
        return l.leg->multiplicities()[static_cast<size_t>(b)];
      })
      .def("multiplets", [](const PyLeg &l) { return l.leg->multiplets(); })
      .def("fuse", [](const PyLeg &l, const PyLeg &o) { return PyLeg{l.leg->fuse(*o.leg)}; })
      .def("conj", [](const PyLeg &l) { return PyLeg{l.leg->conj()}; })
      .def("capped", [](const PyLeg &l, int64_t n) { return PyLeg{l.leg->capped(n)}; })
      .def("__eq__", [](const PyLeg &a, const PyLeg &b) { return a.leg == b.leg; })
      .def("__ne__", [](const PyLeg &a, const PyLeg &b) { return a.leg != b.leg; })
      .def("__hash__", [](const PyLeg &l) { return l.leg->hash(); })
      .def("_address", [](const PyLeg &l) { return reinterpret_cast<uintptr_t>(l.leg.get()); })
      .def("__repr__", [](const PyLeg &l) {
        std::string out = "Leg(" + l.leg->cat()->name() + ", [";
        for (int64_t b = 0; b < l.leg->nblocks(); ++b) {
          if (b) out += ", ";
          out += render(l.leg->sectors()[static_cast<size_t>(b)]) + ": " + std::to_string(l.leg->multiplicities()[static_cast<size_t>(b)]);
        }
        return out + "])";
      });

  m.def("fusion_trees", [](const PyCategory &cat, const std::vector<PyLeg> &legs, py::handle coupled) {
    py::list out;
    for (const FusionTree &t : fusion_trees(*cat.cat, to_legs(legs), to_sector(coupled))) out.append(tree_object(t));
    return out;
  });
  m.def("reachable_charges", [](const PyCategory &cat, const std::vector<PyLeg> &legs) {
    return sectors_list(reachable_charges(*cat.cat, to_legs(legs)));
  });
  m.def("block_keys", [](const PyCategory &cat, const std::vector<PyLeg> &cod, const std::vector<PyLeg> &dom, py::handle q) {
    py::list out;
    for (const BlockKey &k : block_keys(*cat.cat, to_legs(cod), to_legs(dom), to_sector(q)))
      out.append(py::make_tuple(from_sector(k.codomain), from_sector(k.domain), k.copy));
    return out;
  });
  m.def("side_layout", [](const PyCategory &cat, const std::vector<PyLeg> &legs, py::handle root) {
    LayoutPtr l = layout_of(cat.cat, to_legs(legs), to_sector(root));
    py::list trees;
    for (const FusionTree &t : l->trees) trees.append(tree_object(t));
    return py::make_tuple(trees, l->offsets);
  });

  py::class_<PyStructure>(m, "Structure")
      .def(py::init([](const PyCategory &cat, const std::vector<PyLeg> &cod, const std::vector<PyLeg> &dom, py::handle q) {
             return PyStructure{Structure::make(cat.cat, to_legs(cod), to_legs(dom), to_sector(q))};
           }),
           py::arg("cat"), py::arg("codomain"), py::arg("domain"), py::arg("qtotal"))
      .def_property_readonly("cat", [](const PyStructure &s) { return PyCategory{s.s->cat()}; })
      .def_property_readonly("codomain", [](const PyStructure &s) {
        std::vector<PyLeg> out;
        for (const LegPtr &l : s.s->codomain()) out.push_back(PyLeg{l});
        return out;
      })
      .def_property_readonly("domain", [](const PyStructure &s) {
        std::vector<PyLeg> out;
        for (const LegPtr &l : s.s->domain()) out.push_back(PyLeg{l});
        return out;
      })
      .def_property_readonly("qtotal", [](const PyStructure &s) { return from_sector(s.s->qtotal()); })
      .def_property_readonly("size", [](const PyStructure &s) { return s.s->size(); })
      .def_property_readonly("id", [](const PyStructure &s) { return s.s->id(); })
      .def("nlegs", [](const PyStructure &s) { return s.s->nlegs(); })
      .def("nblocks", [](const PyStructure &s) { return static_cast<int64_t>(s.s->blocks().size()); })
      .def("blocks", [](const PyStructure &s) {
        py::list out;
        for (const BlockInfo &b : s.s->blocks())
          out.append(py::make_tuple(py::make_tuple(from_sector(b.key.codomain), from_sector(b.key.domain), b.key.copy),
                                    b.rows, b.cols, b.offset));
        return out;
      })
      .def("find", [](const PyStructure &s, py::handle r, py::handle d, int64_t copy) {
        return s.s->find(to_sector(r), to_sector(d), copy);
      })
      .def("__eq__", [](const PyStructure &a, const PyStructure &b) { return a.s == b.s; })
      .def("__hash__", [](const PyStructure &s) { return s.s->id(); });

  py::class_<PyProgram>(m, "Program")
      .def_property_readonly("out", [](const PyProgram &p) { return PyStructure{p.p->out}; })
      .def_property_readonly("workspace", [](const PyProgram &p) { return p.p->workspace; })
      .def_property_readonly("plain_workspace", [](const PyProgram &p) { return p.p->plain_workspace; })
      .def_property_readonly("flops", [](const PyProgram &p) { return p.p->flops; })
      .def_property_readonly("move_elements", [](const PyProgram &p) { return p.p->move_elements; })
      .def_property_readonly("hoistable", [](const PyProgram &p) { return p.p->hoistable(); })
      .def_property_readonly("id", [](const PyProgram &p) { return reinterpret_cast<uintptr_t>(p.p.get()); })
      .def("op_kinds", [](const PyProgram &p) {
        std::vector<std::string> out;
        for (const Op &op : p.p->ops)
          out.push_back(op.kind == OpKind::move ? "move" : op.kind == OpKind::gemm ? "gemm" : op.kind == OpKind::copy ? "copy" : "add");
        return out;
      })
      .def("ops",
           [](const PyProgram &p) {
             py::list out;
             for (const Op &op : p.p->ops)
               out.append(py::make_tuple(op.kind == OpKind::move ? "move" : op.kind == OpKind::gemm ? "gemm"
                                         : op.kind == OpKind::copy ? "copy" : "add",
                                         op.a, op.b, op.dst, op.inputs));
             return out;
           })
      .def("slots",
           [](const PyProgram &p) {
             py::list out;
             for (const Slot &s : p.p->slots)
               out.append(py::make_tuple(s.kind == Slot::operand ? "operand" : s.kind == Slot::output ? "output" : "workspace",
                                         s.index, s.offset, s.structure->size()));
             return out;
           })
      .def("move_tables",
           [](const PyProgram &p, int64_t which) {
             // The descriptor table of the which-th leg move, for inspection.
             int64_t seen = 0;
             for (const Op &op : p.p->ops) {
               if (op.kind != OpKind::move) continue;
               if (seen++ != which) continue;
               py::list tasks;
               for (const tt_gss_task &t : op.move->tasks) {
                 std::vector<int64_t> ext(t.extent, t.extent + t.ndim), ds(t.dst_stride, t.dst_stride + t.ndim);
                 tasks.append(py::make_tuple(t.dst_offset, ext, ds, t.first_source, t.num_sources));
               }
               py::list sources;
               for (const tt_gss_source &s : op.move->sources) {
                 std::vector<int64_t> ss(s.src_stride, s.src_stride + TT_MAX_DIMS);
                 sources.append(py::make_tuple(s.src_offset, ss, std::complex<double>(s.coef_re, s.coef_im)));
               }
               return py::make_tuple(tasks, sources);
             }
             throw std::out_of_range("the program has no such leg move");
           })
      .def("gemm_tables",
           [](const PyProgram &p, int64_t which) {
             int64_t seen = 0;
             for (const Op &op : p.p->ops) {
               if (op.kind != OpKind::gemm) continue;
               if (seen++ != which) continue;
               py::list groups, terms;
               for (const tt_gemm_group &g : op.gemm->groups)
                 groups.append(py::make_tuple(g.c_offset, g.m, g.n, g.first_term, g.num_terms));
               for (const tt_gemm_term &t : op.gemm->terms) terms.append(py::make_tuple(t.a_offset, t.b_offset, t.k));
               return py::make_tuple(groups, terms);
             }
             throw std::out_of_range("the program has no such block product");
           })
      .def("run",
           [](const PyProgram &p, int dtype, const std::vector<uintptr_t> &operands, uintptr_t out, uintptr_t workspace,
              uintptr_t stream, int mode, uint32_t skip, int64_t first, int64_t last, int ws_dtype, int compute) {
             std::vector<const void *> ptrs;
             for (uintptr_t x : operands) ptrs.push_back(reinterpret_cast<const void *>(x));
             py::gil_scoped_release release;
             p.p->run(dtype, ptrs, reinterpret_cast<void *>(out), reinterpret_cast<void *>(workspace),
                      reinterpret_cast<void *>(stream), mode == 1 ? Mode::plain : Mode::batched, skip, first, last,
                      ws_dtype, compute);
           },
           py::arg("dtype"), py::arg("operands"), py::arg("out"), py::arg("workspace"), py::arg("stream"),
           py::arg("mode") = 0, py::arg("skip") = 0, py::arg("first") = 0, py::arg("last") = -1,
           py::arg("ws_dtype") = -1, py::arg("compute") = -1)
      .def("krylov",
           [](const PyProgram &p, const std::vector<uintptr_t> &operands, int64_t slot, uintptr_t basis, int64_t ldv,
              int64_t steps, uintptr_t image, uintptr_t scratch, uintptr_t alphas, uintptr_t betas, uintptr_t workspace,
              uintptr_t stream, uint32_t skip, int ws_dtype, int compute) {
             // The Krylov steps of a Lanczos run whose action is this program with the basis row in
             // operand `slot` (tentorch/linalg/lanczos.py:27-49): for k < steps, image = program(v_k),
             // then tt_lanczos_orthonormalize makes v_{k+1} and the step's alpha and beta, all on the
             // stream with no synchronisation.  The operands other than `slot` do not change between
             // steps, so after the first step their hoisted moves are skipped.
             if (slot < 0 || slot >= static_cast<int64_t>(operands.size()))
               throw std::invalid_argument("krylov: the basis operand is out of range");
             if (steps < 1 || steps > 31) throw std::invalid_argument("krylov: 1..31 steps");
             const int64_t n = p.p->operands[static_cast<size_t>(slot)]->size();
             if (p.p->out->size() != n) throw std::invalid_argument("krylov: the action has to keep the vector's size");
             if (ldv < n) throw std::invalid_argument("krylov: basis rows shorter than the vector");
             std::vector<const void *> ptrs;
             for (uintptr_t x : operands) ptrs.push_back(reinterpret_cast<const void *>(x));
             double *V = reinterpret_cast<double *>(basis);
             const uint32_t fixed = p.p->hoistable() & ~(1u << slot);
             py::gil_scoped_release release;
             for (int64_t k = 0; k < steps; ++k) {
               ptrs[static_cast<size_t>(slot)] = V + k * ldv;
               p.p->run(TT_F64, ptrs, reinterpret_cast<void *>(image), reinterpret_cast<void *>(workspace),
                        reinterpret_cast<void *>(stream), Mode::batched, k == 0 ? skip : (skip | fixed), 0, -1,
                        ws_dtype, compute);
               if (tt_lanczos_orthonormalize(n, static_cast<int>(k), V, ldv, reinterpret_cast<double *>(image),
                                             reinterpret_cast<double *>(scratch), reinterpret_cast<double *>(alphas) + k,
                                             reinterpret_cast<double *>(betas) + k, reinterpret_cast<void *>(stream)) != 0)
                 throw std::runtime_error("lanczos orthonormalize: " + device_error());
             }
           },
           py::arg("operands"), py::arg("slot"), py::arg("basis"), py::arg("ldv"), py::arg("steps"), py::arg("image"),
           py::arg("scratch"), py::arg("alphas"), py::arg("betas"), py::arg("workspace"), py::arg("stream"),
           py::arg("skip") = 0, py::arg("ws_dtype") = -1, py::arg("compute") = -1)
      .def("prepare", [](const PyProgram &p) { p.p->prepare(); })
      .def("gemm_config",
           [](const PyProgram &p, int64_t op) {
             const Op &o = gemm_op(p, op);
             return tt_ggemm_plan_config(o.gemm->device());
           })
      .def("set_gemm_config", [](const PyProgram &p, int64_t op, int config) { gemm_op(p, op).gemm->set_config(config); })
      .def("gemm_features",
           [](const PyProgram &p, int64_t op, int element_size) {
             const std::vector<float> f = gemm_op(p, op).gemm->features(element_size);
             const int64_t rows = static_cast<int64_t>(f.size()) / kGemmFeatures;
             py::array_t<float> out({rows, static_cast<int64_t>(kGemmFeatures)});
             std::copy(f.begin(), f.end(), out.mutable_data());
             return out;
           })
      .def("op_costs",
           [](const PyProgram &p) {
             // Per op: (kind, algorithmic flops, elements written, elements read).
             py::list out;
             for (const Op &op : p.p->ops) {
               if (op.kind == OpKind::gemm)
                 out.append(py::make_tuple("gemm", op.gemm->flops, op.gemm->out->size(), op.gemm->above->size() + op.gemm->under->size()));
               else if (op.kind == OpKind::move)
                 out.append(py::make_tuple("move", 0.0, op.move->target->size(), op.move->source->size()));
               else
                 out.append(py::make_tuple(op.kind == OpKind::copy ? "copy" : "add", 0.0, p.p->out->size(), p.p->out->size()));
             }
             return out;
           });

  m.def("choose_tails",
        [](py::array_t<double, py::array::c_style | py::array::forcecast> norms,
           py::array_t<int64_t, py::array::c_style | py::array::forcecast> norms_at,
           py::array_t<int64_t, py::array::c_style | py::array::forcecast> ranks,
           py::array_t<int64_t, py::array::c_style | py::array::forcecast> first, double floor, double reliable,
           double resolved) {
          const int64_t blocks = static_cast<int64_t>(ranks.size());
          if (norms_at.size() != blocks + 1 || first.size() != blocks)
            throw std::invalid_argument("choose_tails: arrays of unequal length");
          return choose_tails(norms.data(), norms_at.data(), blocks, ranks.data(), first.data(), floor, reliable,
                              resolved);
        });
  m.def("rank_split",
        [](const PyStructure &s, const std::vector<int64_t> &blocks,
           py::array_t<double, py::array::c_style | py::array::forcecast> values, const std::vector<int64_t> &at,
           int64_t max_bond, double tol) {
          if (values.size() < (at.empty() ? 0 : at.back())) throw std::invalid_argument("rank_split: too few values");
          return ranking_tuple(rank_split(*s.s, blocks, values.data(), at, max_bond, tol));
        });
  m.def("rank_split",
        [](const PyCategory &c, const py::list &sectors,
           py::array_t<double, py::array::c_style | py::array::forcecast> values, const std::vector<int64_t> &at,
           int64_t max_bond, double tol) {
          if (values.size() < (at.empty() ? 0 : at.back())) throw std::invalid_argument("rank_split: too few values");
          std::vector<Sector> ss;
          for (const auto &h : sectors) ss.push_back(to_sector(h));
          return ranking_tuple(rank_split(*c.cat, ss, values.data(), at, max_bond, tol));
        });
  py::class_<TailPass>(m, "TailPass")
      .def(py::init<const std::vector<int64_t> &, const std::vector<int64_t> &, const std::vector<int64_t> &,
                    const std::vector<int64_t> &, const std::vector<int64_t> &, const std::vector<int64_t> &,
                    const std::vector<int64_t> &>(),
           py::arg("g"), py::arg("n"), py::arg("ld"), py::arg("pq_at"), py::arg("block"), py::arg("first"),
           py::arg("tail"))
      .def_property_readonly("work_doubles", &TailPass::work_doubles)
      .def_property_readonly("gram_doubles", &TailPass::gram_doubles)
      .def_property_readonly("norm_doubles", &TailPass::norm_doubles)
      .def_property_readonly("stage_bytes", &TailPass::stage_bytes)
      .def_property_readonly("orders", &TailPass::orders)
      .def_property_readonly("gram_offsets", &TailPass::gram_offsets)
      .def("fill",
           [](TailPass &t, uintptr_t pq, int64_t pq_bytes, uintptr_t work, uintptr_t grams, uintptr_t vectors,
              uintptr_t norms, uintptr_t stage_device, uintptr_t stage_host,
              py::array_t<double, py::array::c_style | py::array::forcecast> values,
              py::array_t<int64_t, py::array::c_style | py::array::forcecast> values_at) {
             t.fill(pq, pq_bytes, work, grams, vectors, norms, stage_device, reinterpret_cast<void *>(stage_host),
                    values.data(), values_at.data());
           })
      .def("tables", [](const TailPass &t) {
        py::list out;
        for (const TailPass::Table &x : t.tables()) out.append(py::make_tuple(x.kind, x.count, x.items, x.at));
        return out;
      });

  m.def("plan_einsum", [](const std::string &spec, const std::vector<PyStructure> &ops, py::object charge) {
    std::vector<StructurePtr> ss;
    for (const PyStructure &s : ops) ss.push_back(s.s);
    if (charge.is_none()) return PyProgram{plan_einsum(spec, ss, nullptr)};
    const Sector q = to_sector(charge);
    return PyProgram{plan_einsum(spec, ss, &q)};
  });
  m.def("plan_contract", [](const PyStructure &l, const PyStructure &r, const std::vector<int64_t> &ll,
                            const std::vector<int64_t> &rl, py::object charge) {
    if (charge.is_none()) return PyProgram{plan_contract(l.s, r.s, ll, rl, nullptr)};
    const Sector q = to_sector(charge);
    return PyProgram{plan_contract(l.s, r.s, ll, rl, &q)};
  });
  m.def("plan_place", [](const PyStructure &s, const std::vector<int64_t> &above, py::object below) {
    const std::vector<int64_t> b = below.is_none() ? complement(s.s->nlegs(), above) : below.cast<std::vector<int64_t>>();
    return PyProgram{plan_place(s.s, above, b)};
  });
  m.def("plan_conj", [](const PyStructure &s) { return PyProgram{plan_conj(s.s)}; });
  m.def("clear_caches", [] {
    clear_program_cache();
    clear_move_cache();
    clear_engine_caches();
  });
}
