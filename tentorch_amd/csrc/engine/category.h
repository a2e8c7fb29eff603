// Fusion categories of the host engine: sector labels, fusion outcomes, and the F and R matrices
// that recouple fusion trees.  Interface contract follows the reference's `FusionCategory`
// (tentorch/symmetry/cat/cat.h:216-429): outcome order fixes F/R row order, F rows are the left
// bracketing ((a b) c) -> d addressed (e, inner, outer) and columns the right bracketing
// (a (b c)) -> d addressed (f, inner, outer) (cat.cpp:62-96), R rows/columns are (e, copy)
// (cat.cpp:98-109).  Symbols are held on the host as dense complex matrices.
#pragma once

#include <array>
#include <complex>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

namespace tt {

using cplx = std::complex<double>;

// A sector: up to four integers held inline (the reference's `SectorOf<4>`, cat.h:41-128).
struct Sector {
  std::array<int64_t, 4> v{};
  int32_t n = 0;

  Sector() = default;
  Sector(std::initializer_list<int64_t> values) {
    if (values.size() > 4) throw std::length_error("a sector holds at most four entries");
    for (int64_t x : values) v[n++] = x;
  }
  explicit Sector(const std::vector<int64_t> &values) {
    if (values.size() > 4) throw std::length_error("a sector holds at most four entries");
    for (int64_t x : values) v[n++] = x;
  }
  int64_t operator[](int i) const { return v[i]; }
  int64_t at(int i) const {
    if (i < 0 || i >= n) throw std::out_of_range("sector entry out of range");
    return v[i];
  }
  std::vector<int64_t> to_vector() const { return std::vector<int64_t>(v.begin(), v.begin() + n); }
  bool operator==(const Sector &o) const {
    if (n != o.n) return false;
    for (int i = 0; i < n; ++i)
      if (v[i] != o.v[i]) return false;
    return true;
  }
  bool operator!=(const Sector &o) const { return !(*this == o); }
  // Lexicographic, shorter first on a common prefix (cat.h:118-128).
  bool operator<(const Sector &o) const {
    const int common = n < o.n ? n : o.n;
    for (int i = 0; i < common; ++i)
      if (v[i] != o.v[i]) return v[i] < o.v[i];
    return n < o.n;
  }
};

std::string render(const Sector &s);

struct SectorHash {
  size_t operator()(const Sector &s) const {
    uint64_t h = 0xcbf29ce484222325ULL ^ static_cast<uint64_t>(s.n);
    for (int i = 0; i < s.n; ++i) h = (h ^ static_cast<uint64_t>(s.v[i])) * 0x100000001b3ULL;
    return static_cast<size_t>(h);
  }
};

struct Outcome {
  Sector sector;
  int64_t multiplicity;
};

struct FIndex {
  Sector sector;
  int64_t inner;
  int64_t outer;
};

// A dense row-major complex matrix.
struct Matrix {
  int64_t rows = 0, cols = 0;
  std::vector<cplx> data;
  cplx operator()(int64_t r, int64_t c) const { return data[static_cast<size_t>(r * cols + c)]; }
  cplx &operator()(int64_t r, int64_t c) { return data[static_cast<size_t>(r * cols + c)]; }
};

class Category {
 public:
  explicit Category(std::string name) : name_(std::move(name)) {}
  virtual ~Category() = default;

  const std::string &name() const { return name_; }
  virtual bool abelian() const { return false; }
  // Whether every R symbol is the 1x1 one, so an exchange of two legs is a plain relabelling.
  virtual bool trivial_braiding() const { return false; }
  // Whether every F and R entry is real (the scalar type the reference's `dtype()` reports).
  virtual bool real_symbols() const { return true; }

  const std::vector<Outcome> &outcomes(const Sector &a, const Sector &b) const;
  int64_t n_symbol(const Sector &a, const Sector &b, const Sector &c) const;
  std::vector<FIndex> f_rows(const Sector &a, const Sector &b, const Sector &c, const Sector &d) const;
  std::vector<FIndex> f_columns(const Sector &a, const Sector &b, const Sector &c, const Sector &d) const;
  std::vector<std::pair<Sector, int64_t>> r_rows(const Sector &b, const Sector &c) const;

  const Matrix &F(const Sector &a, const Sector &b, const Sector &c, const Sector &d) const;
  const Matrix &R(const Sector &b, const Sector &c) const;

  virtual double q_dim(const Sector &a) const = 0;
  virtual Sector dual(const Sector &a) const = 0;
  virtual Sector trivial() const { return Sector{0}; }
  virtual double frobenius_schur_sign(const Sector &) const { return 1.0; }
  virtual void clear_cache() const;

 protected:
  virtual std::vector<Outcome> compute_outcomes(const Sector &a, const Sector &b) const = 0;
  virtual Matrix compute_F(const Sector &a, const Sector &b, const Sector &c, const Sector &d) const = 0;
  virtual Matrix compute_R(const Sector &b, const Sector &c) const = 0;

 private:
  struct PairHash {
    size_t operator()(const std::pair<Sector, Sector> &p) const {
      return SectorHash{}(p.first) * 1000003u ^ SectorHash{}(p.second);
    }
  };
  std::string name_;
  mutable std::unordered_map<std::pair<Sector, Sector>, std::vector<Outcome>, PairHash> outcomes_;
  // Direct-mapped view of outcomes_ for one-integer sectors in [-kSmall, kSmall): the fusion tables
  // of a contraction are looked up hundreds of thousands of times while it is planned.
  static constexpr int64_t kSmall = 64;
  mutable std::vector<const std::vector<Outcome> *> small_;
  mutable std::map<std::tuple<Sector, Sector, Sector, Sector>, Matrix> f_;
  mutable std::map<std::pair<Sector, Sector>, Matrix> r_;
};

// Abelian groups (tentorch/symmetry/cat/abelian/abelian.cpp): one outcome, CG = 1, so F and R are
// the 1x1 identity whenever they exist.
class AbelianGroup : public Category {
 public:
  using Category::Category;
  bool abelian() const override { return true; }
  bool trivial_braiding() const override { return true; }
  double q_dim(const Sector &) const override { return 1.0; }
  Sector dual(const Sector &a) const override { return inverse(a); }
  virtual Sector combine(const Sector &a, const Sector &b) const = 0;
  virtual Sector inverse(const Sector &a) const = 0;

 protected:
  std::vector<Outcome> compute_outcomes(const Sector &a, const Sector &b) const override;
  Matrix compute_F(const Sector &a, const Sector &b, const Sector &c, const Sector &d) const override;
  Matrix compute_R(const Sector &b, const Sector &c) const override;
};

class NoSymmetry : public AbelianGroup {
 public:
  NoSymmetry() : AbelianGroup("NoSymmetry") {}
  Sector combine(const Sector &, const Sector &) const override { return Sector{0}; }
  Sector inverse(const Sector &) const override { return Sector{0}; }
};

class U1Symmetry : public AbelianGroup {
 public:
  U1Symmetry() : AbelianGroup("U(1)") {}
  Sector combine(const Sector &a, const Sector &b) const override { return Sector{a.at(0) + b.at(0)}; }
  Sector inverse(const Sector &a) const override { return Sector{-a.at(0)}; }
};

class ZNSymmetry : public AbelianGroup {
 public:
  explicit ZNSymmetry(int64_t modulus);
  int64_t modulus() const { return modulus_; }
  Sector combine(const Sector &a, const Sector &b) const override;
  Sector inverse(const Sector &a) const override;

 private:
  int64_t modulus_;
};

// Fermion parity (abelian.cpp:83-115): Koszul sign -1 for odd x odd.
class FermionParity : public Category {
 public:
  FermionParity() : Category("FermionParity") {}
  bool abelian() const override { return true; }
  double q_dim(const Sector &) const override { return 1.0; }
  Sector dual(const Sector &a) const override { return a; }
  double frobenius_schur_sign(const Sector &a) const override { return a.at(0) == 1 ? -1.0 : 1.0; }

 protected:
  std::vector<Outcome> compute_outcomes(const Sector &a, const Sector &b) const override;
  Matrix compute_F(const Sector &a, const Sector &b, const Sector &c, const Sector &d) const override;
  Matrix compute_R(const Sector &b, const Sector &c) const override;
};

// A category realised by group representations: F and R are overlaps of Clebsch-Gordan tensors
// divided by the quantum dimension of the coupled sector (tentorch/symmetry/cat/group.cpp:44-130).
class CGGroup : public Category {
 public:
  using Category::Category;
  // CG tensor of one vertex, shape (dim a, dim b, dim c), row-major, real.
  virtual std::vector<double> clebsch_gordan(const Sector &a, const Sector &b, const Sector &c, int64_t copy,
                                             int64_t *da, int64_t *db, int64_t *dc) const = 0;

 protected:
  Matrix compute_F(const Sector &a, const Sector &b, const Sector &c, const Sector &d) const override;
  Matrix compute_R(const Sector &b, const Sector &c) const override;
};

// SU(2), sectors labelled by twice the spin (tentorch/symmetry/cat/su2/su2.cpp).
class SU2Symmetry : public CGGroup {
 public:
  SU2Symmetry() : CGGroup("SU(2)") {}
  double q_dim(const Sector &a) const override { return static_cast<double>(a.at(0) + 1); }
  Sector dual(const Sector &a) const override { return a; }
  double frobenius_schur_sign(const Sector &a) const override { return a.at(0) % 2 == 0 ? 1.0 : -1.0; }
  std::vector<double> clebsch_gordan(const Sector &a, const Sector &b, const Sector &c, int64_t copy,
                                     int64_t *da, int64_t *db, int64_t *dc) const override;
  void clear_cache() const override;

 protected:
  std::vector<Outcome> compute_outcomes(const Sector &a, const Sector &b) const override;

 private:
  mutable std::map<std::pair<int64_t, int64_t>, std::map<int64_t, std::vector<double>>> tables_;
};

// The Fibonacci category (tentorch/symmetry/cat/fibonacci/fibonacci.cpp): closed-form F and R.
class FibonacciCategory : public Category {
 public:
  FibonacciCategory() : Category("FibonacciCategory") {}
  bool real_symbols() const override { return false; }
  double q_dim(const Sector &a) const override;
  Sector dual(const Sector &a) const override { return a; }

 protected:
  std::vector<Outcome> compute_outcomes(const Sector &a, const Sector &b) const override;
  Matrix compute_F(const Sector &a, const Sector &b, const Sector &c, const Sector &d) const override;
  Matrix compute_R(const Sector &b, const Sector &c) const override;
};

using CategoryPtr = std::shared_ptr<const Category>;

CategoryPtr no_symmetry();
CategoryPtr u1_symmetry();
CategoryPtr z_n_symmetry(int64_t modulus);
CategoryPtr fermion_parity();
CategoryPtr su2_symmetry();
CategoryPtr fibonacci_category();

}  // namespace tt
