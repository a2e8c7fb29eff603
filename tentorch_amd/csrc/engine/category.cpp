#include "category.h"

#include <cmath>

namespace tt {

std::string render(const Sector &s) {
  std::string out = "(";
  for (int i = 0; i < s.n; ++i) {
    if (i) out += ", ";
    out += std::to_string(s.v[i]);
  }
  if (s.n == 1) out += ",";
  return out + ")";
}

const std::vector<Outcome> &Category::outcomes(const Sector &a, const Sector &b) const {
  const bool small = a.n == 1 && b.n == 1 && a.v[0] >= -kSmall && a.v[0] < kSmall && b.v[0] >= -kSmall && b.v[0] < kSmall;
  size_t slot = 0;
  if (small) {
    if (small_.empty()) small_.assign(static_cast<size_t>(4 * kSmall * kSmall), nullptr);
    slot = static_cast<size_t>((a.v[0] + kSmall) * 2 * kSmall + (b.v[0] + kSmall));
    if (small_[slot]) return *small_[slot];
  }
  const auto key = std::make_pair(a, b);
  auto found = outcomes_.find(key);
  if (found == outcomes_.end()) found = outcomes_.emplace(key, compute_outcomes(a, b)).first;
  if (small) small_[slot] = &found->second;  // node-based map: the element never moves
  return found->second;
}

int64_t Category::n_symbol(const Sector &a, const Sector &b, const Sector &c) const {
  for (const Outcome &o : outcomes(a, b))
    if (o.sector == c) return o.multiplicity;
  return 0;
}

// Row addressing of F (cat.cpp:62-78).
std::vector<FIndex> Category::f_rows(const Sector &a, const Sector &b, const Sector &c, const Sector &d) const {
  std::vector<FIndex> rows;
  for (const Outcome &e : outcomes(a, b)) {
    const int64_t copies = n_symbol(e.sector, c, d);
    for (int64_t inner = 0; inner < e.multiplicity; ++inner)
      for (int64_t outer = 0; outer < copies; ++outer) rows.push_back({e.sector, inner, outer});
  }
  return rows;
}

// Column addressing of F (cat.cpp:80-96).
std::vector<FIndex> Category::f_columns(const Sector &a, const Sector &b, const Sector &c,
                                        const Sector &d) const {
  std::vector<FIndex> cols;
  for (const Outcome &f : outcomes(b, c)) {
    const int64_t copies = n_symbol(a, f.sector, d);
    for (int64_t inner = 0; inner < f.multiplicity; ++inner)
      for (int64_t outer = 0; outer < copies; ++outer) cols.push_back({f.sector, inner, outer});
  }
  return cols;
}

std::vector<std::pair<Sector, int64_t>> Category::r_rows(const Sector &b, const Sector &c) const {
  std::vector<std::pair<Sector, int64_t>> rows;
  for (const Outcome &e : outcomes(b, c))
    for (int64_t copy = 0; copy < e.multiplicity; ++copy) rows.emplace_back(e.sector, copy);
  return rows;
}

const Matrix &Category::F(const Sector &a, const Sector &b, const Sector &c, const Sector &d) const {
  const auto key = std::make_tuple(a, b, c, d);
  auto found = f_.find(key);
  if (found == f_.end()) found = f_.emplace(key, compute_F(a, b, c, d)).first;
  return found->second;
}

const Matrix &Category::R(const Sector &b, const Sector &c) const {
  const auto key = std::make_pair(b, c);
  auto found = r_.find(key);
  if (found == r_.end()) found = r_.emplace(key, compute_R(b, c)).first;
  return found->second;
}

void Category::clear_cache() const {
  small_.clear();
  outcomes_.clear();
  f_.clear();
  r_.clear();
}

namespace {
Matrix ones(int64_t side) {
  Matrix m;
  m.rows = m.cols = side;
  m.data.assign(static_cast<size_t>(side * side), cplx(1.0, 0.0));
  return m;
}
int64_t residue(int64_t x, int64_t n) {
  const int64_t r = x % n;
  return r < 0 ? r + n : r;
}
}  // namespace

std::vector<Outcome> AbelianGroup::compute_outcomes(const Sector &a, const Sector &b) const {
  return {{combine(a, b), 1}};
}

// CG = 1 (abelian.cpp:49-54), so the overlap of the two bracketings is 1 when a+b+c = d.
Matrix AbelianGroup::compute_F(const Sector &a, const Sector &b, const Sector &c, const Sector &d) const {
  return ones(combine(combine(a, b), c) == d ? 1 : 0);
}

Matrix AbelianGroup::compute_R(const Sector &, const Sector &) const { return ones(1); }

ZNSymmetry::ZNSymmetry(int64_t modulus)
    : AbelianGroup(modulus > 0 ? "Z_" + std::to_string(modulus)
                               : throw std::invalid_argument("The modulus of Z_N must be positive, not " +
                                                             std::to_string(modulus) + ".")),
      modulus_(modulus) {}

Sector ZNSymmetry::combine(const Sector &a, const Sector &b) const {
  return Sector{residue(a.at(0) + b.at(0), modulus_)};
}
Sector ZNSymmetry::inverse(const Sector &a) const { return Sector{residue(-a.at(0), modulus_)}; }

std::vector<Outcome> FermionParity::compute_outcomes(const Sector &a, const Sector &b) const {
  return {{Sector{residue(a.at(0) + b.at(0), 2)}, 1}};
}
Matrix FermionParity::compute_F(const Sector &a, const Sector &b, const Sector &c, const Sector &d) const {
  return ones(residue(a.at(0) + b.at(0) + c.at(0), 2) == d.at(0) ? 1 : 0);
}
Matrix FermionParity::compute_R(const Sector &b, const Sector &c) const {
  Matrix m = ones(1);
  if (b.at(0) == 1 && c.at(0) == 1) m.data[0] = cplx(-1.0, 0.0);
  return m;
}

// F = <left tree | right tree> / d_d over CG tensors (group.cpp:44-83).
Matrix CGGroup::compute_F(const Sector &a, const Sector &b, const Sector &c, const Sector &d) const {
  const auto rows = f_rows(a, b, c, d);
  const auto cols = f_columns(a, b, c, d);
  Matrix out;
  out.rows = static_cast<int64_t>(rows.size());
  out.cols = static_cast<int64_t>(cols.size());
  out.data.assign(static_cast<size_t>(out.rows * out.cols), cplx(0.0, 0.0));
  if (!out.rows || !out.cols) return out;
  int64_t da, db, dc, dd, x1, x2, x3;
  auto vec_left = [&](const FIndex &r) {
    int64_t ea, eb, ee, f1, f2, f3;
    std::vector<double> ab = clebsch_gordan(a, b, r.sector, r.inner, &ea, &eb, &ee);
    std::vector<double> ec = clebsch_gordan(r.sector, c, d, r.outer, &f1, &f2, &f3);
    da = ea; db = eb; dc = f2; dd = f3;
    std::vector<double> v(static_cast<size_t>(da * db * dc * dd), 0.0);
    for (int64_t i = 0; i < da; ++i)
      for (int64_t j = 0; j < db; ++j)
        for (int64_t e = 0; e < ee; ++e) {
          const double x = ab[static_cast<size_t>((i * db + j) * ee + e)];
          if (x == 0.0) continue;
          for (int64_t k = 0; k < dc; ++k)
            for (int64_t l = 0; l < dd; ++l)
              v[static_cast<size_t>(((i * db + j) * dc + k) * dd + l)] +=
                  x * ec[static_cast<size_t>((e * dc + k) * dd + l)];
        }
    return v;
  };
  auto vec_right = [&](const FIndex &col) {
    int64_t eb, ec_, ef, g1, g2, g3;
    std::vector<double> bc = clebsch_gordan(b, c, col.sector, col.inner, &eb, &ec_, &ef);
    std::vector<double> af = clebsch_gordan(a, col.sector, d, col.outer, &g1, &g2, &g3);
    const int64_t na = g1, nb = eb, nc = ec_, nd = g3;
    std::vector<double> v(static_cast<size_t>(na * nb * nc * nd), 0.0);
    for (int64_t j = 0; j < nb; ++j)
      for (int64_t k = 0; k < nc; ++k)
        for (int64_t f = 0; f < ef; ++f) {
          const double x = bc[static_cast<size_t>((j * nc + k) * ef + f)];
          if (x == 0.0) continue;
          for (int64_t i = 0; i < na; ++i)
            for (int64_t l = 0; l < nd; ++l)
              v[static_cast<size_t>(((i * nb + j) * nc + k) * nd + l)] +=
                  x * af[static_cast<size_t>((i * ef + f) * nd + l)];
        }
    return v;
  };
  std::vector<std::vector<double>> left, right;
  for (const FIndex &r : rows) left.push_back(vec_left(r));
  for (const FIndex &col : cols) right.push_back(vec_right(col));
  (void)x1; (void)x2; (void)x3;
  const double scale = q_dim(d);
  for (int64_t r = 0; r < out.rows; ++r)
    for (int64_t col = 0; col < out.cols; ++col) {
      double s = 0.0;
      const auto &L = left[static_cast<size_t>(r)];
      const auto &Rv = right[static_cast<size_t>(col)];
      for (size_t i = 0; i < L.size(); ++i) s += L[i] * Rv[i];
      out(r, col) = cplx(s / scale, 0.0);
    }
  return out;
}

// R block per outcome e: <straight | swapped> / d_e (group.cpp:93-130).
Matrix CGGroup::compute_R(const Sector &b, const Sector &c) const {
  const auto rows = r_rows(b, c);
  Matrix out;
  out.rows = out.cols = static_cast<int64_t>(rows.size());
  out.data.assign(static_cast<size_t>(out.rows * out.cols), cplx(0.0, 0.0));
  for (int64_t start = 0; start < out.rows;) {
    const Sector e = rows[static_cast<size_t>(start)].first;
    int64_t end = start;
    while (end < out.rows && rows[static_cast<size_t>(end)].first == e) ++end;
    const int64_t copies = end - start;
    if (n_symbol(c, b, e) != copies)
      throw std::runtime_error("R_symbol needs N^e_bc == N^e_cb for every outcome e of b (x) c");
    for (int64_t i = 0; i < copies; ++i)
      for (int64_t j = 0; j < copies; ++j) {
        int64_t nb, nc, ne, mc, mb, me;
        std::vector<double> s = clebsch_gordan(b, c, e, i, &nb, &nc, &ne);
        std::vector<double> w = clebsch_gordan(c, b, e, j, &mc, &mb, &me);
        double acc = 0.0;
        for (int64_t x = 0; x < nb; ++x)
          for (int64_t y = 0; y < nc; ++y)
            for (int64_t z = 0; z < ne; ++z)
              acc += s[static_cast<size_t>((x * nc + y) * ne + z)] * w[static_cast<size_t>((y * nb + x) * ne + z)];
        out(start + i, start + j) = cplx(acc / q_dim(e), 0.0);
      }
    start = end;
  }
  return out;
}

std::vector<Outcome> SU2Symmetry::compute_outcomes(const Sector &a, const Sector &b) const {
  std::vector<Outcome> out;
  for (int64_t jj = std::llabs(a.at(0) - b.at(0)); jj <= a.at(0) + b.at(0); jj += 2) out.push_back({Sector{jj}, 1});
  return out;
}

namespace {
// CG tables of SU(2) by lowering highest weights (su2.cpp, `su2::coefficients`): the level-M state of
// total spin J is lowered from level M+2, and the state new at a level is the part of the
// highest-weight direction (m1 = j1) orthogonal to the lowered ones, with a positive m1 = j1 entry.
std::map<int64_t, std::vector<double>> su2_tables(int64_t j1, int64_t j2) {
  const int64_t lo = std::llabs(j1 - j2), hi = j1 + j2, n1 = j1 + 1;
  auto index = [&](int64_t J, int64_t k1, int64_t k2, int64_t k3) { return ((k1 * (j2 + 1) + k2) * (J + 1) + k3); };
  std::map<int64_t, std::vector<double>> out;
  out[hi] = std::vector<double>(static_cast<size_t>((j1 + 1) * (j2 + 1) * (hi + 1)), 0.0);
  out[hi][static_cast<size_t>(index(hi, j1, j2, hi))] = 1.0;
  std::map<int64_t, std::vector<double>> previous;
  previous[hi] = std::vector<double>(static_cast<size_t>(n1), 0.0);
  previous[hi][static_cast<size_t>(j1)] = 1.0;
  auto lower = [&](const std::vector<double> &vec, int64_t MM) {
    std::vector<double> res(static_cast<size_t>(n1), 0.0);
    for (int64_t k1 = 0; k1 < n1; ++k1) {
      const int64_t m1 = 2 * k1 - j1, m2 = MM + 2 - m1;
      if (std::llabs(m2) > j2) continue;
      if (std::llabs(m2 - 2) <= j2) res[static_cast<size_t>(k1)] += vec[static_cast<size_t>(k1)] * std::sqrt(double((j2 + m2) * (j2 - m2 + 2))) / 2;
      if (k1 > 0 && std::llabs(m1 - 2) <= j1)
        res[static_cast<size_t>(k1 - 1)] += vec[static_cast<size_t>(k1)] * std::sqrt(double((j1 + m1) * (j1 - m1 + 2))) / 2;
    }
    return res;
  };
  for (int64_t MM = hi - 2; MM >= -hi; MM -= 2) {
    std::map<int64_t, std::vector<double>> level;
    for (auto it = previous.rbegin(); it != previous.rend(); ++it) {
      if (it->first < std::llabs(MM)) continue;
      const double scale = std::sqrt(double((it->first + MM + 2) * (it->first - MM))) / 2;
      std::vector<double> vec = lower(it->second, MM);
      for (double &x : vec) x /= scale;
      level[it->first] = vec;
    }
    if (MM >= lo && MM >= 0) {
      std::vector<double> vec(static_cast<size_t>(n1), 0.0);
      vec[static_cast<size_t>(j1)] = 1.0;
      for (auto it = level.rbegin(); it != level.rend(); ++it) {
        double dot = 0.0;
        for (int64_t k = 0; k < n1; ++k) dot += it->second[static_cast<size_t>(k)] * vec[static_cast<size_t>(k)];
        for (int64_t k = 0; k < n1; ++k) vec[static_cast<size_t>(k)] -= dot * it->second[static_cast<size_t>(k)];
      }
      double norm = 0.0;
      for (double x : vec) norm += x * x;
      norm = std::sqrt(norm);
      if (norm <= 1e-12) throw std::runtime_error("degenerate SU(2) orthogonalisation");
      for (double &x : vec) x /= norm;
      level[MM] = vec;
    }
    previous = level;
    for (const auto &[JJ, vec] : level) {
      auto &t = out[JJ];
      if (t.empty()) t.assign(static_cast<size_t>((j1 + 1) * (j2 + 1) * (JJ + 1)), 0.0);
      for (int64_t k1 = 0; k1 < n1; ++k1) {
        const int64_t m1 = 2 * k1 - j1, m2 = MM - m1;
        if (std::llabs(m2) <= j2 && vec[static_cast<size_t>(k1)] != 0.0)
          t[static_cast<size_t>(index(JJ, k1, (m2 + j2) / 2, (MM + JJ) / 2))] += vec[static_cast<size_t>(k1)];
      }
    }
  }
  return out;
}
}  // namespace

std::vector<double> SU2Symmetry::clebsch_gordan(const Sector &a, const Sector &b, const Sector &c, int64_t copy,
                                                int64_t *da, int64_t *db, int64_t *dc) const {
  if (copy != 0) throw std::runtime_error("SU(2) couples each pair of spins once");
  const auto key = std::make_pair(a.at(0), b.at(0));
  auto found = tables_.find(key);
  if (found == tables_.end()) found = tables_.emplace(key, su2_tables(a.at(0), b.at(0))).first;
  auto t = found->second.find(c.at(0));
  if (t == found->second.end()) throw std::runtime_error("the spins do not fuse to the requested one");
  *da = a.at(0) + 1;
  *db = b.at(0) + 1;
  *dc = c.at(0) + 1;
  return t->second;
}

void SU2Symmetry::clear_cache() const {
  tables_.clear();
  Category::clear_cache();
}

std::vector<Outcome> FibonacciCategory::compute_outcomes(const Sector &a, const Sector &b) const {
  if (a.at(0) == 0) return {{b, 1}};
  if (b.at(0) == 0) return {{a, 1}};
  return {{Sector{0}, 1}, {Sector{1}, 1}};
}

double FibonacciCategory::q_dim(const Sector &a) const { return a.at(0) == 1 ? (1.0 + std::sqrt(5.0)) / 2.0 : 1.0; }

// The golden matrix for tau tau tau -> tau, ones elsewhere (fibonacci.cpp:30-53).
Matrix FibonacciCategory::compute_F(const Sector &a, const Sector &b, const Sector &c, const Sector &d) const {
  const auto rows = f_rows(a, b, c, d);
  const auto cols = f_columns(a, b, c, d);
  Matrix m;
  m.rows = static_cast<int64_t>(rows.size());
  m.cols = static_cast<int64_t>(cols.size());
  m.data.assign(static_cast<size_t>(m.rows * m.cols), cplx(1.0, 0.0));
  if (a.at(0) == 1 && b.at(0) == 1 && c.at(0) == 1 && d.at(0) == 1 && m.rows && m.cols) {
    const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
    const double golden[2][2] = {{1.0 / phi, 1.0 / std::sqrt(phi)}, {1.0 / std::sqrt(phi), -1.0 / phi}};
    for (int64_t i = 0; i < m.rows; ++i)
      for (int64_t j = 0; j < m.cols; ++j)
        m(i, j) = cplx(golden[rows[static_cast<size_t>(i)].sector.at(0)][cols[static_cast<size_t>(j)].sector.at(0)], 0.0);
  }
  return m;
}

// R phases e^{-4 pi i/5} (to 1) and e^{3 pi i/5} (to tau) for tau x tau (fibonacci.cpp:55-69).
Matrix FibonacciCategory::compute_R(const Sector &b, const Sector &c) const {
  const auto rows = r_rows(b, c);
  Matrix m;
  m.rows = m.cols = static_cast<int64_t>(rows.size());
  m.data.assign(static_cast<size_t>(m.rows * m.cols), cplx(0.0, 0.0));
  const double pi = std::acos(-1.0);
  for (int64_t i = 0; i < m.rows; ++i) {
    cplx phase(1.0, 0.0);
    if (b.at(0) == 1 && c.at(0) == 1)
      phase = rows[static_cast<size_t>(i)].first.at(0) == 0 ? std::polar(1.0, -4.0 * pi / 5.0) : std::polar(1.0, 3.0 * pi / 5.0);
    m(i, i) = phase;
  }
  return m;
}

CategoryPtr no_symmetry() {
  static const CategoryPtr c = std::make_shared<NoSymmetry>();
  return c;
}
CategoryPtr u1_symmetry() {
  static const CategoryPtr c = std::make_shared<U1Symmetry>();
  return c;
}
CategoryPtr z_n_symmetry(int64_t modulus) {
  static std::map<int64_t, CategoryPtr> kept;
  auto &c = kept[modulus];
  if (!c) c = std::make_shared<ZNSymmetry>(modulus);
  return c;
}
CategoryPtr fermion_parity() {
  static const CategoryPtr c = std::make_shared<FermionParity>();
  return c;
}
CategoryPtr su2_symmetry() {
  static const CategoryPtr c = std::make_shared<SU2Symmetry>();
  return c;
}
CategoryPtr fibonacci_category() {
  static const CategoryPtr c = std::make_shared<FibonacciCategory>();
  return c;
}

}  // namespace tt
