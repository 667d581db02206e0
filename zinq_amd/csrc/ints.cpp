// Host Gaussian integral engine: overlap, kinetic, nuclear attraction and electron repulsion
// integrals over contracted Cartesian Gaussians with pure (real solid harmonic) shells for l >= 2.
//
// It stands in for libint2, which the reference calls through src/libint.cpp; the exported ABI is the
// one of src/libint.h (see include/zinq_ints.h).  Method: McMurchie-Davidson Hermite expansion
// (McMurchie & Davidson, J. Comput. Phys. 26, 218 (1978); Helgaker, Jorgensen & Olsen, "Molecular
// Electronic-Structure Theory", ch. 9), Boys function from a Taylor-interpolated table built from its
// convergent series, with downward recursion; libm exp/sqrt only.
//
// Shell quartets are enumerated canonically as in src/libint.cpp:115-119 (i <= j, k >= i,
// l >= (i == k ? j : k)), computed in parallel over bra shell pairs, and scattered either to the dense
// physicist layout of src/libint.cpp:143-150 or to canonical 8-fold packed storage.

#include "../../include/zinq_ints.h"

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "basis_sets.inc"

namespace {

thread_local std::string g_last_error;

constexpr int kMaxL = 4;                // up to g shells
constexpr int kMaxLTot = 4 * kMaxL;     // (gg|gg)
constexpr double kPi = 3.14159265358979323846;

int g_threads = 0;

int ncart(int l) { return (l + 1) * (l + 2) / 2; }
int ntuv(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

double dfact(int n) {  // n!! with (-1)!! = 0!! = 1
  double r = 1;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

double binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  double r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

// ---------------------------------------------------------------- Boys function F_m(T)
struct BoysTable {
  static constexpr double kTMax = 40.0;
  static constexpr double kDT = 0.05;
  static constexpr int kNT = 801;
  static constexpr int kOrder = 6;
  static constexpr int kM = kMaxLTot + kOrder + 2;
  std::vector<double> f;  // [kNT][kM]
  BoysTable() : f(size_t(kNT) * kM) {
    for (int k = 0; k < kNT; ++k) {
      const double T = k * kDT;
      for (int m = 0; m < kM; ++m) {
        // F_m(T) = exp(-T) * sum_i (2T)^i / ((2m+1)(2m+3)...(2m+2i+1)); all terms positive.
        double term = 1.0 / (2 * m + 1), sum = term;
        for (int i = 1; i < 2000; ++i) {
          term *= 2 * T / (2 * m + 2 * i + 1);
          sum += term;
          if (term < 1e-18 * sum) break;
        }
        f[size_t(k) * kM + m] = std::exp(-T) * sum;
      }
    }
  }
  // Fills F[0..mmax].
  void eval(int mmax, double T, double *F) const {
    if (T < kTMax) {
      const int k = int(T / kDT + 0.5);
      const double d = k * kDT - T;  // expand around the grid point: F_m(T) = sum_j F_{m+j}(Tk) (Tk-T)^j/j!
      const double *row = &f[size_t(k) * kM + mmax];
      double acc = 0, pw = 1, fact = 1;
      for (int j = 0; j <= kOrder; ++j) {
        acc += row[j] * pw / fact;
        pw *= d;
        fact *= j + 1;
      }
      F[mmax] = acc;
      const double ex = std::exp(-T);
      for (int m = mmax; m > 0; --m) F[m - 1] = (2 * T * F[m] + ex) / (2 * m - 1);
    } else {
      const double ex = std::exp(-T);
      F[0] = 0.5 * std::sqrt(kPi / T);
      for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - ex) / (2 * T);
    }
  }
};

const BoysTable &boys() {
  static const BoysTable t;
  return t;
}

// ---------------------------------------------------------------- Hermite index tables
struct TuvTable {
  std::vector<int> idx;         // [(kMaxLTot+1)^3] -> index, -1 if t+u+v > kMaxLTot
  std::vector<int> t, u, v;     // ordered by t+u+v, so entries of order <= L are the first ntuv(L)
  TuvTable() {
    const int D = kMaxLTot + 1;
    idx.assign(D * D * D, -1);
    for (int L = 0; L <= kMaxLTot; ++L)
      for (int a = L; a >= 0; --a)
        for (int b = L - a; b >= 0; --b) {
          const int c = L - a - b;
          idx[(a * D + b) * D + c] = int(t.size());
          t.push_back(a), u.push_back(b), v.push_back(c);
        }
  }
  int at(int a, int b, int c) const {
    const int D = kMaxLTot + 1;
    return idx[(a * D + b) * D + c];
  }
};

const TuvTable &tuv() {
  static const TuvTable t;
  return t;
}

// Cartesian components of a shell, libint order: lx descending, then ly descending.
void cart_list(int l, std::vector<int> &out) {
  out.clear();
  for (int lx = l; lx >= 0; --lx)
    for (int ly = l - lx; ly >= 0; --ly) out.push_back(lx), out.push_back(ly), out.push_back(l - lx - ly);
}

// Real solid harmonics (m = -l..l) in terms of the Cartesian monomials of cart_list, normalised so
// that each pure function has unit norm when every Cartesian component carries the normalisation of
// x^l.  Coefficient formula: Helgaker, Jorgensen & Olsen eq. 6.4.47-6.4.48.
std::vector<double> pure_matrix(int l) {
  std::vector<int> cl;
  cart_list(l, cl);
  const int nc = ncart(l);
  std::vector<double> T(size_t(2 * l + 1) * nc, 0.0);
  auto cidx = [&](int a, int b, int c) {
    for (int i = 0; i < nc; ++i)
      if (cl[3 * i] == a && cl[3 * i + 1] == b && cl[3 * i + 2] == c) return i;
    return -1;
  };
  for (int m = -l; m <= l; ++m) {
    const int am = std::abs(m);
    const int km = m < 0 ? 1 : 0;
    for (int t = 0; t <= (l - am) / 2; ++t)
      for (int u = 0; u <= t; ++u)
        for (int k = km; k <= am; k += 2) {
          const int sgn = ((t + (k - km) / 2) % 2 == 0) ? 1 : -1;
          const double c = sgn * std::pow(0.25, t) * binom(l, t) * binom(l - t, am + t) * binom(t, u) * binom(am, k);
          const int px = 2 * t + am - 2 * u - k, py = 2 * u + k, pz = l - 2 * t - am;
          T[size_t(m + l) * nc + cidx(px, py, pz)] += c;
        }
  }
  // Normalise with the one-centre monomial overlap relative to <x^l|x^l>.
  auto mono = [&](int i, int j) {
    double r = 1;
    for (int d = 0; d < 3; ++d) {
      const int s = cl[3 * i + d] + cl[3 * j + d];
      if (s % 2) return 0.0;
      r *= dfact(s - 1);
    }
    return r / dfact(2 * l - 1);
  };
  for (int m = 0; m < 2 * l + 1; ++m) {
    double nrm = 0;
    for (int i = 0; i < nc; ++i)
      for (int j = 0; j < nc; ++j) nrm += T[size_t(m) * nc + i] * T[size_t(m) * nc + j] * mono(i, j);
    for (int i = 0; i < nc; ++i) T[size_t(m) * nc + i] /= std::sqrt(nrm);
  }
  return T;
}

const std::vector<double> &pure_T(int l) {
  static const std::vector<double> tabs[kMaxL + 1] = {pure_matrix(0), pure_matrix(1), pure_matrix(2), pure_matrix(3),
                                                      pure_matrix(4)};
  return tabs[l];
}

// ---------------------------------------------------------------- basis / system
struct RawShell {
  int l;
  std::vector<double> exps, coefs;
};

struct Shell {
  int l = 0, atom = 0, bf0 = 0;
  bool pure = false;
  double O[3] = {0, 0, 0};
  std::vector<double> exps, coefs;  // coefs include primitive and contraction normalisation
  int nbf() const { return pure ? 2 * l + 1 : ncart(l); }
};

struct Basis {
  std::vector<std::vector<RawShell>> by_z;  // indexed by atomic number
};

const char *kSymbols[] = {"X",  "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si",
                          "P",  "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu",
                          "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru",
                          "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe"};
constexpr int kNumSymbols = sizeof(kSymbols) / sizeof(kSymbols[0]);

int symbol_to_z(std::string s) {
  if (s.empty()) return -1;
  s[0] = char(std::toupper(s[0]));
  for (size_t i = 1; i < s.size(); ++i) s[i] = char(std::tolower(s[i]));
  for (int z = 1; z < kNumSymbols; ++z)
    if (s == kSymbols[z]) return z;
  return -1;
}

std::string lower(std::string s) {
  for (auto &c : s) c = char(std::tolower(c));
  return s;
}

bool load_builtin(const std::string &name, Basis &b) {
  int id = -1;
  for (int i = 0; i < zq_builtin_basis_count; ++i)
    if (lower(name) == zq_builtin_basis_names[i]) id = i;
  if (id < 0) return false;
  b.by_z.assign(kNumSymbols, {});
  for (int s = 0; s < zq_builtin_shell_count; ++s) {
    const int *r = zq_builtin_shells[s];
    if (r[0] != id) continue;
    RawShell rs;
    rs.l = r[2];
    for (int p = 0; p < r[3]; ++p) {
      rs.exps.push_back(zq_builtin_prims[r[4] + p][0]);
      rs.coefs.push_back(zq_builtin_prims[r[4] + p][1]);
    }
    b.by_z[r[1]].push_back(rs);
  }
  return true;
}

// Gaussian94 basis file reader (element blocks "Sym 0", shells "L nprim scale", "****" terminator).
bool load_g94(const std::string &path, Basis &b) {
  std::ifstream f(path);
  if (!f) return false;
  b.by_z.assign(kNumSymbols, {});
  std::string line;
  std::vector<std::string> lines;
  while (std::getline(f, line)) {
    size_t a = line.find_first_not_of(" \t\r");
    if (a == std::string::npos || line[a] == '!') continue;
    size_t e = line.find_last_not_of(" \t\r");
    lines.push_back(line.substr(a, e - a + 1));
  }
  auto num = [](std::string s) {
    for (auto &c : s)
      if (c == 'D' || c == 'd') c = 'E';
    return std::strtod(s.c_str(), nullptr);
  };
  size_t i = 0;
  while (i < lines.size()) {
    std::istringstream hs(lines[i]);
    std::string sym, zero;
    hs >> sym >> zero;
    const int z = symbol_to_z(sym);
    if (z < 0 || zero != "0") {
      ++i;
      continue;
    }
    ++i;
    while (i < lines.size() && lines[i] != "****") {
      std::istringstream ss(lines[i]);
      std::string kind;
      int nprim = 0;
      ss >> kind >> nprim;
      kind = lower(kind);
      std::vector<std::vector<double>> rows;
      for (int p = 0; p < nprim && i + 1 + p < lines.size(); ++p) {
        std::istringstream rs(lines[i + 1 + p]);
        std::vector<double> row;
        std::string tok;
        while (rs >> tok) row.push_back(num(tok));
        rows.push_back(row);
      }
      i += nprim + 1;
      auto add = [&](int l, int col) {
        RawShell s;
        s.l = l;
        for (auto &r : rows) s.exps.push_back(r[0]), s.coefs.push_back(r[col]);
        b.by_z[z].push_back(s);
      };
      if (kind == "sp") {
        add(0, 1), add(1, 2);
      } else {
        const std::string letters = "spdfg";
        const size_t l = letters.find(kind[0]);
        if (l == std::string::npos || kind.size() != 1) return false;
        add(int(l), 1);
      }
    }
    ++i;
  }
  return true;
}

}  // namespace

struct SystemData {
  std::vector<int> Z;
  std::vector<double> xyz;  // bohr
  std::vector<Shell> shells;
  int nbf = 0;
  int max_l = 0;
};

namespace {

bool build_shells(SystemData &sys, const Basis &b) {
  sys.shells.clear();
  int bf = 0;
  for (size_t a = 0; a < sys.Z.size(); ++a) {
    const int z = sys.Z[a];
    if (z <= 0 || z >= int(b.by_z.size()) || b.by_z[z].empty()) {
      g_last_error = "basis has no functions for atomic number " + std::to_string(z);
      return false;
    }
    for (const RawShell &rs : b.by_z[z]) {
      if (rs.l > kMaxL) {
        g_last_error = "angular momentum above g is not supported";
        return false;
      }
      Shell s;
      s.l = rs.l;
      s.pure = rs.l >= 2;
      s.atom = int(a);
      for (int d = 0; d < 3; ++d) s.O[d] = sys.xyz[3 * a + d];
      s.exps = rs.exps;
      const int l = rs.l;
      // primitive normalisation of x^l exp(-a r^2), then renormalise the contraction
      for (size_t p = 0; p < rs.exps.size(); ++p) {
        const double e = rs.exps[p];
        s.coefs.push_back(rs.coefs[p] * std::pow(2 * e / kPi, 0.75) * std::pow(4 * e, 0.5 * l) / std::sqrt(dfact(2 * l - 1)));
      }
      double nrm = 0;
      for (size_t p = 0; p < s.exps.size(); ++p)
        for (size_t q = 0; q < s.exps.size(); ++q) {
          const double g = s.exps[p] + s.exps[q];
          nrm += s.coefs[p] * s.coefs[q] * dfact(2 * l - 1) / std::pow(2 * g, l) * std::pow(kPi / g, 1.5);
        }
      for (auto &c : s.coefs) c /= std::sqrt(nrm);
      s.bf0 = bf;
      bf += s.nbf();
      sys.max_l = std::max(sys.max_l, l);
      sys.shells.push_back(s);
    }
  }
  sys.nbf = bf;
  return true;
}

bool load_basis(const std::string &basis, Basis &b) {
  const std::string pre = "builtin:";
  if (basis.compare(0, pre.size(), pre) == 0) {
    if (!load_builtin(basis.substr(pre.size()), b)) {
      g_last_error = "unknown builtin basis " + basis;
      return false;
    }
    return true;
  }
  if (!load_g94(basis, b)) {
    g_last_error = "cannot read basis file " + basis;
    return false;
  }
  return true;
}

// ---------------------------------------------------------------- Hermite expansion coefficients
// E[i][j][t] for i <= la, j <= lb, t <= i + j (1D, without the exp(-mu X^2) factor).
struct E1D {
  int la, lb, nt;
  std::vector<double> e;
  double at(int i, int j, int t) const {
    if (t < 0 || t > i + j) return 0.0;
    return e[(size_t(i) * (lb + 1) + j) * nt + t];
  }
  double &ref(int i, int j, int t) { return e[(size_t(i) * (lb + 1) + j) * nt + t]; }
  void build(int la_, int lb_, double p, double XPA, double XPB) {
    la = la_, lb = lb_, nt = la + lb + 1;
    e.assign(size_t(la + 1) * (lb + 1) * nt, 0.0);
    const double h = 0.5 / p;
    ref(0, 0, 0) = 1.0;
    for (int i = 0; i < la; ++i)
      for (int t = 0; t <= i + 1; ++t) ref(i + 1, 0, t) = h * at(i, 0, t - 1) + XPA * at(i, 0, t) + (t + 1) * at(i, 0, t + 1);
    for (int i = 0; i <= la; ++i)
      for (int j = 0; j < lb; ++j)
        for (int t = 0; t <= i + j + 1; ++t)
          ref(i, j + 1, t) = h * at(i, j, t - 1) + XPB * at(i, j, t) + (t + 1) * at(i, j, t + 1);
  }
};

// Hermite Coulomb integrals R_tuv(alpha, X, Y, Z) for t+u+v <= L, stored in TuvTable order.
void hermite_R(int L, double alpha, double X, double Y, double Z, double pref_boys, double *R, double *work) {
  const TuvTable &tt = tuv();
  double F[kMaxLTot + 1];
  boys().eval(L, alpha * (X * X + Y * Y + Z * Z), F);
  const int nt = ntuv(L);
  // work: (L+1) levels of nt entries; level n holds R^{(n)} for order <= L - n.
  double m2a = 1;
  for (int n = 0; n <= L; ++n) {
    work[size_t(n) * nt] = pref_boys * m2a * F[n];
    m2a *= -2 * alpha;
  }
  for (int n = L - 1; n >= 0; --n) {
    double *cur = work + size_t(n) * nt;
    const double *up = work + size_t(n + 1) * nt;
    const int nmax = ntuv(L - n);
    for (int k = 1; k < nmax; ++k) {
      const int t = tt.t[k], u = tt.u[k], v = tt.v[k];
      double r;
      if (t > 0) {
        r = X * up[tt.at(t - 1, u, v)];
        if (t > 1) r += (t - 1) * up[tt.at(t - 2, u, v)];
      } else if (u > 0) {
        r = Y * up[tt.at(t, u - 1, v)];
        if (u > 1) r += (u - 1) * up[tt.at(t, u - 2, v)];
      } else {
        r = Z * up[tt.at(t, u, v - 1)];
        if (v > 1) r += (v - 1) * up[tt.at(t, u, v - 2)];
      }
      cur[k] = r;
    }
  }
  std::memcpy(R, work, sizeof(double) * nt);
}

// ---------------------------------------------------------------- shell pairs
struct PrimPair {
  double p, P[3], pref;  // pref = exp(-mu |AB|^2) c_a c_b
};

struct ShellPair {
  int a = 0, b = 0, la = 0, lb = 0, nab = 0, nt = 0;
  std::vector<PrimPair> prims;
  std::vector<double> E;  // [prim][nab cart pairs][nt hermite] : Ex*Ey*Ez
};

void make_pair(const Shell &A, const Shell &B, ShellPair &sp) {
  sp.la = A.l, sp.lb = B.l;
  const int L = A.l + B.l;
  sp.nt = ntuv(L);
  sp.nab = ncart(A.l) * ncart(B.l);
  sp.prims.clear();
  sp.E.clear();
  std::vector<int> ca, cb;
  cart_list(A.l, ca), cart_list(B.l, cb);
  const double AB2 = (A.O[0] - B.O[0]) * (A.O[0] - B.O[0]) + (A.O[1] - B.O[1]) * (A.O[1] - B.O[1]) +
                     (A.O[2] - B.O[2]) * (A.O[2] - B.O[2]);
  const TuvTable &tt = tuv();
  E1D ex[3];
  for (size_t i = 0; i < A.exps.size(); ++i)
    for (size_t j = 0; j < B.exps.size(); ++j) {
      const double a = A.exps[i], b = B.exps[j], p = a + b;
      PrimPair pp;
      pp.p = p;
      pp.pref = std::exp(-a * b / p * AB2) * A.coefs[i] * B.coefs[j];
      // primitive-pair screening far below libint's 1e-14 engine precision
      if (std::fabs(pp.pref) * std::pow(kPi / p, 1.5) < 1e-22) continue;
      for (int d = 0; d < 3; ++d) {
        pp.P[d] = (a * A.O[d] + b * B.O[d]) / p;
        ex[d].build(A.l, B.l, p, pp.P[d] - A.O[d], pp.P[d] - B.O[d]);
      }
      sp.prims.push_back(pp);
      const size_t off = sp.E.size();
      sp.E.resize(off + size_t(sp.nab) * sp.nt, 0.0);
      double *E = &sp.E[off];
      for (int x = 0; x < ncart(A.l); ++x)
        for (int y = 0; y < ncart(B.l); ++y) {
          const int ax = ca[3 * x], ay = ca[3 * x + 1], az = ca[3 * x + 2];
          const int bx = cb[3 * y], by = cb[3 * y + 1], bz = cb[3 * y + 2];
          double *row = E + size_t(x * ncart(B.l) + y) * sp.nt;
          for (int t = 0; t <= ax + bx; ++t)
            for (int u = 0; u <= ay + by; ++u)
              for (int v = 0; v <= az + bz; ++v)
                row[tt.at(t, u, v)] = ex[0].at(ax, bx, t) * ex[1].at(ay, by, u) * ex[2].at(az, bz, v);
        }
    }
}

// Apply the pure transformation to one index of a row-major block with dims [n0][n1][n2][n3].
// idx selects which dimension; shells with pure == false are left alone.
void pure_transform(std::vector<double> &blk, int dims[4], int idx, const Shell &s, std::vector<double> &tmp) {
  if (!s.pure) return;
  const std::vector<double> &T = pure_T(s.l);
  const int nc = ncart(s.l), np = 2 * s.l + 1;
  int outer = 1, inner = 1;
  for (int d = 0; d < idx; ++d) outer *= dims[d];
  for (int d = idx + 1; d < 4; ++d) inner *= dims[d];
  tmp.assign(size_t(outer) * np * inner, 0.0);
  for (int o = 0; o < outer; ++o)
    for (int m = 0; m < np; ++m)
      for (int c = 0; c < nc; ++c) {
        const double t = T[size_t(m) * nc + c];
        if (t == 0.0) continue;
        const double *src = &blk[(size_t(o) * nc + c) * inner];
        double *dst = &tmp[(size_t(o) * np + m) * inner];
        for (int i = 0; i < inner; ++i) dst[i] += t * src[i];
      }
  blk.swap(tmp);
  dims[idx] = np;
}

// ---------------------------------------------------------------- one-electron integrals
enum class OneE { Overlap, Kinetic, Nuclear };

void one_electron(const SystemData &sys, OneE kind, double *I) {
  const int n = sys.nbf;
  std::vector<double> blk, tmp;
  std::vector<int> ca, cb;
  const TuvTable &tt = tuv();
  std::vector<double> R, work;
  for (size_t si = 0; si < sys.shells.size(); ++si)
    for (size_t sj = si; sj < sys.shells.size(); ++sj) {
      const Shell &A = sys.shells[si], &B = sys.shells[sj];
      cart_list(A.l, ca), cart_list(B.l, cb);
      const int na = ncart(A.l), nb = ncart(B.l);
      blk.assign(size_t(na) * nb, 0.0);
      const double AB2 = (A.O[0] - B.O[0]) * (A.O[0] - B.O[0]) + (A.O[1] - B.O[1]) * (A.O[1] - B.O[1]) +
                         (A.O[2] - B.O[2]) * (A.O[2] - B.O[2]);
      E1D ex[3];
      for (size_t i = 0; i < A.exps.size(); ++i)
        for (size_t j = 0; j < B.exps.size(); ++j) {
          const double a = A.exps[i], b = B.exps[j], p = a + b;
          const double K = std::exp(-a * b / p * AB2) * A.coefs[i] * B.coefs[j];
          double P[3];
          const int extra = kind == OneE::Kinetic ? 2 : 0;
          for (int d = 0; d < 3; ++d) {
            P[d] = (a * A.O[d] + b * B.O[d]) / p;
            ex[d].build(A.l, B.l + extra, p, P[d] - A.O[d], P[d] - B.O[d]);
          }
          const double sq = std::sqrt(kPi / p);
          if (kind == OneE::Nuclear) {
            const int L = A.l + B.l;
            R.resize(ntuv(L));
            work.resize(size_t(L + 1) * ntuv(L));
            std::vector<double> acc(ntuv(L), 0.0);
            for (size_t c = 0; c < sys.Z.size(); ++c) {
              hermite_R(L, p, P[0] - sys.xyz[3 * c], P[1] - sys.xyz[3 * c + 1], P[2] - sys.xyz[3 * c + 2],
                        -double(sys.Z[c]), R.data(), work.data());
              for (int k = 0; k < ntuv(L); ++k) acc[k] += R[k];
            }
            const double f = 2 * kPi / p * K;
            for (int x = 0; x < na; ++x)
              for (int y = 0; y < nb; ++y) {
                double s = 0;
                for (int t = 0; t <= ca[3 * x] + cb[3 * y]; ++t)
                  for (int u = 0; u <= ca[3 * x + 1] + cb[3 * y + 1]; ++u)
                    for (int v = 0; v <= ca[3 * x + 2] + cb[3 * y + 2]; ++v)
                      s += ex[0].at(ca[3 * x], cb[3 * y], t) * ex[1].at(ca[3 * x + 1], cb[3 * y + 1], u) *
                           ex[2].at(ca[3 * x + 2], cb[3 * y + 2], v) * acc[tt.at(t, u, v)];
                blk[size_t(x) * nb + y] += f * s;
              }
          } else {
            for (int x = 0; x < na; ++x)
              for (int y = 0; y < nb; ++y) {
                double S[3], D[3];
                for (int d = 0; d < 3; ++d) {
                  const int ia = ca[3 * x + d], jb = cb[3 * y + d];
                  S[d] = ex[d].at(ia, jb, 0) * sq;
                  if (kind == OneE::Kinetic) {
                    double v = -2 * b * (2 * jb + 1) * S[d] + 4 * b * b * ex[d].at(ia, jb + 2, 0) * sq;
                    if (jb >= 2) v += jb * (jb - 1) * ex[d].at(ia, jb - 2, 0) * sq;
                    D[d] = v;
                  }
                }
                double val = S[0] * S[1] * S[2];
                if (kind == OneE::Kinetic) val = -0.5 * (D[0] * S[1] * S[2] + S[0] * D[1] * S[2] + S[0] * S[1] * D[2]);
                blk[size_t(x) * nb + y] += K * val;
              }
          }
        }
      int dims[4] = {1, na, nb, 1};
      pure_transform(blk, dims, 1, A, tmp);
      pure_transform(blk, dims, 2, B, tmp);
      for (int x = 0; x < dims[1]; ++x)
        for (int y = 0; y < dims[2]; ++y) {
          const double v = blk[size_t(x) * dims[2] + y];
          I[size_t(A.bf0 + x) * n + (B.bf0 + y)] = v;
          I[size_t(B.bf0 + y) * n + (A.bf0 + x)] = v;
        }
    }
}

// ---------------------------------------------------------------- electron repulsion integrals
struct EriWork {
  std::vector<double> R, Rw, Rmat, W, cart, tmp, Ewk;
};

// Contracted (ab|cd) block for shell pairs ab, cd in Cartesian components [na][nb][nc][nd], then the
// pure transformation of all four indices.  Returns dims in `dims`.
void eri_quartet(const SystemData &sys, const ShellPair &ab, const ShellPair &cd, EriWork &w, int dims[4]) {
  const Shell &A = sys.shells[ab.a], &B = sys.shells[ab.b], &C = sys.shells[cd.a], &D = sys.shells[cd.b];
  const int Lab = ab.la + ab.lb, Lcd = cd.la + cd.lb, L = Lab + Lcd;
  const int ntab = ntuv(Lab), ntcd = ntuv(Lcd), nt = ntuv(L);
  const TuvTable &tt = tuv();
  w.cart.assign(size_t(ab.nab) * cd.nab, 0.0);
  w.R.resize(nt);
  w.Rw.resize(size_t(L + 1) * nt);
  w.Rmat.resize(size_t(ntab) * ntcd);
  w.W.resize(size_t(ntab) * cd.nab);
  // ket Hermite sign (-1)^(tau+nu+phi) and the combined index table
  static thread_local std::vector<int> comb;
  comb.resize(size_t(ntab) * ntcd);
  for (int i = 0; i < ntab; ++i)
    for (int j = 0; j < ntcd; ++j)
      comb[size_t(i) * ntcd + j] = tt.at(tt.t[i] + tt.t[j], tt.u[i] + tt.u[j], tt.v[i] + tt.v[j]);
  const double c0 = 2.0 * std::pow(kPi, 2.5);
  for (size_t ip = 0; ip < ab.prims.size(); ++ip) {
    const PrimPair &P = ab.prims[ip];
    std::fill(w.W.begin(), w.W.end(), 0.0);
    for (size_t kp = 0; kp < cd.prims.size(); ++kp) {
      const PrimPair &Q = cd.prims[kp];
      const double p = P.p, q = Q.p, alpha = p * q / (p + q);
      const double pref = c0 / (p * q * std::sqrt(p + q)) * P.pref * Q.pref;
      hermite_R(L, alpha, P.P[0] - Q.P[0], P.P[1] - Q.P[1], P.P[2] - Q.P[2], pref, w.R.data(), w.Rw.data());
      for (int i = 0; i < ntab; ++i)
        for (int j = 0; j < ntcd; ++j) {
          const double sgn = ((tt.t[j] + tt.u[j] + tt.v[j]) & 1) ? -1.0 : 1.0;
          w.Rmat[size_t(i) * ntcd + j] = sgn * w.R[comb[size_t(i) * ntcd + j]];
        }
      // W[i][cd] += sum_j Rmat[i][j] * Ecd[cd][j]
      const double *Ecd = &cd.E[kp * size_t(cd.nab) * cd.nt];
      for (int i = 0; i < ntab; ++i) {
        const double *rm = &w.Rmat[size_t(i) * ntcd];
        double *wr = &w.W[size_t(i) * cd.nab];
        for (int c = 0; c < cd.nab; ++c) {
          const double *e = Ecd + size_t(c) * cd.nt;
          double s = 0;
          for (int j = 0; j < ntcd; ++j) s += rm[j] * e[j];
          wr[c] += s;
        }
      }
    }
    // cart[ab][cd] += sum_i Eab[ab][i] * W[i][cd]
    const double *Eab = &ab.E[ip * size_t(ab.nab) * ab.nt];
    for (int x = 0; x < ab.nab; ++x) {
      const double *e = Eab + size_t(x) * ab.nt;
      double *out = &w.cart[size_t(x) * cd.nab];
      for (int i = 0; i < ntab; ++i) {
        const double ev = e[i];
        if (ev == 0.0) continue;
        const double *wr = &w.W[size_t(i) * cd.nab];
        for (int c = 0; c < cd.nab; ++c) out[c] += ev * wr[c];
      }
    }
  }
  dims[0] = ncart(A.l), dims[1] = ncart(B.l), dims[2] = ncart(C.l), dims[3] = ncart(D.l);
  pure_transform(w.cart, dims, 0, A, w.tmp);
  pure_transform(w.cart, dims, 1, B, w.tmp);
  pure_transform(w.cart, dims, 2, C, w.tmp);
  pure_transform(w.cart, dims, 3, D, w.tmp);
}

int thread_count() {
  if (g_threads > 0) return g_threads;
  for (const char *env : {"ZQ_NUM_THREADS", "OMP_NUM_THREADS"}) {
    const char *v = std::getenv(env);
    if (v && std::atoi(v) > 0) return std::atoi(v);
  }
  const unsigned h = std::thread::hardware_concurrency();
  return h ? int(h) : 1;
}

template <class Sink>
void two_electron(const SystemData &sys, Sink sink) {
  const int ns = int(sys.shells.size());
  // shell pair tables (i <= j) shared read-only by all threads
  std::vector<ShellPair> pairs(size_t(ns) * (ns + 1) / 2);
  auto pidx = [&](int i, int j) { return size_t(j) * (j + 1) / 2 + i; };  // i <= j
  for (int j = 0; j < ns; ++j)
    for (int i = 0; i <= j; ++i) {
      ShellPair &sp = pairs[pidx(i, j)];
      sp.a = i, sp.b = j;
      make_pair(sys.shells[i], sys.shells[j], sp);
    }
  // bra work items (i, j) with i <= j, enumerated as in src/libint.cpp:115-119
  std::vector<std::pair<int, int>> bras;
  for (int i = 0; i < ns; ++i)
    for (int j = i; j < ns; ++j) bras.push_back({i, j});
  std::atomic<size_t> next{0};
  auto worker = [&]() {
    EriWork w;
    int dims[4];
    for (;;) {
      const size_t it = next.fetch_add(1);
      if (it >= bras.size()) break;
      const int i = bras[it].first, j = bras[it].second;
      const ShellPair &ab = pairs[pidx(i, j)];
      for (int k = i; k < ns; ++k)
        for (int l = (i == k ? j : k); l < ns; ++l) {
          const ShellPair &cd = pairs[pidx(k, l)];
          if (ab.prims.empty() || cd.prims.empty()) continue;
          eri_quartet(sys, ab, cd, w, dims);
          sink(sys.shells[i], sys.shells[j], sys.shells[k], sys.shells[l], w.cart.data(), dims);
        }
    }
  };
  const int nth = std::max(1, std::min<int>(thread_count(), int(bras.size())));
  std::vector<std::thread> th;
  for (int t = 1; t < nth; ++t) th.emplace_back(worker);
  worker();
  for (auto &t : th) t.join();
}

SystemData *finish(SystemData *sys, const Basis &b) {
  if (!build_shells(*sys, b)) {
    delete sys;
    return nullptr;
  }
  return sys;
}

}  // namespace

extern "C" {

SystemData *libint_init(const char *system, const char *basis) {
  g_last_error.clear();
  if (!system || !basis) return nullptr;
  std::ifstream f(system);
  if (!f) {
    g_last_error = std::string("cannot open ") + system;
    return nullptr;
  }
  int nat = 0;
  std::string line;
  if (!std::getline(f, line)) return nullptr;
  nat = std::atoi(line.c_str());
  std::getline(f, line);  // comment
  auto *sys = new SystemData;
  for (int a = 0; a < nat; ++a) {
    if (!std::getline(f, line)) {
      g_last_error = "truncated xyz file";
      delete sys;
      return nullptr;
    }
    std::istringstream ss(line);
    std::string sym;
    double x, y, z;
    ss >> sym >> x >> y >> z;
    const int Z = symbol_to_z(sym);
    if (Z < 0 || ss.fail()) {
      g_last_error = "bad xyz line: " + line;
      delete sys;
      return nullptr;
    }
    sys->Z.push_back(Z);
    for (double c : {x, y, z}) sys->xyz.push_back(c / ZQ_BOHR_TO_ANGSTROM);
  }
  Basis b;
  if (!load_basis(basis, b)) {
    delete sys;
    return nullptr;
  }
  return finish(sys, b);
}

SystemData *zq_ints_from_atoms(size_t nat, const int *Z, const double *xyz_bohr, const char *basis) {
  g_last_error.clear();
  if (!Z || !xyz_bohr || !basis || nat == 0) return nullptr;
  auto *sys = new SystemData;
  sys->Z.assign(Z, Z + nat);
  sys->xyz.assign(xyz_bohr, xyz_bohr + 3 * nat);
  Basis b;
  if (!load_basis(basis, b)) {
    delete sys;
    return nullptr;
  }
  return finish(sys, b);
}

void libint_deinit(SystemData *sys) { delete sys; }

SystemData *libint_clone(SystemData *sys) { return sys ? new SystemData(*sys) : nullptr; }

size_t libint_nbf(SystemData *sys) { return sys ? size_t(sys->nbf) : 0; }

size_t libint_nat(SystemData *sys) { return sys ? sys->Z.size() : 0; }

size_t zq_ints_nel(SystemData *sys) {
  size_t n = 0;
  if (sys)
    for (int z : sys->Z) n += size_t(z);
  return n;
}

void libint_atoms(int *atoms, SystemData *sys) {
  if (!sys) return;
  for (size_t i = 0; i < sys->Z.size(); ++i) atoms[i] = sys->Z[i];
}

void libint_coors(double *coors, SystemData *sys) {
  if (!sys) return;
  std::memcpy(coors, sys->xyz.data(), sizeof(double) * sys->xyz.size());
}

void libint_bf2at(int *map, SystemData *sys) {
  if (!sys) return;
  for (const Shell &s : sys->shells)
    for (int k = 0; k < s.nbf(); ++k) map[s.bf0 + k] = s.atom;
}

void libint_update_coords(SystemData *sys, const double *coords) {
  if (!sys) return;
  sys->xyz.assign(coords, coords + sys->xyz.size());
  for (Shell &s : sys->shells)
    for (int d = 0; d < 3; ++d) s.O[d] = sys->xyz[3 * s.atom + d];
}

void libint_overlap(double *I, SystemData *sys) {
  if (sys) one_electron(*sys, OneE::Overlap, I);
}

void libint_kinetic(double *I, SystemData *sys) {
  if (sys) one_electron(*sys, OneE::Kinetic, I);
}

void libint_nuclear(double *I, SystemData *sys) {
  if (sys) one_electron(*sys, OneE::Nuclear, I);
}

void libint_coulomb(double *I, SystemData *sys) {
  if (!sys) return;
  const size_t n = size_t(sys->nbf);
  two_electron(*sys, [&](const Shell &A, const Shell &B, const Shell &C, const Shell &D, const double *v, const int *d) {
    size_t idx = 0;
    for (int m = 0; m < d[0]; ++m) {
      const size_t b1 = A.bf0 + m;
      for (int nn = 0; nn < d[1]; ++nn) {
        const size_t b2 = B.bf0 + nn;
        for (int o = 0; o < d[2]; ++o) {
          const size_t b3 = C.bf0 + o;
          for (int p = 0; p < d[3]; ++p) {
            const size_t b4 = D.bf0 + p;
            const double val = v[idx++];
            // (b1 b2|b3 b4) = <b1 b3|b2 b4>, all 8 permutations (src/libint.cpp:143-150)
            I[((b1 * n + b3) * n + b2) * n + b4] = val;
            I[((b1 * n + b4) * n + b2) * n + b3] = val;
            I[((b2 * n + b3) * n + b1) * n + b4] = val;
            I[((b2 * n + b4) * n + b1) * n + b3] = val;
            I[((b3 * n + b1) * n + b4) * n + b2] = val;
            I[((b3 * n + b2) * n + b4) * n + b1] = val;
            I[((b4 * n + b1) * n + b3) * n + b2] = val;
            I[((b4 * n + b2) * n + b3) * n + b1] = val;
          }
        }
      }
    }
  });
}

void zq_ints_coulomb_packed(double *out, SystemData *sys) {
  if (!sys) return;
  two_electron(*sys, [&](const Shell &A, const Shell &B, const Shell &C, const Shell &D, const double *v, const int *d) {
    size_t idx = 0;
    for (int m = 0; m < d[0]; ++m) {
      const size_t b1 = A.bf0 + m;
      for (int nn = 0; nn < d[1]; ++nn) {
        const size_t b2 = B.bf0 + nn;
        const size_t p = std::max(b1, b2), q = std::min(b1, b2), I = p * (p + 1) / 2 + q;
        for (int o = 0; o < d[2]; ++o) {
          const size_t b3 = C.bf0 + o;
          for (int pp = 0; pp < d[3]; ++pp) {
            const size_t b4 = D.bf0 + pp;
            const size_t r = std::max(b3, b4), s = std::min(b3, b4), J = r * (r + 1) / 2 + s;
            const size_t hi = std::max(I, J), lo = std::min(I, J);
            out[hi * (hi + 1) / 2 + lo] = v[idx++];
          }
        }
      }
    }
  });
}

double zq_ints_nrep(SystemData *sys) {
  if (!sys) return 0.0;
  double e = 0;
  for (size_t i = 0; i < sys->Z.size(); ++i)
    for (size_t j = 0; j < i; ++j) {
      const double dx = sys->xyz[3 * i] - sys->xyz[3 * j], dy = sys->xyz[3 * i + 1] - sys->xyz[3 * j + 1],
                   dz = sys->xyz[3 * i + 2] - sys->xyz[3 * j + 2];
      e += double(sys->Z[i]) * sys->Z[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
    }
  return e;
}

void zq_ints_set_threads(int nthreads) { g_threads = nthreads; }

const char *zq_ints_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
