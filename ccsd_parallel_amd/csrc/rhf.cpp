// Closed-shell RHF input producer for the CCSD drop-in (declared in
// include/ccsd_rhf.h).  Host FP64 code.
//
// Integrals: McMurchie-Davidson over Cartesian Gaussians (Hermite expansion
// coefficients E_t^{ij}, Hermite Coulomb integrals R_tuv, Boys function),
// the role of hartree_integrals.cpp:118-796 in the reference.  SCF: core-
// Hamiltonian guess, symmetric orthogonalisation, Pulay DIIS on the
// commutator F P S - S P F (the reference's DIIS, hartree_fock_solver.cpp:
// 1252-1326), converged on |dE| and the commutator rather than the
// reference's change of the lowest orbital energy (:940), so the orbitals
// handed to CCSD are stationary to ~1e-10.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ccsd_rhf.h"

namespace {

constexpr int LMAX = 2;            // up to d functions in the engine (the basis data uses s and p)
constexpr int NR = 4 * LMAX + 1;   // Hermite orders for an ERI
const double kPi = 3.14159265358979323846;

struct ShellDef {
  int L;
  std::vector<double> exps, coefs;
};

// Basis data of the reference (fill_alpha.cpp): 6-31G at :16403-16424 (H) and
// :16919-17010 (O); STO-3G from :23296 on.  The published Pople sets.
std::vector<ShellDef> element_shells(int Z, const std::string& basis) {
  if (basis == "STO-3G") {
    if (Z == 1) return {{0, {3.42525091, 0.62391373, 0.1688554}, {0.15432897, 0.53532814, 0.44463454}}};
    if (Z == 8)
      return {{0, {130.70932, 23.808861, 6.4436083}, {0.15432897, 0.53532814, 0.44463454}},
              {0, {5.0331513, 1.1695961, 0.380389}, {-0.09996723, 0.39951283, 0.70011547}},
              {1, {5.0331513, 1.1695961, 0.380389}, {0.15591627, 0.60768372, 0.39195739}}};
  } else if (basis == "6-31G") {
    if (Z == 1)
      return {{0, {18.731137, 2.8253937, 0.6401217}, {0.0334946, 0.23472695, 0.81375733}},
              {0, {0.1612778}, {1.0}}};
    if (Z == 8)
      return {{0, {5484.6717, 825.23495, 188.04696, 52.9645, 16.89757, 5.7996353},
               {0.0018311, 0.0139501, 0.0684451, 0.2327143, 0.470193, 0.3585209}},
              {0, {15.539616, 3.5999336, 1.0137618}, {-0.1107775, -0.1480263, 1.130767}},
              {1, {15.539616, 3.5999336, 1.0137618}, {0.0708743, 0.3397528, 0.7271586}},
              {0, {0.2700058}, {1.0}},
              {1, {0.2700058}, {1.0}}};
  }
  return {};
}

// One contracted Cartesian function x^l y^m z^n sum_k c_k exp(-a_k r^2) at A;
// coefficients include primitive and contraction normalisation.
struct BasisFn {
  double A[3];
  int l[3];
  std::vector<double> a, c;
};

double dfact(int n) {  // (n)!! with (-1)!! = 1
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

// Hermite expansion coefficients along one axis: E[i][j][t] for i <= li, j <= lj,
// Gaussian exponents a (centre A) and b (centre B), Q = A - B.
struct ETab {
  double e[LMAX + 1][LMAX + 3][2 * LMAX + 4];
};
void hermite_e(int li, int lj, double Q, double a, double b, ETab& E) {
  const double p = a + b, q = a * b / p;
  std::memset(&E, 0, sizeof(E));
  E.e[0][0][0] = std::exp(-q * Q * Q);
  auto at = [&](int i, int j, int t) -> double {
    if (t < 0 || t > i + j) return 0.0;
    return E.e[i][j][t];
  };
  for (int i = 0; i < li; ++i)
    for (int t = 0; t <= i + 1; ++t)
      E.e[i + 1][0][t] = at(i, 0, t - 1) / (2 * p) - (q * Q / a) * at(i, 0, t) + (t + 1) * at(i, 0, t + 1);
  for (int i = 0; i <= li; ++i)
    for (int j = 0; j < lj; ++j)
      for (int t = 0; t <= i + j + 1; ++t)
        E.e[i][j + 1][t] = at(i, j, t - 1) / (2 * p) + (q * Q / b) * at(i, j, t) + (t + 1) * at(i, j, t + 1);
}

// Boys function F_n(T) for n = 0..nmax: power series with downward recursion
// for moderate T, erf and upward recursion for large T.
void boys(int nmax, double T, double* F) {
  if (T < 1e-15) {
    for (int n = 0; n <= nmax; ++n) F[n] = 1.0 / (2 * n + 1);
    return;
  }
  const double ex = std::exp(-T);
  if (T > 35.0) {
    F[0] = 0.5 * std::sqrt(kPi / T) * std::erf(std::sqrt(T));
    for (int n = 0; n < nmax; ++n) F[n + 1] = ((2 * n + 1) * F[n] - ex) / (2 * T);
    return;
  }
  double term = 1.0 / (2 * nmax + 1), sum = term;
  for (int k = 1; k < 400; ++k) {
    term *= 2 * T / (2 * nmax + 2 * k + 1);
    sum += term;
    if (term < 1e-17 * sum) break;
  }
  F[nmax] = ex * sum;
  for (int n = nmax; n > 0; --n) F[n - 1] = (2 * T * F[n] + ex) / (2 * n - 1);
}

// Hermite Coulomb integrals R^0_{tuv}(p, X, Y, Z) for t+u+v <= N.
struct RTab {
  double r[NR][NR][NR];
};
void hermite_r(int N, double p, double X, double Y, double Z, RTab& out) {
  static thread_local double R[NR][NR][NR][NR];  // [n][t][u][v]
  double F[NR];
  boys(N, p * (X * X + Y * Y + Z * Z), F);
  double f = 1.0;
  for (int n = 0; n <= N; ++n) {
    R[n][0][0][0] = f * F[n];
    f *= -2 * p;
  }
  for (int L = 1; L <= N; ++L)
    for (int t = 0; t <= L; ++t)
      for (int u = 0; u <= L - t; ++u) {
        const int v = L - t - u;
        for (int n = 0; n <= N - L; ++n) {
          double val;
          if (t > 0) val = (t > 1 ? (t - 1) * R[n + 1][t - 2][u][v] : 0.0) + X * R[n + 1][t - 1][u][v];
          else if (u > 0) val = (u > 1 ? (u - 1) * R[n + 1][t][u - 2][v] : 0.0) + Y * R[n + 1][t][u - 1][v];
          else val = (v > 1 ? (v - 1) * R[n + 1][t][u][v - 2] : 0.0) + Z * R[n + 1][t][u][v - 1];
          R[n][t][u][v] = val;
        }
      }
  for (int t = 0; t <= N; ++t)
    for (int u = 0; u <= N - t; ++u)
      for (int v = 0; v <= N - t - u; ++v) out.r[t][u][v] = R[0][t][u][v];
}

struct Molecule {
  int n_atoms = 0;
  std::vector<int> Z;
  std::vector<double> R;  // bohr
};

int build_basis(const Molecule& mol, const std::string& basis, std::vector<BasisFn>& fns) {
  fns.clear();
  for (int at = 0; at < mol.n_atoms; ++at) {
    std::vector<ShellDef> sh = element_shells(mol.Z[at], basis);
    if (sh.empty()) return CCSD_ERR_ARG;
    for (const ShellDef& s : sh) {
      // Cartesian components in the reference's order: x, y, z for p (fill_alpha.cpp Potenser rows)
      std::vector<std::array<int, 3>> comps;
      for (int lx = s.L; lx >= 0; --lx)
        for (int ly = s.L - lx; ly >= 0; --ly) comps.push_back({lx, ly, s.L - lx - ly});
      for (auto& lc : comps) {
        BasisFn f;
        for (int x = 0; x < 3; ++x) { f.A[x] = mol.R[3 * at + x]; f.l[x] = lc[x]; }
        f.a = s.exps;
        f.c.resize(s.exps.size());
        const double dd = dfact(2 * lc[0] - 1) * dfact(2 * lc[1] - 1) * dfact(2 * lc[2] - 1);
        for (size_t k = 0; k < s.exps.size(); ++k) {
          const double a = s.exps[k];
          f.c[k] = s.coefs[k] * std::pow(2 * a / kPi, 0.75) * std::pow(4 * a, 0.5 * s.L) / std::sqrt(dd);
        }
        // contracted normalisation: <f|f> = 1
        double ss = 0;
        for (size_t i = 0; i < f.a.size(); ++i)
          for (size_t j = 0; j < f.a.size(); ++j) {
            const double p = f.a[i] + f.a[j];
            double v = std::pow(kPi / p, 1.5);
            for (int x = 0; x < 3; ++x) v *= dfact(2 * f.l[x] - 1) / std::pow(2 * p, f.l[x]);
            ss += f.c[i] * f.c[j] * v;
          }
        for (double& c : f.c) c /= std::sqrt(ss);
        fns.push_back(f);
      }
    }
  }
  return CCSD_OK;
}

// One-electron matrices S, T (kinetic), V (nuclear attraction), row-major n x n.
void one_electron(const Molecule& mol, const std::vector<BasisFn>& fns, std::vector<double>& S,
                  std::vector<double>& T, std::vector<double>& V) {
  const int n = (int)fns.size();
  S.assign((size_t)n * n, 0.0);
  T.assign((size_t)n * n, 0.0);
  V.assign((size_t)n * n, 0.0);
  for (int mu = 0; mu < n; ++mu)
    for (int nu = 0; nu <= mu; ++nu) {
      const BasisFn &f = fns[mu], &g = fns[nu];
      double s = 0, t = 0, v = 0;
      for (size_t i = 0; i < f.a.size(); ++i)
        for (size_t j = 0; j < g.a.size(); ++j) {
          const double a = f.a[i], b = g.a[j], p = a + b, cc = f.c[i] * g.c[j];
          ETab E[3];
          double P[3];
          for (int x = 0; x < 3; ++x) {
            hermite_e(f.l[x], g.l[x] + 2, f.A[x] - g.A[x], a, b, E[x]);
            P[x] = (a * f.A[x] + b * g.A[x]) / p;
          }
          const double pre = std::pow(kPi / p, 1.5);
          double s1[3], t1[3];
          for (int x = 0; x < 3; ++x) {
            const int li = f.l[x], lj = g.l[x];
            s1[x] = E[x].e[li][lj][0];
            const double sm2 = lj >= 2 ? E[x].e[li][lj - 2][0] : 0.0;
            t1[x] = -0.5 * (lj * (lj - 1) * sm2 - 2 * b * (2 * lj + 1) * s1[x] + 4 * b * b * E[x].e[li][lj + 2][0]);
          }
          s += cc * pre * s1[0] * s1[1] * s1[2];
          t += cc * pre * (t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2]);
          const int L = f.l[0] + f.l[1] + f.l[2] + g.l[0] + g.l[1] + g.l[2];
          for (int at = 0; at < mol.n_atoms; ++at) {
            RTab R;
            hermite_r(L, p, P[0] - mol.R[3 * at], P[1] - mol.R[3 * at + 1], P[2] - mol.R[3 * at + 2], R);
            double acc = 0;
            for (int tx = 0; tx <= f.l[0] + g.l[0]; ++tx)
              for (int ty = 0; ty <= f.l[1] + g.l[1]; ++ty)
                for (int tz = 0; tz <= f.l[2] + g.l[2]; ++tz)
                  acc += E[0].e[f.l[0]][g.l[0]][tx] * E[1].e[f.l[1]][g.l[1]][ty] * E[2].e[f.l[2]][g.l[2]][tz] *
                         R.r[tx][ty][tz];
            v -= cc * mol.Z[at] * 2 * kPi / p * acc;
          }
        }
      S[(size_t)mu * n + nu] = S[(size_t)nu * n + mu] = s;
      T[(size_t)mu * n + nu] = T[(size_t)nu * n + mu] = t;
      V[(size_t)mu * n + nu] = V[(size_t)nu * n + mu] = v;
    }
}

// Primitive pair data reused over all quartets.
struct PairPrim {
  double p, P[3], cc;
  ETab E[3];
};
struct Pair {
  int mu, nu, lt[3];
  std::vector<PairPrim> prim;
};

// AO electron repulsion integrals (mu nu | la si), full n^4 array, 8-fold symmetry.
void two_electron(const std::vector<BasisFn>& fns, std::vector<double>& eri) {
  const int n = (int)fns.size();
  const size_t n2 = (size_t)n * n;
  eri.assign(n2 * n2, 0.0);
  std::vector<Pair> pairs;
  for (int mu = 0; mu < n; ++mu)
    for (int nu = 0; nu <= mu; ++nu) {
      const BasisFn &f = fns[mu], &g = fns[nu];
      Pair pr;
      pr.mu = mu; pr.nu = nu;
      for (int x = 0; x < 3; ++x) pr.lt[x] = f.l[x] + g.l[x];
      for (size_t i = 0; i < f.a.size(); ++i)
        for (size_t j = 0; j < g.a.size(); ++j) {
          PairPrim pp;
          const double a = f.a[i], b = g.a[j];
          pp.p = a + b;
          pp.cc = f.c[i] * g.c[j];
          for (int x = 0; x < 3; ++x) {
            hermite_e(f.l[x], g.l[x], f.A[x] - g.A[x], a, b, pp.E[x]);
            pp.P[x] = (a * f.A[x] + b * g.A[x]) / pp.p;
          }
          pr.prim.push_back(pp);
        }
      pairs.push_back(pr);
    }
  const double pre0 = 2 * std::pow(kPi, 2.5);
  for (size_t ia = 0; ia < pairs.size(); ++ia)
    for (size_t ib = 0; ib <= ia; ++ib) {
      const Pair &A = pairs[ia], &B = pairs[ib];
      const BasisFn &f = fns[A.mu], &g = fns[A.nu], &h = fns[B.mu], &k = fns[B.nu];
      const int N = A.lt[0] + A.lt[1] + A.lt[2] + B.lt[0] + B.lt[1] + B.lt[2];
      double val = 0;
      for (const PairPrim& x : A.prim)
        for (const PairPrim& y : B.prim) {
          const double alpha = x.p * y.p / (x.p + y.p);
          RTab R;
          hermite_r(N, alpha, x.P[0] - y.P[0], x.P[1] - y.P[1], x.P[2] - y.P[2], R);
          double acc = 0;
          for (int t = 0; t <= A.lt[0]; ++t)
            for (int u = 0; u <= A.lt[1]; ++u)
              for (int v = 0; v <= A.lt[2]; ++v) {
                const double eab = x.E[0].e[f.l[0]][g.l[0]][t] * x.E[1].e[f.l[1]][g.l[1]][u] *
                                   x.E[2].e[f.l[2]][g.l[2]][v];
                if (eab == 0.0) continue;
                double inner = 0;
                for (int tau = 0; tau <= B.lt[0]; ++tau)
                  for (int nu = 0; nu <= B.lt[1]; ++nu)
                    for (int ph = 0; ph <= B.lt[2]; ++ph) {
                      const double ecd = y.E[0].e[h.l[0]][k.l[0]][tau] * y.E[1].e[h.l[1]][k.l[1]][nu] *
                                         y.E[2].e[h.l[2]][k.l[2]][ph];
                      const double sg = ((tau + nu + ph) & 1) ? -1.0 : 1.0;
                      inner += sg * ecd * R.r[t + tau][u + nu][v + ph];
                    }
                acc += eab * inner;
              }
          val += x.cc * y.cc * pre0 / (x.p * y.p * std::sqrt(x.p + y.p)) * acc;
        }
      const int i = A.mu, j = A.nu, kk = B.mu, l = B.nu;
      const int idx[8][4] = {{i, j, kk, l}, {j, i, kk, l}, {i, j, l, kk}, {j, i, l, kk},
                             {kk, l, i, j}, {l, kk, i, j}, {kk, l, j, i}, {l, kk, j, i}};
      for (auto& q : idx) eri[(((size_t)q[0] * n + q[1]) * n + q[2]) * n + q[3]] = val;
    }
}

// Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (row-major).
// Eigenvalues ascending in w; eigenvectors as columns of U (U[r*n + c]), each
// with its largest-magnitude component positive.
void sym_eig(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& U) {
  U.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) U[(size_t)i * n + i] = 1.0;
  auto a = [&](int r, int c) -> double& { return A[(size_t)r * n + c]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, tot = 0;
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c) {
        tot += a(r, c) * a(r, c);
        if (r != c) off += a(r, c) * a(r, c);
      }
    if (off <= 1e-30 * std::max(tot, 1e-300)) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a(p, q);
        if (std::fabs(apq) < 1e-300) continue;
        // rotation angle zeroing a(p,q): tan(2 theta) = 2 a_pq / (a_qq - a_pp)
        const double theta = (a(q, q) - a(p, p)) / (2 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // columns p, q
          const double akp = a(k, p), akq = a(k, q);
          a(k, p) = c * akp - s * akq;
          a(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double apk = a(p, k), aqk = a(q, k);
          a(p, k) = c * apk - s * aqk;
          a(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double ukp = U[(size_t)k * n + p], ukq = U[(size_t)k * n + q];
          U[(size_t)k * n + p] = c * ukp - s * ukq;
          U[(size_t)k * n + q] = s * ukp + c * ukq;
        }
      }
  }
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return a(x, x) < a(y, y); });
  std::vector<double> U2((size_t)n * n);
  w.resize(n);
  for (int c = 0; c < n; ++c) {
    const int src = order[c];
    w[c] = a(src, src);
    int big = 0;
    for (int r = 1; r < n; ++r)
      if (std::fabs(U[(size_t)r * n + src]) > std::fabs(U[(size_t)big * n + src]) + 1e-12) big = r;
    const double sg = U[(size_t)big * n + src] < 0 ? -1.0 : 1.0;
    for (int r = 0; r < n; ++r) U2[(size_t)r * n + c] = sg * U[(size_t)r * n + src];
  }
  U.swap(U2);
}

void matmul(int n, const double* A, bool ta, const double* B, bool tb, double* C) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0;
      for (int k = 0; k < n; ++k) s += (ta ? A[(size_t)k * n + i] : A[(size_t)i * n + k]) *
                                       (tb ? B[(size_t)j * n + k] : B[(size_t)k * n + j]);
      C[(size_t)i * n + j] = s;
    }
}

// Solve the small dense system M x = r (Gaussian elimination, partial pivoting).
bool solve_dense(int m, std::vector<double> M, std::vector<double>& x) {
  for (int c = 0; c < m; ++c) {
    int piv = c;
    for (int r = c + 1; r < m; ++r)
      if (std::fabs(M[(size_t)r * (m + 1) + c]) > std::fabs(M[(size_t)piv * (m + 1) + c])) piv = r;
    if (std::fabs(M[(size_t)piv * (m + 1) + c]) < 1e-300) return false;
    if (piv != c)
      for (int k = 0; k <= m; ++k) std::swap(M[(size_t)c * (m + 1) + k], M[(size_t)piv * (m + 1) + k]);
    for (int r = 0; r < m; ++r) {
      if (r == c) continue;
      const double f = M[(size_t)r * (m + 1) + c] / M[(size_t)c * (m + 1) + c];
      for (int k = c; k <= m; ++k) M[(size_t)r * (m + 1) + k] -= f * M[(size_t)c * (m + 1) + k];
    }
  }
  x.resize(m);
  for (int r = 0; r < m; ++r) x[r] = M[(size_t)r * (m + 1) + m] / M[(size_t)r * (m + 1) + r];
  return true;
}

struct ScfOut {
  double e_hf = 0, e_nuc = 0, comm = 0;
  int iterations = 0, converged = 0, nocc = 0;
  std::vector<double> C, eps, eri;
};

int run_scf(const Molecule& mol, const std::string& basis, int n_elec, double tol, bool print, ScfOut& out) {
  if (mol.n_atoms <= 0 || n_elec <= 0 || (n_elec & 1)) return CCSD_ERR_ARG;
  std::vector<BasisFn> fns;
  int rc = build_basis(mol, basis, fns);
  if (rc != CCSD_OK) return rc;
  const int n = (int)fns.size(), nocc = n_elec / 2;
  if (nocc > n) return CCSD_ERR_ARG;
  const size_t nn = (size_t)n * n;
  std::vector<double> S, T, V;
  one_electron(mol, fns, S, T, V);
  two_electron(fns, out.eri);
  const std::vector<double>& eri = out.eri;
  double enuc = 0;
  for (int a = 0; a < mol.n_atoms; ++a)
    for (int b = 0; b < a; ++b) {
      double d2 = 0;
      for (int x = 0; x < 3; ++x) d2 += std::pow(mol.R[3 * a + x] - mol.R[3 * b + x], 2);
      if (d2 < 1e-20) return CCSD_ERR_ARG;
      enuc += mol.Z[a] * mol.Z[b] / std::sqrt(d2);
    }
  std::vector<double> H(nn);
  for (size_t i = 0; i < nn; ++i) H[i] = T[i] + V[i];
  // symmetric orthogonalisation X = U s^-1/2 U^T
  std::vector<double> sw, SU;
  sym_eig(n, S, sw, SU);
  if (sw[0] < 1e-9) return CCSD_ERR_ARG;  // linearly dependent basis
  std::vector<double> X(nn, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0;
      for (int k = 0; k < n; ++k) s += SU[(size_t)i * n + k] * SU[(size_t)j * n + k] / std::sqrt(sw[k]);
      X[(size_t)i * n + j] = s;
    }
  std::vector<double> tmp(nn), tmp2(nn), Fp(nn), Cp, eps, C(nn);
  auto diagonalise = [&](const std::vector<double>& F) {
    matmul(n, X.data(), true, F.data(), false, tmp.data());
    matmul(n, tmp.data(), false, X.data(), false, Fp.data());
    sym_eig(n, Fp, eps, Cp);
    matmul(n, X.data(), false, Cp.data(), false, C.data());
  };
  std::vector<double> P(nn), F(nn), G(nn);
  auto density = [&]() {
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double s = 0;
        for (int k = 0; k < nocc; ++k) s += 2 * C[(size_t)i * n + k] * C[(size_t)j * n + k];
        P[(size_t)i * n + j] = s;
      }
  };
  auto fock = [&]() {
    for (int mu = 0; mu < n; ++mu)
      for (int nu = 0; nu < n; ++nu) {
        double g = 0;
        for (int la = 0; la < n; ++la)
          for (int si = 0; si < n; ++si)
            g += P[(size_t)la * n + si] * (eri[(((size_t)mu * n + nu) * n + la) * n + si] -
                                           0.5 * eri[(((size_t)mu * n + la) * n + nu) * n + si]);
        F[(size_t)mu * n + nu] = H[(size_t)mu * n + nu] + g;
      }
    double e = 0;
    for (size_t i = 0; i < nn; ++i) e += 0.5 * P[i] * (H[i] + F[i]);
    return e;
  };
  diagonalise(H);
  density();
  const double tol_e = std::min(tol, 1e-10) * 1e-2, tol_c = std::min(tol, 1e-10);
  const int max_iter = 300, ndiis = 8;
  std::vector<std::vector<double>> fs, es;
  double e_old = 0, e = 0, comm = 0;
  int it = 0, conv = 0;
  for (it = 1; it <= max_iter; ++it) {
    e = fock();
    // error X^T (F P S - S P F) X
    std::vector<double> FPS(nn), SPF(nn), err(nn);
    matmul(n, F.data(), false, P.data(), false, tmp.data());
    matmul(n, tmp.data(), false, S.data(), false, FPS.data());
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) SPF[(size_t)i * n + j] = FPS[(size_t)j * n + i];
    for (size_t i = 0; i < nn; ++i) tmp2[i] = FPS[i] - SPF[i];
    matmul(n, X.data(), true, tmp2.data(), false, tmp.data());
    matmul(n, tmp.data(), false, X.data(), false, err.data());
    comm = 0;
    for (double v : err) comm = std::max(comm, std::fabs(v));
    if (print) std::printf("Energi: %.12f Steg: %d\n", e + enuc, it);
    if (it > 1 && std::fabs(e - e_old) < tol_e && comm < tol_c) {
      conv = 1;
      break;
    }
    e_old = e;
    fs.push_back(F);
    es.push_back(err);
    if ((int)fs.size() > ndiis) { fs.erase(fs.begin()); es.erase(es.begin()); }
    std::vector<double> Fx = F;
    const int m = (int)fs.size();
    if (m >= 2) {
      // bordered Pulay system [B -1; -1 0] [c; l] = [0; -1]
      std::vector<double> M((size_t)(m + 1) * (m + 2), 0.0), cvec;
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
          double s = 0;
          for (size_t k = 0; k < nn; ++k) s += es[i][k] * es[j][k];
          M[(size_t)i * (m + 2) + j] = s;
        }
      for (int i = 0; i < m; ++i) {
        M[(size_t)i * (m + 2) + m] = -1.0;
        M[(size_t)m * (m + 2) + i] = -1.0;
      }
      M[(size_t)m * (m + 2) + m + 1] = -1.0;
      if (solve_dense(m + 1, M, cvec)) {
        std::fill(Fx.begin(), Fx.end(), 0.0);
        for (int i = 0; i < m; ++i)
          for (size_t k = 0; k < nn; ++k) Fx[k] += cvec[i] * fs[i][k];
      }
    }
    diagonalise(Fx);
    density();
  }
  // canonical orbitals of the converged Fock matrix
  diagonalise(F);
  density();
  e = fock();
  out.e_hf = e + enuc;
  out.e_nuc = enuc;
  out.comm = comm;
  out.iterations = std::min(it, max_iter);
  out.converged = conv;
  out.nocc = nocc;
  out.C = C;
  out.eps = eps;
  return conv ? CCSD_OK : CCSD_ERR_NOT_CONVERGED;
}

void mo_transform(int n, const std::vector<double>& C, const std::vector<double>& ao, double* mo) {
  const size_t n2 = (size_t)n * n, n3 = n2 * n;
  std::vector<double> a(ao), b(n2 * n2);
  // transform the last index, then rotate the indices left; four times
  for (int pass = 0; pass < 4; ++pass) {
    for (size_t ijk = 0; ijk < n3; ++ijk)
      for (int s = 0; s < n; ++s) {
        double v = 0;
        for (int l = 0; l < n; ++l) v += a[ijk * n + l] * C[(size_t)l * n + s];
        // (i j k s) -> stored as (s i j k)
        b[(size_t)s * n3 + ijk] = v;
      }
    a.swap(b);
  }
  std::memcpy(mo, a.data(), n2 * n2 * sizeof(double));
}

int molecule_from(int n_atoms, const int* Z, const double* R, Molecule& mol) {
  if (n_atoms <= 0 || !Z || !R) return CCSD_ERR_ARG;
  mol.n_atoms = n_atoms;
  mol.Z.assign(Z, Z + n_atoms);
  mol.R.assign(R, R + 3 * n_atoms);
  return CCSD_OK;
}

}  // namespace

extern "C" {

int ccsd_basis_size(int n_atoms, const int* Z, const char* basis) {
  if (n_atoms <= 0 || !Z || !basis) return CCSD_ERR_ARG;
  int n = 0;
  for (int a = 0; a < n_atoms; ++a) {
    std::vector<ShellDef> sh = element_shells(Z[a], basis);
    if (sh.empty()) return CCSD_ERR_ARG;
    for (const ShellDef& s : sh) n += (s.L + 1) * (s.L + 2) / 2;
  }
  return n;
}

int ccsd_ao_one_electron(int n_atoms, const int* Z, const double* R_bohr, const char* basis, double* S, double* T,
                         double* V, int cap_n) {
  Molecule mol;
  int rc = molecule_from(n_atoms, Z, R_bohr, mol);
  if (rc != CCSD_OK || !basis) return CCSD_ERR_ARG;
  std::vector<BasisFn> fns;
  rc = build_basis(mol, basis, fns);
  if (rc != CCSD_OK) return rc;
  const int n = (int)fns.size();
  if (n > cap_n) return CCSD_ERR_ARG;
  std::vector<double> s, t, v;
  one_electron(mol, fns, s, t, v);
  if (S) std::memcpy(S, s.data(), s.size() * 8);
  if (T) std::memcpy(T, t.data(), t.size() * 8);
  if (V) std::memcpy(V, v.data(), v.size() * 8);
  return CCSD_OK;
}

int ccsd_rhf(int n_atoms, const int* Z, const double* R_bohr, const char* basis, int n_electrons, double tol,
             ccsd_rhf_result* out, double* C, double* eps, double* mo_eri, double* ao_eri, int cap_n) {
  Molecule mol;
  if (!basis || !out || !(tol > 0)) return CCSD_ERR_ARG;
  int rc = molecule_from(n_atoms, Z, R_bohr, mol);
  if (rc != CCSD_OK) return rc;
  const int nb = ccsd_basis_size(n_atoms, Z, basis);
  if (nb < 0) return nb;
  if ((C || eps || mo_eri || ao_eri) && cap_n < nb) return CCSD_ERR_ARG;
  ScfOut s;
  rc = run_scf(mol, basis, n_electrons, tol, false, s);
  if (rc != CCSD_OK && rc != CCSD_ERR_NOT_CONVERGED) return rc;
  out->e_hf = s.e_hf;
  out->e_nuc = s.e_nuc;
  out->commutator = s.comm;
  out->n_basis = nb;
  out->n_occ = s.nocc;
  out->iterations = s.iterations;
  out->converged = s.converged;
  if (C) std::memcpy(C, s.C.data(), s.C.size() * 8);
  if (eps) std::memcpy(eps, s.eps.data(), s.eps.size() * 8);
  if (ao_eri) std::memcpy(ao_eri, s.eri.data(), s.eri.size() * 8);
  if (mo_eri) mo_transform(nb, s.C, s.eri, mo_eri);
  return rc;
}

}  // extern "C"

RHF_Hartree_Fock::RHF_Hartree_Fock(int n_N, ccsd_compat::vec Z, ccsd_compat::mat R, std::string basis, int n_elec,
                                   bool print_stuff, int, int, bool)
    : n_atoms_(n_N), n_elec_(n_elec), basis_(basis), print_(print_stuff) {
  for (int a = 0; a < n_N; ++a) {
    Z_.push_back((int)std::lround(Z(a)));
    for (int x = 0; x < 3; ++x) R_.push_back(R(a, x));
  }
}

double RHF_Hartree_Fock::get_Energy(double toler, int do_uhf) {
  if (do_uhf != 0) {  // closed-shell only; the reference's UHF branch is outside this path
    status_ = CCSD_ERR_ARG;
    return std::nan("");
  }
  Molecule mol;
  status_ = molecule_from(n_atoms_, Z_.data(), R_.data(), mol);
  if (status_ != CCSD_OK) return std::nan("");
  ScfOut s;
  status_ = run_scf(mol, basis_, n_elec_, toler > 0 ? toler : 1e-10, print_, s);
  res_.e_hf = s.e_hf;
  res_.e_nuc = s.e_nuc;
  res_.commutator = s.comm;
  res_.n_basis = (int)s.eps.size();
  res_.n_occ = s.nocc;
  res_.iterations = s.iterations;
  res_.converged = s.converged;
  C_ = s.C;
  eps_ = s.eps;
  ao_eri_.swap(s.eri);
  return (status_ == CCSD_OK || status_ == CCSD_ERR_NOT_CONVERGED) ? s.e_hf : std::nan("");
}

int RHF_Hartree_Fock::ReturnMatrixSize() { return (int)eps_.size(); }

ccsd_compat::mat RHF_Hartree_Fock::ReturnC() {
  const int n = (int)eps_.size();
  ccsd_compat::mat M(n, n);
  for (int mu = 0; mu < n; ++mu)
    for (int p = 0; p < n; ++p) M(mu, p) = C_[(size_t)mu * n + p];
  return M;
}

ccsd_compat::mat RHF_Hartree_Fock::return_eigval_F() {
  const int n = (int)eps_.size();
  ccsd_compat::mat M(n, 1);
  for (int p = 0; p < n; ++p) M(p, 0) = eps_[p];
  return M;
}

ccsd_compat::mat RHF_Hartree_Fock::Return_Field_Q(int j, int l) {
  const int n = (int)eps_.size();
  ccsd_compat::mat M(n, n);
  if (j < 0 || l < 0 || j >= n || l >= n) return M;
  for (int p = 0; p < n; ++p)
    for (int q = 0; q < n; ++q) M(p, q) = ao_eri_[(((size_t)p * n + q) * n + j) * n + l];
  return M;
}
