// Host-code driver for the AddressSanitizer / UndefinedBehaviorSanitizer build
// (test infrastructure; built with g++ by tests/test_host_sanitizers.py).
//
// Runs, on small inputs, every host routine that sits beside the device path:
//   - the RHF producer (include/ccsd_rhf.h): STO-3G water to a converged SCF with
//     MO integrals, checked against Crawford's published E_SCF;
//   - the closed-shell term sequence through the contraction planner and the
//     host executor (tests/cpp/rccsd_cpu.cpp), with and without a frozen core
//     and with more occupied than virtual orbitals;
//   - single planned contractions, including the planner's error paths;
//   - the packed spin-orbital CPU residual (oracle/ccsd_cpu.cpp).
// Any invalid access or undefined behaviour aborts with the sanitizer's report.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/ccsd_rhf.h"

extern "C" {
int rcpu_rhs(int n, int nocc, int nf, const double* mo, const double* t1, const double* t2, double* R1, double* R2,
             long* permuted);
int rcpu_ein(const char* spec, double alpha, double* A, int na, const long* ea, const long* sa, double* B, int nb,
             const long* eb, const long* sb, double beta, double* Cp, int nc, const long* ec, const long* scn,
             long scratch, char* err, int cap);
int cpu_ccsd_residual(int O_, int V_, const double* fock, const double* oooo, const double* ooov,
                      const double* oovv, const double* ovov, const double* ovvv, const double* vvvv,
                      const double* t1, const double* t2p, double* r1, double* r2p, int threads);
}

namespace {

unsigned long long g_state = 0x9e3779b97f4a7c15ull;
double urand() {  // splitmix64 in [-0.5, 0.5)
  unsigned long long z = (g_state += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (double)(z >> 11) / 9007199254740992.0 - 0.5;
}
std::vector<double> rnd(size_t n, double s) {
  std::vector<double> v(n);
  for (auto& x : v) x = s * urand();
  return v;
}

int fail(const char* what) {
  fprintf(stderr, "FAIL: %s\n", what);
  return 1;
}

// (pq|rs) with the 8-fold permutational symmetry of real orbitals
std::vector<double> random_mo(int n) {
  std::vector<double> mo((size_t)n * n * n * n);
  for (int p = 0; p < n; ++p)
    for (int q = 0; q <= p; ++q)
      for (int r = 0; r < n; ++r)
        for (int s = 0; s <= r; ++s) {
          if (p * n + q < r * n + s) continue;
          const double v = 0.05 * urand();
          const int idx[8][4] = {{p, q, r, s}, {q, p, r, s}, {p, q, s, r}, {q, p, s, r},
                                 {r, s, p, q}, {s, r, p, q}, {r, s, q, p}, {s, r, q, p}};
          for (auto& t : idx) mo[(((size_t)t[0] * n + t[1]) * n + t[2]) * n + t[3]] = v;
        }
  return mo;
}

int closed_shell(int n, int nocc, int nf, const std::vector<double>& mo) {
  const int o = nocc - nf, v = n - nocc;
  auto t1 = rnd((size_t)o * v, 0.02);
  std::vector<double> t2((size_t)o * o * v * v);
  for (int i = 0; i < o; ++i)
    for (int j = 0; j < o; ++j)
      for (int a = 0; a < v; ++a)
        for (int b = 0; b < v; ++b) {  // t2[i][j][a][b] = t2[j][i][b][a]
          const size_t e = (((size_t)i * o + j) * v + a) * v + b, f = (((size_t)j * o + i) * v + b) * v + a;
          if (e <= f) t2[e] = t2[f] = 0.02 * urand();
        }
  std::vector<double> R1((size_t)o * v), R2((size_t)o * o * v * v);
  long permuted = 0;
  if (rcpu_rhs(n, nocc, nf, mo.data(), t1.data(), t2.data(), R1.data(), R2.data(), &permuted) != 0) return 1;
  for (double x : R2)
    if (!std::isfinite(x)) return 1;
  return 0;
}

}  // namespace

int main() {
  // -- RHF: STO-3G water (Crawford's geometry), MO integrals
  const int Z[3] = {8, 1, 1};
  const double R[9] = {0.0, -0.143225816552, 0.0, 1.638036840407, 1.136548822547, 0.0,
                       -1.638036840407, 1.136548822547, 0.0};
  const int n = ccsd_basis_size(3, Z, "STO-3G");
  if (n != 7) return fail("basis size");
  std::vector<double> C((size_t)n * n), eps(n), mo((size_t)n * n * n * n), ao((size_t)n * n * n * n);
  ccsd_rhf_result res;
  if (ccsd_rhf(3, Z, R, "STO-3G", 10, 1e-10, &res, C.data(), eps.data(), mo.data(), ao.data(), n) != 0)
    return fail("ccsd_rhf");
  if (std::fabs(res.e_hf - (-74.942079928192)) > 1e-8 || !res.converged) return fail("STO-3G water E_SCF");
  std::vector<double> S((size_t)n * n), T((size_t)n * n), V((size_t)n * n);
  if (ccsd_ao_one_electron(3, Z, R, "STO-3G", S.data(), T.data(), V.data(), n) != 0) return fail("one-electron");
  if (ccsd_basis_size(3, Z, "no-such-basis") >= 0) return fail("unknown basis accepted");

  // -- closed-shell term sequence: water (o > v), frozen core, random integrals
  if (closed_shell(n, 5, 0, mo)) return fail("closed shell, water");
  if (closed_shell(n, 5, 1, mo)) return fail("closed shell, water, frozen core");
  {
    const auto m9 = random_mo(9);
    if (closed_shell(9, 4, 1, m9)) return fail("closed shell, n=9");
    const auto m8 = random_mo(8);
    if (closed_shell(8, 3, 0, m8)) return fail("closed shell, n=8");
  }

  // -- single planned contractions (strided views) and the planner's refusals
  {
    auto A = rnd(6 * 11, 1.0), B = rnd(7 * 3, 1.0), Cm = rnd(6 * 8, 1.0);
    const long ea[2] = {6, 7}, sa[2] = {11, 1}, eb[2] = {7, 3}, sb[2] = {1, 7}, ec[2] = {6, 3}, sc[2] = {8, 1};
    char err[256];
    if (rcpu_ein("ik,kj->ij", 2.0, A.data() + 2, 2, ea, sa, B.data(), 2, eb, sb, 1.0, Cm.data() + 1, 2, ec, sc,
                 1 << 12, err, 256) != 0)
      return fail("planned contraction");
    if (rcpu_ein("ik,kj->ijz", 1.0, A.data(), 2, ea, sa, B.data(), 2, eb, sb, 0.0, Cm.data(), 2, ec, sc, 1 << 12,
                 err, 256) == 0)
      return fail("bad spec accepted");
    auto X = rnd(3 * 4 * 5 * 4, 1.0), Y = rnd(2 * 5 * 4 * 4, 1.0), W = rnd(3 * 2, 1.0);
    const long ex[4] = {3, 4, 5, 4}, sx[4] = {80, 20, 4, 1}, ey[4] = {2, 5, 4, 4}, sy[4] = {80, 16, 4, 1};
    const long ew[2] = {3, 2}, sw[2] = {2, 1};
    if (rcpu_ein("kcld,ilcd->ki", 1.0, X.data(), 4, ex, sx, Y.data(), 4, ey, sy, 0.0, W.data(), 2, ew, sw, 1 << 12,
                 err, 256) != 0)
      return fail("permuted contraction");
    if (rcpu_ein("kcld,ilcd->ki", 1.0, X.data(), 4, ex, sx, Y.data(), 4, ey, sy, 0.0, W.data(), 2, ew, sw, 4, err,
                 256) == 0)
      return fail("scratch overflow not refused");
  }

  // -- packed spin-orbital CPU residual
  {
    const long O = 4, Vv = 6, N = O + Vv, po = O * (O - 1) / 2, pv = Vv * (Vv - 1) / 2, ov = O * Vv;
    auto f = rnd(N * N, 0.1);
    for (long p = 0; p < N; ++p) f[p * N + p] = p < O ? -1.0 - 0.1 * p : 0.5 + 0.1 * p;
    auto oooo = rnd(po * po, 0.05), ooov = rnd(po * ov, 0.05), oovv = rnd(po * pv, 0.05);
    auto ovov = rnd(ov * ov, 0.05), ovvv = rnd(ov * pv, 0.05), vvvv = rnd(pv * pv, 0.05);
    auto t1 = rnd(ov, 0.02), t2 = rnd(po * pv, 0.02);
    std::vector<double> r1(ov), r2(po * pv);
    if (cpu_ccsd_residual((int)O, (int)Vv, f.data(), oooo.data(), ooov.data(), oovv.data(), ovov.data(),
                          ovvv.data(), vvvv.data(), t1.data(), t2.data(), r1.data(), r2.data(), 2) != 0)
      return fail("cpu residual");
    if (cpu_ccsd_residual(1, (int)Vv, f.data(), oooo.data(), ooov.data(), oovv.data(), ovov.data(), ovvv.data(),
                          vvvv.data(), t1.data(), t2.data(), r1.data(), r2.data(), 2) == 0)
      return fail("cpu residual accepted O = 1");
  }
  printf("sanitized host run: ok\n");
  return 0;
}
