// Host executor for the closed-shell CCSD term sequence (test infrastructure).
//
// Runs the contraction planner and the term sequence of
// ccsd_parallel_amd/csrc/rplan.hpp on host arrays: every planned permute is a
// strided loop, every planned GEMM descriptor is evaluated by plain loops over
// its batches, rows, columns and k.  The tests compare the result with the
// numpy closed-shell oracle, which checks the planner's strides and the term
// sequence without a GPU.  Built by tests/test_rccsd_oracle.py with g++.
#include <cstdio>
#include <vector>

#include "../../ccsd_parallel_amd/csrc/rplan.hpp"

using namespace ccsd::rplan;

namespace {

void run_copy(const Copy& c) {
  const long n = numel(c.dst);
  long idx[kMaxDims] = {};
  for (long e = 0; e < n; ++e) {
    long r = e, so = 0, dof = 0;
    for (int t = c.dst.nd - 1; t >= 0; --t) {
      idx[t] = r % c.dst.ext[t];
      r /= c.dst.ext[t];
      so += idx[t] * c.src.str[t];
      dof += idx[t] * c.dst.str[t];
    }
    const double v = c.alpha * c.src.p[so];
    c.dst.p[dof] = (c.beta != 0.0) ? v + c.beta * c.dst.p[dof] : v;
  }
}

void run_gemm(const ccsd::GemmDesc<double>& g) {
  for (int b1 = 0; b1 < g.nb1; ++b1)
    for (int b2 = 0; b2 < g.nb2; ++b2) {
      const double* A = g.A + b1 * g.sA1 + b2 * g.sA2;
      const double* B = g.B + b1 * g.sB1 + b2 * g.sB2;
      double* C = g.C + b1 * g.sC1 + b2 * g.sC2;
      for (int m = 0; m < g.M; ++m)
        for (int n = 0; n < g.N; ++n) {
          double s = 0.0;
          for (int k = 0; k < g.K; ++k) s += A[m * g.sAm + k * g.sAk] * B[k * g.sBk + n * g.sBn];
          double* c = C + m * g.sCm + n * g.sCn;
          *c = g.alpha * s + (g.beta != 0.0 ? g.beta * *c : 0.0);
        }
    }
}

struct CpuExec {
  Scratch sc;
  std::vector<double> buf[3];
  std::string err;
  long permuted = 0;  // elements copied by planned permutes

  explicit CpuExec(long cap) {
    for (int k = 0; k < 3; ++k) {
      buf[k].assign(cap, 0.0);
      sc.p[k] = buf[k].data();
      sc.cap[k] = cap;
    }
  }
  int ein(const char* spec, double alpha, const View& A, const View& B, double beta, const View& C) {
    Plan p;
    if (!make_plan(spec, alpha, A, B, beta, C, sc, &p, &err)) return -1;
    for (const Copy& c : p.pre) {
      run_copy(c);
      permuted += numel(c.dst);
    }
    run_gemm(p.g);
    for (const Copy& c : p.post) run_copy(c);
    return 0;
  }
  int axpy(const char* spec, double alpha, const View& A, double beta, const View& C) {
    Copy c;
    if (!make_axpy(spec, alpha, A, beta, C, &c, &err)) return -1;
    run_copy(c);
    return 0;
  }
  static long p1(long p, long q) { return p <= q ? q * (q + 1) / 2 + p : p * (p + 1) / 2 + q; }
  static long p0(long p, long q) { return p < q ? q * (q - 1) / 2 + p : p * (p - 1) / 2 + q; }
  int vpm(const View& W, double* Vp, double* Vm, long v) {
    auto at = [&](long a, long b, long c, long d) {
      return W.p[a * W.str[0] + b * W.str[1] + c * W.str[2] + d * W.str[3]];
    };
    for (long b = 0; b < v; ++b)
      for (long a = 0; a <= b; ++a)
        for (long d = 0; d < v; ++d)
          for (long c = 0; c <= d; ++c) {
            const double x = at(a, c, b, d), y = at(a, d, b, c);
            Vp[p1(a, b) * ldpair(npair1(v)) + p1(c, d)] = x + y;
            if (a < b && c < d) Vm[p0(a, b) * ldpair(npair0(v)) + p0(c, d)] = x - y;
          }
    return 0;
  }
  int ladder_pack(const double* tau, double* tp, double* tm, long o, long v) {
    const long n1 = ldpair(npair1(v)), n0 = ldpair(npair0(v));
    for (long j = 0; j < o; ++j)
      for (long i = 0; i <= j; ++i)
        for (long d = 0; d < v; ++d)
          for (long c = 0; c <= d; ++c) {
            const double x = tau[((i * o + j) * v + c) * v + d], y = tau[((i * o + j) * v + d) * v + c];
            tp[p1(i, j) * n1 + p1(c, d)] = 0.5 * (x + y) * (c == d ? 0.5 : 1.0);
            if (i < j && c < d) tm[p0(i, j) * n0 + p0(c, d)] = 0.5 * (x - y);
          }
    return 0;
  }
  int ladder_unpack(const double* Sl, const double* Al, double* R2, long o, long v) {
    const long n1 = v * (v + 1) / 2, n0 = v * (v - 1) / 2;
    for (long i = 0; i < o; ++i)
      for (long j = 0; j < o; ++j)
        for (long a = 0; a < v; ++a)
          for (long b = 0; b < v; ++b) {
            double x = Sl[p1(i, j) * n1 + p1(a, b)];
            if (i != j && a != b) x += ((i < j) == (a < b) ? 1.0 : -1.0) * Al[p0(i, j) * n0 + p0(a, b)];
            R2[((i * o + j) * v + a) * v + b] += x;
          }
    return 0;
  }
  int xpm(const View& w, double* Xp, double* Xm, long o, long v) {
    const long l1 = ldpair(npair1(v)), l0 = ldpair(npair0(v));
    auto at = [&](long k, long d, long a, long c) { return w.p[k * w.str[0] + d * w.str[1] + a * w.str[2] + c * w.str[3]]; };
    for (long k = 0; k < o; ++k)
      for (long a = 0; a < v; ++a)
        for (long d = 0; d < v; ++d)
          for (long c = 0; c <= d; ++c) {
            const double x = at(k, d, a, c), y = at(k, c, a, d);
            Xp[(k * v + a) * l1 + p1(c, d)] = x + y;
            if (c < d) Xm[(k * v + a) * l0 + p0(c, d)] = x - y;
          }
    return 0;
  }
  int xk_unpack(const double* XS, const double* XA, double* Xk, long o, long v) {
    const long m1 = npair1(o), m0 = npair0(o);
    for (long k = 0; k < o; ++k)
      for (long i = 0; i < o; ++i)
        for (long j = 0; j < o; ++j)
          for (long a = 0; a < v; ++a) {
            double x = XS[(k * v + a) * m1 + p1(i, j)];
            if (i != j) x += (i < j ? 1.0 : -1.0) * XA[(k * v + a) * m0 + p0(i, j)];
            Xk[((k * o + i) * o + j) * v + a] = x;
          }
    return 0;
  }
  int tau(const double* t1, const double* t2, double* tau, long o, long v) {
    for (long i = 0; i < o; ++i)
      for (long j = 0; j < o; ++j)
        for (long a = 0; a < v; ++a)
          for (long b = 0; b < v; ++b) {
            const long e = ((i * o + j) * v + a) * v + b;
            tau[e] = t2[e] + t1[i * v + a] * t1[j * v + b];
          }
    return 0;
  }
};

}  // namespace

extern "C" {

// R1 [o][v], R2 [o][o][v][v] of the closed-shell equations at (t1, t2) for the
// MO integrals mo [n^4] (correlated occupied [nf, nocc), virtual [nocc, n)).
// Returns 0, or -1 with the planner's message printed.  *permuted receives
// the number of elements the planned operand permutes copied.
int rcpu_rhs(int n, int nocc, int nf, const double* mo, const double* t1, const double* t2, double* R1, double* R2,
             long* permuted) {
  const long o = nocc - nf, v = n - nocc;
  const long t2n = o * o * v * v, ov3 = o * v * v * v;
  std::vector<double> st[22];
  const long sizes[22] = {t2n, t2n, t2n, t2n, t2n, t2n, t2n, t2n, t2n, t2n, o * v * o * o, o * o * o * v,
                          o * o * o * o, ov3, ov3, ov3, ov3, ov3, npair1(v) * ldpair(npair1(v)),
                          npair0(v) * ldpair(npair0(v)), o * v * ldpair(npair1(v)), o * v * ldpair(npair0(v))};
  for (int k = 0; k < 22; ++k) st[k].assign(sizes[k], 0.0);
  RStatic S;
  double** sp[22] = {&S.ovov, &S.Lov, &S.Lt, &S.oovv, &S.ovovP, &S.M1, &S.ovovx, &S.ovovT, &S.oovvI, &S.ovovK,
                     &S.ovoo, &S.Y2, &S.oooo, &S.ovvv, &S.ovvvR, &S.ovvvR2, &S.ovvvY, &S.ovvvL, &S.Vp, &S.Vm,
                     &S.Xp, &S.Xm};
  for (int k = 0; k < 22; ++k) *sp[k] = st[k].data();
  std::vector<long> wn;
  rccsd_work_sizes(o, v, &wn);
  std::vector<std::vector<double>> wk(wn.size());
  for (size_t k = 0; k < wn.size(); ++k) wk[k].assign(wn[k], 0.0);
  RWork W;
  double** wp[] = {&W.tau, &W.Tr, &W.Tx, &W.Z, &W.Wr1, &W.Wr2, &W.W12, &W.tA, &W.tB, &W.Y, &W.H, &W.Xk, &W.Woooo,
                   &W.Fov, &W.Foo, &W.Fvv, &W.Loo, &W.Lvv, &W.u, &W.tp, &W.tm, &W.Sl, &W.Al, &W.XS, &W.XA};
  for (size_t k = 0; k < wn.size(); ++k) *wp[k] = wk[k].data();
  CpuExec x(t2n + o * o * o * v + o * o * o * o);  // the device context's scratch size
  std::vector<double> mo_c(mo, mo + (long)n * n * n * n), t1c(t1, t1 + o * v), t2c(t2, t2 + t2n);
  int rc = rccsd_build_static(x, mo_c.data(), n, nocc, nf, S);
  if (rc == 0) rc = rccsd_rhs(x, o, v, S, W, t1c.data(), t2c.data(), R1, R2);
  if (rc != 0) fprintf(stderr, "rcpu_rhs: %s\n", x.err.c_str());
  if (permuted) *permuted = x.permuted;
  return rc;
}

// One contraction C = alpha A.B + beta C through the planner and the host
// executor; each operand is a strided view (nd, ext[nd], str[nd]) of its
// array.  Returns 0, or -1 with the planner's message in err (cap bytes).
int rcpu_ein(const char* spec, double alpha, double* A, int na, const long* ea, const long* sa, double* B, int nb,
             const long* eb, const long* sb, double beta, double* Cp, int nc, const long* ec, const long* scn,
             long scratch, char* err, int cap) {
  auto view = [](double* p, int nd, const long* e, const long* st) {
    View w;
    w.p = p;
    w.nd = nd;
    for (int t = 0; t < nd; ++t) {
      w.ext[t] = e[t];
      w.str[t] = st[t];
    }
    return w;
  };
  CpuExec x(scratch);
  const int rc = x.ein(spec, alpha, view(A, na, ea, sa), view(B, nb, eb, sb), beta, view(Cp, nc, ec, scn));
  if (rc != 0 && err && cap > 0) snprintf(err, cap, "%s", x.err.c_str());
  return rc;
}

}  // extern "C"
