// Contraction planner and the closed-shell CCSD term sequence (host code).
//
// Planner: a contraction "A_labels,B_labels->C_labels" over strided views
// (up to kMaxDims dimensions each) becomes one call of the strided, batched
// GEMM engine (gemm.hpp).  Labels in A, B and C are batch indices, labels in
// A and C only form M, labels in B and C only form N, labels in A and B only
// form K (summed).  Each group must collapse to one stride in every operand
// that carries it (consecutive labels with stride[t] = stride[t+1] *
// extent[t+1]); batch labels may form up to two such runs (the engine's two
// batch dimensions).  An operand that does not collapse is first copied by a
// strided permute into scratch in the canonical order (A: batch, M, K; B:
// batch, K, N; C: batch, M, N), and a permuted C is copied back afterwards.
// The closed-shell program keeps its large static integrals in layouts that
// need no permute; only amplitude-sized operands are ever permuted.
//
// The term sequence (rccsd_rhs) is a template over an executor with
//   int ein(spec, alpha, A, B, beta, C)   C = alpha A.B + beta C
//   int axpy(spec, alpha, A, beta, C)     C = alpha perm(A) + beta C
//   int tau(t1, t2, tau, o, v)            tau = t2 + t1 (x) t1
//   int vpm(VVVV, Vp, Vm, v)               the pair-packed ladder integrals
//   int ladder_pack(tau, tp, tm, o, v)     symmetric / antisymmetric tau pairs
//   int ladder_unpack(Sl, Al, R2, o, v)    R2 += S + A unpacked
//   int xpm(ovvv, Xp, Xm, o, v)            the pair-packed ladder-t1 integrals
//   int xk_unpack(XS, XA, Xk, o, v)        Xk = XS + sgn XA unpacked
// so that the same sequence runs on the device (rccsd.hip) and, as test
// infrastructure, on host arrays.
#pragma once
#include <cstring>
#include <string>
#include <vector>

#include "gemm_desc.hpp"

namespace ccsd {
namespace rplan {

constexpr int kMaxDims = 6;

// A strided view of a double array: element (i_0, .., i_{nd-1}) at p + sum i_t str[t].
struct View {
  double* p = nullptr;
  int nd = 0;
  long ext[kMaxDims] = {};
  long str[kMaxDims] = {};
};

inline View contig(double* p, std::initializer_list<long> e) {
  View v;
  v.p = p;
  v.nd = (int)e.size();
  int t = 0;
  for (long x : e) v.ext[t++] = x;
  long s = 1;
  for (int d = v.nd - 1; d >= 0; --d) {
    v.str[d] = s;
    s *= v.ext[d];
  }
  return v;
}

inline long numel(const View& v) {
  long n = 1;
  for (int t = 0; t < v.nd; ++t) n *= v.ext[t];
  return n;
}

// dst = alpha * src + beta * dst over identical shapes (dimension t of src
// corresponds to dimension t of dst); beta == 0 writes without reading dst.
struct Copy {
  View src, dst;
  double alpha = 1.0, beta = 0.0;
};

struct Plan {
  std::vector<Copy> pre;   // operand permutes into scratch (and C's current value)
  GemmDesc<double> g;
  bool has_gemm = false;
  std::vector<Copy> post;  // scratch C back into C
};

// Scratch the planner may use for permuted operands, one buffer per operand.
struct Scratch {
  double* p[3] = {nullptr, nullptr, nullptr};
  long cap[3] = {0, 0, 0};  // elements
};

namespace detail {

struct Lab {
  char s[kMaxDims + 1] = {};
  int n = 0;
  int find(char c) const {
    for (int t = 0; t < n; ++t)
      if (s[t] == c) return t;
    return -1;
  }
};

inline bool parse(const char* spec, Lab* ops, int nops_in, std::string* err) {
  // "ab,cd->ef" (nops_in = 2) or "ab->ba" (nops_in = 1)
  int k = 0;
  const char* q = spec;
  while (*q) {
    if (*q == ',') {
      if (++k >= nops_in) break;
      ++q;
      continue;
    }
    if (*q == '-' && q[1] == '>') {
      if (k != nops_in - 1) break;
      k = nops_in;
      q += 2;
      continue;
    }
    if (ops[k].n >= kMaxDims) {
      *err = std::string("too many labels in ") + spec;
      return false;
    }
    if (ops[k].find(*q) >= 0) {
      *err = std::string("repeated label in ") + spec;
      return false;
    }
    ops[k].s[ops[k].n++] = *q;
    ++q;
  }
  if (k != nops_in) {
    *err = std::string("bad spec ") + spec;
    return false;
  }
  return true;
}

// Collapse the labels g (in this order) of an operand to one index: ok when
// consecutive labels nest (stride[t] = stride[t+1] * ext[t+1]); labels of
// extent 1 are ignored.  Returns the stride of the collapsed index (0 if empty).
inline bool collapse(const View& v, const Lab& lab, const std::vector<char>& g, long* stride) {
  long prev_str = -1, prev_ext = 1;
  *stride = 0;
  for (char c : g) {
    const int t = lab.find(c);
    if (t < 0) return false;
    if (v.ext[t] == 1) continue;
    if (prev_str >= 0 && prev_str != v.str[t] * v.ext[t]) return false;
    prev_str = v.str[t];
    prev_ext = v.ext[t];
    *stride = v.str[t];
  }
  (void)prev_ext;
  return true;
}

inline long extent_of(const std::vector<char>& g, const Lab* labs, const View* vs, int nops) {
  long e = 1;
  for (char c : g)
    for (int k = 0; k < nops; ++k) {
      const int t = labs[k].find(c);
      if (t >= 0) {
        e *= vs[k].ext[t];
        break;
      }
    }
  return e;
}

// The view `v` reindexed to the label order `order` (a permutation of its labels).
inline View reorder(const View& v, const Lab& lab, const std::vector<char>& order) {
  View r;
  r.p = v.p;
  r.nd = (int)order.size();
  for (int t = 0; t < r.nd; ++t) {
    const int s = lab.find(order[t]);
    r.ext[t] = v.ext[s];
    r.str[t] = v.str[s];
  }
  return r;
}

inline void set_labels(Lab* lab, const std::vector<char>& order) {
  lab->n = (int)order.size();
  for (int t = 0; t < lab->n; ++t) lab->s[t] = order[t];
  lab->s[lab->n] = 0;
}

// Split the batch labels (C order) into at most two runs that collapse in A, B
// and C at once; false when more runs would be needed.
inline bool batch_runs(const std::vector<char>& bl, const Lab* labs, const View* vs,
                       std::vector<std::vector<char>>* runs) {
  runs->clear();
  std::vector<char> cur;
  for (char c : bl) {
    std::vector<char> trial = cur;
    trial.push_back(c);
    bool ok = true;
    for (int k = 0; k < 3 && ok; ++k) {
      long s;
      ok = collapse(vs[k], labs[k], trial, &s);
    }
    if (ok) {
      cur = trial;
    } else {
      runs->push_back(cur);
      cur.assign(1, c);
    }
  }
  if (!cur.empty()) runs->push_back(cur);
  return runs->size() <= 2;
}

}  // namespace detail

// Plan C = alpha A.B + beta C (C must not alias A or B).  On failure returns
// false with *err set.
inline bool make_plan(const char* spec, double alpha, const View& A, const View& B, double beta, const View& C,
                      const Scratch& sc, Plan* out, std::string* err) {
  using namespace detail;
  Lab labs[3];
  if (!parse(spec, labs, 2, err)) return false;
  View vs[3] = {A, B, C};
  for (int k = 0; k < 3; ++k)
    if (labs[k].n != vs[k].nd) {
      *err = std::string("rank mismatch in ") + spec;
      return false;
    }
  // extents agree across operands
  for (int k = 0; k < 3; ++k)
    for (int t = 0; t < labs[k].n; ++t)
      for (int q = 0; q < 3; ++q) {
        const int u = labs[q].find(labs[k].s[t]);
        if (u >= 0 && vs[q].ext[u] != vs[k].ext[t]) {
          *err = std::string("extent mismatch in ") + spec;
          return false;
        }
      }
  std::vector<char> bl, ml, nl, kl, kl_b;
  for (int t = 0; t < labs[2].n; ++t) {
    const char c = labs[2].s[t];
    const bool a = labs[0].find(c) >= 0, b = labs[1].find(c) >= 0;
    if (a && b) bl.push_back(c);
    else if (a) ml.push_back(c);
    else if (b) nl.push_back(c);
    else {
      *err = std::string("output label in neither input: ") + spec;
      return false;
    }
  }
  for (int t = 0; t < labs[0].n; ++t) {
    const char c = labs[0].s[t];
    if (labs[2].find(c) >= 0) continue;
    if (labs[1].find(c) < 0) {
      *err = std::string("label summed in one operand only: ") + spec;
      return false;
    }
    kl.push_back(c);
  }
  for (int t = 0; t < labs[1].n; ++t) {
    const char c = labs[1].s[t];
    if (labs[2].find(c) < 0 && labs[0].find(c) < 0) {
      *err = std::string("label summed in one operand only: ") + spec;
      return false;
    }
    if (labs[2].find(c) < 0) kl_b.push_back(c);
  }
  // Label orders of the M, N and K groups: each as it appears in either
  // operand that carries it; the combination that permutes the fewest
  // elements wins (ties keep C's order, so a large static input stays put).
  auto order_in = [&](const std::vector<char>& g, int k) {
    std::vector<char> r;
    for (int t = 0; t < labs[k].n; ++t)
      for (char c : g)
        if (labs[k].s[t] == c) r.push_back(c);
    return r;
  };
  auto cost_for = [&](const std::vector<char>& mo, const std::vector<char>& no, const std::vector<char>& ko,
                      bool need[3]) {
    long s;
    need[0] = !(collapse(vs[0], labs[0], mo, &s) && collapse(vs[0], labs[0], ko, &s));
    need[1] = !(collapse(vs[1], labs[1], ko, &s) && collapse(vs[1], labs[1], no, &s));
    need[2] = !(collapse(vs[2], labs[2], mo, &s) && collapse(vs[2], labs[2], no, &s));
    std::vector<std::vector<char>> runs;
    if (!batch_runs(bl, labs, vs, &runs)) need[0] = need[1] = need[2] = true;
    long c = 0;
    for (int k = 0; k < 3; ++k)
      if (need[k]) c += numel(vs[k]);
    return c;
  };
  const std::vector<char> mos[2] = {ml, order_in(ml, 0)}, nos[2] = {nl, order_in(nl, 1)}, kos[2] = {kl, kl_b};
  long best = -1;
  bool need[3] = {false, false, false};
  std::vector<char> mb, nb, kb;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int k = 0; k < 2; ++k) {
        bool nd[3];
        const long c = cost_for(mos[a], nos[b], kos[k], nd);
        if (best < 0 || c < best) {
          best = c;
          mb = mos[a]; nb = nos[b]; kb = kos[k];
          for (int q = 0; q < 3; ++q) need[q] = nd[q];
        }
      }
  ml = mb; nl = nb; kl = kb;
  out->pre.clear();
  out->post.clear();
  // canonical orders of permuted operands
  std::vector<char> ord[3];
  ord[0] = bl; ord[0].insert(ord[0].end(), ml.begin(), ml.end()); ord[0].insert(ord[0].end(), kl.begin(), kl.end());
  ord[1] = bl; ord[1].insert(ord[1].end(), kl.begin(), kl.end()); ord[1].insert(ord[1].end(), nl.begin(), nl.end());
  ord[2] = bl; ord[2].insert(ord[2].end(), ml.begin(), ml.end()); ord[2].insert(ord[2].end(), nl.begin(), nl.end());
  for (int k = 0; k < 3; ++k) {
    if (!need[k]) continue;
    const long n = numel(vs[k]);
    if (n > sc.cap[k] || !sc.p[k]) {
      *err = std::string("scratch too small for a permuted operand in ") + spec;
      return false;
    }
    View src = reorder(vs[k], labs[k], ord[k]);
    View dst = src;
    dst.p = sc.p[k];
    long s = 1;
    for (int d = dst.nd - 1; d >= 0; --d) {
      dst.str[d] = s;
      s *= dst.ext[d];
    }
    if (k < 2 || beta != 0.0) out->pre.push_back(Copy{src, dst, 1.0, 0.0});
    if (k == 2) out->post.push_back(Copy{dst, src, 1.0, 0.0});
    vs[k] = dst;
    set_labels(&labs[k], ord[k]);
  }
  GemmDesc<double> g;
  long s;
  g.M = (int)extent_of(ml, labs, vs, 3);
  g.N = (int)extent_of(nl, labs, vs, 3);
  g.K = (int)extent_of(kl, labs, vs, 3);
  collapse(vs[0], labs[0], ml, &g.sAm);
  collapse(vs[0], labs[0], kl, &g.sAk);
  collapse(vs[1], labs[1], kl, &g.sBk);
  collapse(vs[1], labs[1], nl, &g.sBn);
  collapse(vs[2], labs[2], ml, &g.sCm);
  collapse(vs[2], labs[2], nl, &g.sCn);
  // unit-extent groups: any stride will do; give the engine its preferred unit stride
  if (g.M == 1) g.sAm = g.sCm = 1;
  if (g.N == 1) g.sBn = g.sCn = 1;
  if (g.K == 1) g.sAk = g.sBk = 1;
  std::vector<std::vector<char>> runs;
  if (!batch_runs(bl, labs, vs, &runs)) {
    *err = std::string("batch labels do not collapse in ") + spec;
    return false;
  }
  if (!runs.empty()) {
    g.nb2 = (int)extent_of(runs.back(), labs, vs, 3);
    collapse(vs[0], labs[0], runs.back(), &g.sA2);
    collapse(vs[1], labs[1], runs.back(), &g.sB2);
    collapse(vs[2], labs[2], runs.back(), &g.sC2);
    if (runs.size() == 2) {
      g.nb1 = (int)extent_of(runs[0], labs, vs, 3);
      collapse(vs[0], labs[0], runs[0], &g.sA1);
      collapse(vs[1], labs[1], runs[0], &g.sB1);
      collapse(vs[2], labs[2], runs[0], &g.sC1);
    }
  }
  (void)s;
  g.A = vs[0].p;
  g.B = vs[1].p;
  g.C = vs[2].p;
  g.alpha = alpha;
  g.beta = beta;
  out->g = g;
  out->has_gemm = true;
  return true;
}

// Plan C = alpha perm(A) + beta C for "A_labels->C_labels" (same label set).
inline bool make_axpy(const char* spec, double alpha, const View& A, double beta, const View& C, Copy* out,
                      std::string* err) {
  using namespace detail;
  Lab labs[2];
  if (!parse(spec, labs, 1, err)) return false;
  if (labs[0].n != A.nd || labs[1].n != C.nd || labs[0].n != labs[1].n) {
    *err = std::string("rank mismatch in ") + spec;
    return false;
  }
  std::vector<char> order(labs[1].s, labs[1].s + labs[1].n);
  for (char c : order)
    if (labs[0].find(c) < 0) {
      *err = std::string("label sets differ in ") + spec;
      return false;
    }
  View src = reorder(A, labs[0], order);
  for (int t = 0; t < C.nd; ++t)
    if (src.ext[t] != C.ext[t]) {
      *err = std::string("extent mismatch in ") + spec;
      return false;
    }
  *out = Copy{src, C, alpha, beta};
  return true;
}

// ---------------------------------------------------------------------------
// Closed-shell CCSD (spin-adapted, canonical RHF) in spatial orbitals.
// Amplitudes t1[i][a] = t_{i alpha}^{a alpha}, t2[i][j][a][b] =
// t_{i alpha j beta}^{a alpha b beta}; every static block is built once from
// the MO integrals (pq|rs) in the layout its contraction reads directly.

struct RStatic {
  // o^2 v^2
  double *ovov = nullptr;   // [k][c][l][d] = (kc|ld)
  double *Lov = nullptr;    // [k][c][l][d] = 2 (kc|ld) - (kd|lc)
  double *Lt = nullptr;     // [k][l][c][d] = Lov[k][c][l][d]
  double *oovv = nullptr;   // [k][i][a][c] = (ki|ac)
  double *ovovP = nullptr;  // [k][l][c][d] = (kc|ld)
  double *M1 = nullptr;     // [l][d][k][c] = (ld|kc) - 1/2 (lc|kd)
  double *ovovx = nullptr;  // [l][d][k][c] = (lc|kd)
  double *ovovT = nullptr;  // [i][j][a][b] = (ia|jb)
  double *oovvI = nullptr;  // [i][a][k][c] = (ki|ac)
  double *ovovK = nullptr;  // [k][i][a][c] = (kc|ia)
  // o^3 v, o^4
  double *ovoo = nullptr;   // [l][c][k][i] = (lc|ki)
  double *Y2 = nullptr;     // [i][l][k][c] = -2 (lc|ki) + (kc|li)
  double *oooo = nullptr;   // [k][i][l][j] = (ki|lj)
  // o v^3
  double *ovvv = nullptr;   // [k][d][a][c] = (kd|ac)
  // t1 part of the ladder over packed pairs: Xp[(ka)][(cd)] = (kd|ac) + (kc|ad)
  // (c <= d), Xm[(ka)][(cd)] = (kd|ac) - (kc|ad) (c < d); rows of pitch ldpair
  double *Xp = nullptr;     // [o v][ldpair(v(v+1)/2)]
  double *Xm = nullptr;     // [o v][ldpair(v(v-1)/2)]
  double *ovvvR = nullptr;  // [a][k][c][d] = (kc|ad)
  double *ovvvR2 = nullptr; // [a][k][c][d] = (kd|ac)
  double *ovvvY = nullptr;  // [k][c][d][a] = 2 (kd|ac) - (kc|ad)
  double *ovvvL = nullptr;  // [k][d][a][c] = 2 (kd|ac) - (kc|ad)
  // pp ladder over pair-packed indices (a <= b, c <= d / a < b, c < d):
  // Vp[(ab)][(cd)] = (ac|bd) + (ad|bc), Vm[(ab)][(cd)] = (ac|bd) - (ad|bc);
  // rows of pitch ldpair(columns)
  double *Vp = nullptr;     // [v(v+1)/2][ldpair(v(v+1)/2)]
  double *Vm = nullptr;     // [v(v-1)/2][ldpair(v(v-1)/2)]
};

// Per-iteration work arrays.
struct RWork {
  double *tau = nullptr;    // [i][j][a][b]
  double *Tr = nullptr;     // [i][a][j][b] = t2[i][j][a][b]
  double *Tx = nullptr;     // [i][a][l][d] = t2[i][l][d][a]
  double *Z = nullptr;      // [i][a][l][d] = 1/2 t2[i][l][d][a] + t1[i][d] t1[l][a]
  double *Wr1 = nullptr;    // [i][a][k][c] = Wvoov[a][k][i][c]
  double *Wr2 = nullptr;    // [i][a][k][c] = Wvovo[a][k][c][i]
  double *W12 = nullptr;    // 2 Wr1 - Wr2
  double *tA = nullptr;     // [i][a][j][b] ring products
  double *tB = nullptr;     // [i][b][j][a] ring products
  double *Y = nullptr;      // [i][j][a][b] scratch products
  double *H = nullptr;      // [k][i][b][j]
  double *Xk = nullptr;     // [k][i][j][a] (o^3 v)
  double *Woooo = nullptr;  // [k][l][i][j]
  double *Fov = nullptr, *Foo = nullptr, *Fvv = nullptr, *Loo = nullptr, *Lvv = nullptr, *u = nullptr;
  // symmetric / antisymmetric ladder: tp [o(o+1)/2][ldpair(v(v+1)/2)], tm [o(o-1)/2][ldpair(v(v-1)/2)],
  // results Sl [o(o+1)/2][v(v+1)/2], Al [o(o-1)/2][v(v-1)/2]
  double *tp = nullptr, *tm = nullptr, *Sl = nullptr, *Al = nullptr;
  // Xk over pairs: XS [o v][o(o+1)/2], XA [o v][o(o-1)/2]
  double *XS = nullptr, *XA = nullptr;
};

// Pair counts and pair-packed indices: p1(p, q) = q(q+1)/2 + p for p <= q,
// p0(p, q) = q(q-1)/2 + p for p < q.
inline long npair1(long n) { return n * (n + 1) / 2; }
inline long npair0(long n) { return n * (n - 1) / 2; }
// Row pitch of the pair-packed ladder arrays: a multiple of 8 elements, so that
// rows start 64-byte aligned and the GEMM engine can load element pairs.
inline long ldpair(long n) { return (n + 7) / 8 * 8; }

inline View rows(double* p, long m, long k, long ld) {
  View w;
  w.p = p;
  w.nd = 2;
  w.ext[0] = m; w.ext[1] = k;
  w.str[0] = ld; w.str[1] = 1;
  return w;
}

// Build every static block from a view of the MO integrals mo[p][q][r][s]
// (n^4, chemists' notation); correlated occupied orbitals [nf, nocc), virtuals
// [nocc, n).
template <class X>
int rccsd_build_static(X& x, double* mo, int n, int nocc, int nf, const RStatic& S) {
  const long o = nocc - nf, v = n - nocc;
  const long n3 = (long)n * n * n, n2 = (long)n * n;
  auto blk = [&](int r0, int r1, int r2, int r3, long e0, long e1, long e2, long e3) {
    View w;
    w.p = mo + r0 * n3 + r1 * n2 + (long)r2 * n + r3;
    w.nd = 4;
    w.ext[0] = e0; w.ext[1] = e1; w.ext[2] = e2; w.ext[3] = e3;
    w.str[0] = n3; w.str[1] = n2; w.str[2] = n; w.str[3] = 1;
    return w;
  };
  const View OVOV = blk(nf, nocc, nf, nocc, o, v, o, v);
  const View OOVV = blk(nf, nf, nocc, nocc, o, o, v, v);
  const View OVOO = blk(nf, nocc, nf, nf, o, v, o, o);
  const View OOOO = blk(nf, nf, nf, nf, o, o, o, o);
  const View OVVV = blk(nf, nocc, nocc, nocc, o, v, v, v);
  const View VVVV = blk(nocc, nocc, nocc, nocc, v, v, v, v);
  const View ovov = contig(S.ovov, {o, v, o, v}), Lov = contig(S.Lov, {o, v, o, v});
  const View ovoo = contig(S.ovoo, {o, v, o, o}), ovvv = contig(S.ovvv, {o, v, v, v});
  int rc;
#define RT(call) \
  if ((rc = (call)) != 0) return rc
  RT(x.axpy("kcld->kcld", 1.0, OVOV, 0.0, ovov));
  RT(x.axpy("kcld->kcld", 2.0, ovov, 0.0, Lov));
  RT(x.axpy("kdlc->kcld", -1.0, ovov, 1.0, Lov));
  RT(x.axpy("kcld->klcd", 1.0, Lov, 0.0, contig(S.Lt, {o, o, v, v})));
  RT(x.axpy("kiac->kiac", 1.0, OOVV, 0.0, contig(S.oovv, {o, o, v, v})));
  RT(x.axpy("kcld->klcd", 1.0, ovov, 0.0, contig(S.ovovP, {o, o, v, v})));
  RT(x.axpy("ldkc->ldkc", 1.0, ovov, 0.0, contig(S.M1, {o, v, o, v})));
  RT(x.axpy("lckd->ldkc", -0.5, ovov, 1.0, contig(S.M1, {o, v, o, v})));
  RT(x.axpy("lckd->ldkc", 1.0, ovov, 0.0, contig(S.ovovx, {o, v, o, v})));
  RT(x.axpy("iajb->ijab", 1.0, ovov, 0.0, contig(S.ovovT, {o, o, v, v})));
  RT(x.axpy("lcki->lcki", 1.0, OVOO, 0.0, ovoo));
  RT(x.axpy("lcki->ilkc", -2.0, ovoo, 0.0, contig(S.Y2, {o, o, o, v})));
  RT(x.axpy("kcli->ilkc", 1.0, ovoo, 1.0, contig(S.Y2, {o, o, o, v})));
  RT(x.axpy("kilj->kilj", 1.0, OOOO, 0.0, contig(S.oooo, {o, o, o, o})));
  RT(x.axpy("kdac->kdac", 1.0, OVVV, 0.0, ovvv));
  RT(x.xpm(ovvv, S.Xp, S.Xm, o, v));
  RT(x.axpy("kiac->iakc", 1.0, OOVV, 0.0, contig(S.oovvI, {o, v, o, v})));
  RT(x.axpy("kcia->kiac", 1.0, ovov, 0.0, contig(S.ovovK, {o, o, v, v})));
  RT(x.axpy("kcad->akcd", 1.0, ovvv, 0.0, contig(S.ovvvR, {v, o, v, v})));
  RT(x.axpy("kdac->akcd", 1.0, ovvv, 0.0, contig(S.ovvvR2, {v, o, v, v})));
  RT(x.axpy("kdac->kcda", 2.0, ovvv, 0.0, contig(S.ovvvY, {o, v, v, v})));
  RT(x.axpy("kcad->kcda", -1.0, ovvv, 1.0, contig(S.ovvvY, {o, v, v, v})));
  RT(x.axpy("kdac->kdac", 2.0, ovvv, 0.0, contig(S.ovvvL, {o, v, v, v})));
  RT(x.axpy("kcad->kdac", -1.0, ovvv, 1.0, contig(S.ovvvL, {o, v, v, v})));
  RT(x.vpm(VVVV, S.Vp, S.Vm, v));
#undef RT
  return 0;
}

// R1 [o][v] and R2 [o][o][v][v] (right-hand sides of D t = R) at t1, t2.
template <class X>
int rccsd_rhs(X& x, long o, long v, const RStatic& S, const RWork& W, double* t1p, double* t2p, double* R1p,
              double* R2p) {
  const View t1 = contig(t1p, {o, v}), t2 = contig(t2p, {o, o, v, v});
  const View R1 = contig(R1p, {o, v}), R2 = contig(R2p, {o, o, v, v});
  const View tau = contig(W.tau, {o, o, v, v});
  const View Tr = contig(W.Tr, {o, v, o, v}), Tx = contig(W.Tx, {o, v, o, v}), Z = contig(W.Z, {o, v, o, v});
  const View Wr1 = contig(W.Wr1, {o, v, o, v}), Wr2 = contig(W.Wr2, {o, v, o, v}), W12 = contig(W.W12, {o, v, o, v});
  const View tA = contig(W.tA, {o, v, o, v}), tB = contig(W.tB, {o, v, o, v}), Y = contig(W.Y, {o, o, v, v});
  const View H = contig(W.H, {o, o, v, o}), Xk = contig(W.Xk, {o, o, o, v}), Woooo = contig(W.Woooo, {o, o, o, o});
  const View Fov = contig(W.Fov, {o, v}), Foo = contig(W.Foo, {o, o}), Fvv = contig(W.Fvv, {v, v});
  const View Loo = contig(W.Loo, {o, o}), Lvv = contig(W.Lvv, {v, v}), u = contig(W.u, {v, v});
  const View ovov = contig(S.ovov, {o, v, o, v}), Lov = contig(S.Lov, {o, v, o, v}), Lt = contig(S.Lt, {o, o, v, v});
  const View oovv = contig(S.oovv, {o, o, v, v}), ovovP = contig(S.ovovP, {o, o, v, v});
  const View M1 = contig(S.M1, {o, v, o, v}), ovovx = contig(S.ovovx, {o, v, o, v});
  const View ovovT = contig(S.ovovT, {o, o, v, v}), ovoo = contig(S.ovoo, {o, v, o, o});
  const View Y2 = contig(S.Y2, {o, o, o, v}), oooo = contig(S.oooo, {o, o, o, o}), ovvv = contig(S.ovvv, {o, v, v, v});
  const View ovvvR = contig(S.ovvvR, {v, o, v, v});
  const View oovvI = contig(S.oovvI, {o, v, o, v}), ovovK = contig(S.ovovK, {o, o, v, v});
  const View ovvvR2 = contig(S.ovvvR2, {v, o, v, v}), ovvvY = contig(S.ovvvY, {o, v, v, v});
  const View ovvvL = contig(S.ovvvL, {o, v, v, v});
  int rc;
#define RT(call) \
  if ((rc = (call)) != 0) return rc
  RT(x.tau(t1p, t2p, W.tau, o, v));
  RT(x.axpy("ijab->iajb", 1.0, t2, 0.0, Tr));
  RT(x.axpy("ilda->iald", 1.0, t2, 0.0, Tx));
  // -- one-particle intermediates
  RT(x.ein("kcld,ld->kc", 1.0, Lov, t1, 0.0, Fov));
  RT(x.ein("klcd,ilcd->ki", 1.0, Lt, tau, 0.0, Foo));
  RT(x.ein("lkda,lkdc->ac", -1.0, tau, Lt, 0.0, Fvv));   // tau[k][l][a][d] = tau[l][k][d][a]
  RT(x.axpy("ki->ki", 1.0, Foo, 0.0, Loo));
  RT(x.ein("lcki,lc->ki", 2.0, ovoo, t1, 1.0, Loo));
  RT(x.ein("kcli,lc->ki", -1.0, ovoo, t1, 1.0, Loo));
  RT(x.axpy("ac->ac", 1.0, Fvv, 0.0, Lvv));
  RT(x.ein("kdac,kd->ac", 1.0, ovvvL, t1, 1.0, Lvv));
  // -- T1
  RT(x.ein("ac,ic->ia", 1.0, Fvv, t1, 0.0, R1));
  RT(x.ein("ki,ka->ia", -1.0, Foo, t1, 1.0, R1));
  RT(x.ein("kc,iakc->ia", 2.0, Fov, Tr, 1.0, R1));        // Tr[i][a][k][c] = t2[k][i][c][a]
  RT(x.ein("kc,iakc->ia", -1.0, Fov, Tx, 1.0, R1));       // Tx[i][a][k][c] = t2[i][k][c][a]
  RT(x.ein("kc,ka->ca", 1.0, Fov, t1, 0.0, u));
  RT(x.ein("ic,ca->ia", 1.0, t1, u, 1.0, R1));
  RT(x.ein("kcia,kc->ia", 2.0, ovov, t1, 1.0, R1));
  RT(x.ein("iakc,kc->ia", -1.0, oovvI, t1, 1.0, R1));
  RT(x.ein("ikcd,kcda->ia", 1.0, tau, ovvvY, 1.0, R1));
  RT(x.ein("ilkc,lkca->ia", 1.0, Y2, tau, 1.0, R1));      // tau[k][l][a][c] = tau[l][k][c][a]
  // -- T2: integrals, hh ladder, pp ladder
  RT(x.axpy("ijab->ijab", 1.0, ovovT, 0.0, R2));
  RT(x.axpy("kilj->klij", 1.0, oooo, 0.0, Woooo));
  RT(x.ein("lcki,jc->klij", 1.0, ovoo, t1, 1.0, Woooo));
  RT(x.ein("kclj,ic->klij", 1.0, ovoo, t1, 1.0, Woooo));
  RT(x.ein("klcd,ijcd->klij", 1.0, ovovP, tau, 1.0, Woooo));
  RT(x.ein("klij,klab->ijab", 1.0, Woooo, tau, 1.0, R2));
  // pp ladder sum_cd (ac|bd) tau_ij^cd = S + A over pairs: S_ij^ab = sum_{c<=d} Vp tp
  // (symmetric in a,b and in i,j), A_ij^ab = sum_{c<d} Vm tm (antisymmetric in
  // a,b and in i,j); tp carries tau_ij^cc / 2 on the diagonal c = d
  const long n1 = npair1(v), n0 = npair0(v), l1 = ldpair(n1), l0 = ldpair(n0);
  RT(x.ladder_pack(W.tau, W.tp, W.tm, o, v));
  RT(x.ein("pk,qk->pq", 1.0, rows(W.tp, npair1(o), n1, l1), rows(S.Vp, n1, n1, l1), 0.0,
           contig(W.Sl, {npair1(o), n1})));
  RT(x.ein("pk,qk->pq", 1.0, rows(W.tm, npair0(o), n0, l0), rows(S.Vm, n0, n0, l0), 0.0,
           contig(W.Al, {npair0(o), n0})));
  RT(x.ladder_unpack(W.Sl, W.Al, R2p, o, v));
  // t1 parts of the ladder and of (kc|ai) t: Xk[k][i][j][a] = sum_cd (kd|ac) tau_ij^cd
  //   + sum_c (kc|ia) t_jc + (ia|kj); then -P sum_k Xk t_kb.  The tau term runs
  //   over the ladder's packed pairs: XS = Xp tp^T (i <= j), XA = Xm tm^T (i < j)
  RT(x.ein("qk,pk->qp", 1.0, rows(S.Xp, o * v, n1, l1), rows(W.tp, npair1(o), n1, l1), 0.0,
           contig(W.XS, {o * v, npair1(o)})));
  RT(x.ein("qk,pk->qp", 1.0, rows(S.Xm, o * v, n0, l0), rows(W.tm, npair0(o), n0, l0), 0.0,
           contig(W.XA, {o * v, npair0(o)})));
  RT(x.xk_unpack(W.XS, W.XA, W.Xk, o, v));
  RT(x.ein("kiac,jc->kija", 1.0, ovovK, t1, 1.0, Xk));
  RT(x.axpy("iakj->kija", 1.0, ovoo, 1.0, Xk));
  // The P-symmetrised terms accumulate in Y, then R2 += Y + P Y once:
  // -Xk t1; (ai|bc) t_jc - t_ka (ki|bc) t_jc; Lvv, Loo
  RT(x.ein("kija,kb->ijab", -1.0, Xk, t1, 0.0, Y));
  RT(x.ein("iabc,jc->ijab", 1.0, ovvv, t1, 1.0, Y));
  RT(x.ein("kibc,jc->kibj", 1.0, oovv, t1, 0.0, H));
  RT(x.ein("ka,kibj->ijab", -1.0, t1, H, 1.0, Y));
  RT(x.ein("ijac,bc->ijab", 1.0, t2, Lvv, 1.0, Y));
  RT(x.ein("ki,kjab->ijab", -1.0, Loo, t2, 1.0, Y));
  RT(x.axpy("ijab->ijab", 1.0, Y, 1.0, R2));
  RT(x.axpy("jiba->ijab", 1.0, Y, 1.0, R2));
  // -- ring intermediates Wr1 = Wvoov[a][k][i][c], Wr2 = Wvovo[a][k][c][i] as [(i,a)][(k,c)]
  RT(x.axpy("iald->iald", 0.5, Tx, 0.0, Z));
  RT(x.ein("id,la->iald", 1.0, t1, t1, 1.0, Z));
  RT(x.axpy("iakc->iakc", 1.0, ovov, 0.0, Wr1));           // (ai|kc) = (ia|kc)
  RT(x.ein("id,akcd->iakc", 1.0, t1, ovvvR, 1.0, Wr1));
  RT(x.ein("la,kcli->iakc", -1.0, t1, ovoo, 1.0, Wr1));
  RT(x.ein("iald,ldkc->iakc", 1.0, Tr, M1, 1.0, Wr1));
  RT(x.ein("iald,ldkc->iakc", -1.0, Z, ovov, 1.0, Wr1));
  RT(x.axpy("kiac->iakc", 1.0, oovv, 0.0, Wr2));           // (ki|ac)
  RT(x.ein("id,akcd->iakc", 1.0, t1, ovvvR2, 1.0, Wr2));
  RT(x.ein("la,lcki->iakc", -1.0, t1, ovoo, 1.0, Wr2));
  RT(x.ein("iald,ldkc->iakc", -1.0, Z, ovovx, 1.0, Wr2));
  // -- ring products: tA[(i,a)][(j,b)], tB[(i,b)][(j,a)]
  RT(x.axpy("iakc->iakc", 2.0, Wr1, 0.0, W12));
  RT(x.axpy("iakc->iakc", -1.0, Wr2, 1.0, W12));
  RT(x.ein("iakc,kcjb->iajb", 1.0, W12, Tr, 0.0, tA));
  RT(x.ein("iakc,kcjb->iajb", -1.0, Wr1, Tx, 1.0, tA));    // Tx[k][c][j][b] = t2[k][j][b][c]
  RT(x.ein("ibkc,kcja->ibja", -1.0, Wr2, Tx, 0.0, tB));
  RT(x.axpy("iajb->ijab", 1.0, tA, 1.0, R2));
  RT(x.axpy("jbia->ijab", 1.0, tA, 1.0, R2));
  RT(x.axpy("ibja->ijab", 1.0, tB, 1.0, R2));
  RT(x.axpy("jaib->ijab", 1.0, tB, 1.0, R2));
#undef RT
  return 0;
}

// Element counts of the work arrays for (o, v) (in RWork order).
inline void rccsd_work_sizes(long o, long v, std::vector<long>* n) {
  const long t2 = o * o * v * v;
  *n = {t2, t2, t2, t2, t2, t2, t2, t2, t2, t2, o * o * v * o, o * o * o * v, o * o * o * o, o * v, o * o, v * v,
        o * o, v * v, v * v, npair1(o) * ldpair(npair1(v)), npair0(o) * ldpair(npair0(v)), npair1(o) * npair1(v),
        npair0(o) * npair0(v), o * v * npair1(o), o * v * npair0(o)};
}

}  // namespace rplan
}  // namespace ccsd
