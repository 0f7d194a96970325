// CPU restatement of the reference's contraction loops (test and baseline
// infrastructure only: bench.py times it as the CPU baseline beside the GPU;
// nothing in the product links it).
//
// The reference evaluates every CCSD term as Armadillo element-wise products
// reduced with accu() over compact slices, distributed over MPI ranks by
// output index (ccsd_memory_optimized.cpp:2046-2048).  Two loop shapes carry
// almost all of its flops:
//  * ladder type: one output per (a<b, i<j) is a dot product over the packed
//    virtual pairs c<d, e.g. 1/2 sum_cd <ab||cd> tau_ij^cd (Fill_t2_new,
//    :2093, :2158, :2208, :2277, :2342, :2393); W_mnij, W_mbij and the hh
//    ladder have the same form;
//  * ring type: one output per ((i,a),(j,b)) is a dot product over (m,e),
//    W_5(a,b)(i,j) = sum_em t2(b,j)(e,m) W_4(a,i)(e,m) (Fill_W5, :6240), as is
//    the t2 part of W_mbej (Fill_W4_MPI, :4807-4809).
// Here OpenMP threads take the place of the MPI ranks and split the output
// rows the same way; each output is one dot product over contiguous rows, as
// accu(A % B) is.
#include <omp.h>

#include <cstddef>

extern "C" {

// out[r][q] = sum_k A[r][k] * B[q][k] for r < nr, q < nq (row-major, rows of length nk).
// Ladder type: A = <ab||cd> rows (a<b), B = tau rows (i<j), nk = P2(V).
// Ring type:   A = Tring rows (i,a), B = W_mbej^T rows (j,b), nk = O V.
int cpu_dot_rows(const double* A, const double* B, double* out, long nr, long nq, long nk, int threads) {
  if (!A || !B || !out || nr < 0 || nq < 0 || nk < 0) return -1;
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
  for (long r = 0; r < nr; ++r) {
    const double* a = A + r * nk;
    for (long q = 0; q < nq; ++q) {
      const double* b = B + q * nk;
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      long k = 0;
      for (; k + 4 <= nk; k += 4) {
        s0 += a[k] * b[k];
        s1 += a[k + 1] * b[k + 1];
        s2 += a[k + 2] * b[k + 2];
        s3 += a[k + 3] * b[k + 3];
      }
      for (; k < nk; ++k) s0 += a[k] * b[k];
      out[r * nq + q] = (s0 + s1) + (s2 + s3);
    }
  }
  return 0;
}

int cpu_max_threads() { return omp_get_max_threads(); }

}  // extern "C"

// ---------------------------------------------------------------------------
// Packed spin-orbital CCSD right-hand side, FP64, OpenMP: the same terms as
// oracle/ccsd_ref.py (rhs / intermediates, each citing the reference line it
// follows), evaluated on the packed input blocks of the C ABI instead of
// dense V^4 tensors, so it reaches the O=20 / V=180 workload on a host.
// Every term is written out as loops; contractions that carry O(N^6) work are
// one dot product per output over contiguous rows, as in cpu_dot_rows.
// Independent of the HIP library: it shares no code or layout helper with it.
// ---------------------------------------------------------------------------
#include <cmath>
#include <vector>

namespace {

inline long pidx(long p, long q) { return q * (q - 1) / 2 + p; }  // p < q

// packed pair index of (p, q) with the antisymmetry sign; s = 0 when p == q
inline long ppair(long p, long q, double& s) {
  if (p < q) { s = 1.0; return pidx(p, q); }
  if (p > q) { s = -1.0; return pidx(q, p); }
  s = 0.0;
  return 0;
}

inline double dot(const double* a, const double* b, long n) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  long k = 0;
  for (; k + 4 <= n; k += 4) {
    s0 += a[k] * b[k];
    s1 += a[k + 1] * b[k + 1];
    s2 += a[k + 2] * b[k + 2];
    s3 += a[k + 3] * b[k + 3];
  }
  for (; k < n; ++k) s0 += a[k] * b[k];
  return (s0 + s1) + (s2 + s3);
}

struct Blocks {
  long O, V, N, po, pv, ov;
  const double *f, *oooo, *ooov, *oovv, *ovov, *ovvv, *vvvv;
  // <mn||ie>, <mn||ef>, <ia||jb>, <ma||ef> with the antisymmetry signs
  double g_ooov(long m, long n, long i, long e) const {
    double s;
    long mn = ppair(m, n, s);
    return s == 0.0 ? 0.0 : s * ooov[mn * ov + i * V + e];
  }
  double g_oovv(long m, long n, long e, long f2) const {
    double s1, s2;
    long mn = ppair(m, n, s1), ef = ppair(e, f2, s2);
    return s1 * s2 == 0.0 ? 0.0 : s1 * s2 * oovv[mn * pv + ef];
  }
  double g_ovov(long i, long a, long j, long b) const { return ovov[(i * V + a) * ov + j * V + b]; }
  double g_ovvv(long m, long a, long e, long f2) const {
    double s;
    long ef = ppair(e, f2, s);
    return s == 0.0 ? 0.0 : s * ovvv[(m * V + a) * pv + ef];
  }
};

}  // namespace

extern "C" {

// R1 (O x V) and packed R2 (P2(O) x P2(V)) of D t = R at amplitudes t1 (O x V)
// and packed t2, for the packed blocks of ccsd_amd.h and the full N x N Fock
// matrix (occupied first).  Returns 0, or -1 on a bad argument.
int cpu_ccsd_residual(int O_, int V_, const double* fock, const double* oooo, const double* ooov,
                      const double* oovv, const double* ovov, const double* ovvv, const double* vvvv,
                      const double* t1, const double* t2p, double* r1, double* r2p, int threads) {
  if (O_ < 2 || V_ < 2 || !fock || !oooo || !ooov || !oovv || !ovov || !ovvv || !vvvv || !t1 || !t2p || !r1 ||
      !r2p)
    return -1;
  if (threads > 0) omp_set_num_threads(threads);
  Blocks B;
  B.O = O_; B.V = V_; B.N = O_ + V_;
  B.po = B.O * (B.O - 1) / 2; B.pv = B.V * (B.V - 1) / 2; B.ov = B.O * B.V;
  B.f = fock; B.oooo = oooo; B.ooov = ooov; B.oovv = oovv; B.ovov = ovov; B.ovvv = ovvv; B.vvvv = vvvv;
  const long O = B.O, V = B.V, N = B.N, po = B.po, pv = B.pv, ov = B.ov;
  auto fo = [&](long p, long q) { return fock[p * N + q]; };
  auto T1 = [&](long i, long a) { return t1[i * V + a]; };

  // dense t2 [i][j][a][b] and tau (Fill_tau :1669-1850), packed tau [ij][ab]
  std::vector<double> t2((size_t)O * O * V * V, 0.0), tau((size_t)O * O * V * V, 0.0), taup((size_t)po * pv);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < O; ++i)
    for (long j = 0; j < O; ++j) {
      double sij;
      long ij = ppair(i, j, sij);
      for (long a = 0; a < V; ++a)
        for (long b = 0; b < V; ++b) {
          double sab;
          long ab = ppair(a, b, sab);
          const double t = sij * sab == 0.0 ? 0.0 : sij * sab * t2p[ij * pv + ab];
          const size_t x = ((size_t)(i * O + j) * V + a) * V + b;
          t2[x] = t;
          tau[x] = t + T1(i, a) * T1(j, b) - T1(i, b) * T1(j, a);
        }
    }
  auto t2d = [&](long i, long j, long a, long b) { return t2[((size_t)(i * O + j) * V + a) * V + b]; };
  auto taud = [&](long i, long j, long a, long b) { return tau[((size_t)(i * O + j) * V + a) * V + b]; };
#pragma omp parallel for schedule(static)
  for (long j = 1; j < O; ++j)
    for (long i = 0; i < j; ++i)
      for (long b = 1; b < V; ++b)
        for (long a = 0; a < b; ++a) taup[pidx(i, j) * pv + pidx(a, b)] = taud(i, j, a, b);

  // D[(m,e)][(n,f)] = <mn||ef>
  std::vector<double> D((size_t)ov * ov);
#pragma omp parallel for schedule(static)
  for (long m = 0; m < O; ++m)
    for (long e = 0; e < V; ++e)
      for (long n = 0; n < O; ++n)
        for (long f2 = 0; f2 < V; ++f2) D[(size_t)(m * V + e) * ov + n * V + f2] = B.g_oovv(m, n, e, f2);

  // F_me = f_me + t_n^f <mn||ef>  (:1351, Fill_F2 :1596-1602)
  std::vector<double> Fme((size_t)ov);
#pragma omp parallel for schedule(static)
  for (long me = 0; me < ov; ++me) Fme[me] = fo(me / V, O + me % V) + dot(&D[(size_t)me * ov], t1, ov);

  // F_mi = (1-d) f_mi + t_n^e <mn||ie> + 1/2 t_in^ef <mn||ef> + F_me t_i^e  (:708.., :1309, :1329, Fill_F2)
  std::vector<double> Fmi((size_t)O * O);
#pragma omp parallel for schedule(static)
  for (long x = 0; x < O * O; ++x) {
    const long m = x / O, i = x % O;
    double s = m == i ? 0.0 : fo(m, i);
    for (long n = 0; n < O; ++n)
      for (long e = 0; e < V; ++e) s += T1(n, e) * B.g_ooov(m, n, i, e);
    for (long n = 0; n < O; ++n)
      for (long e = 0; e < V; ++e)
        for (long f2 = 0; f2 < V; ++f2) s += 0.5 * t2d(i, n, e, f2) * D[(size_t)(m * V + e) * ov + n * V + f2];
    for (long e = 0; e < V; ++e) s += Fme[m * V + e] * T1(i, e);
    Fmi[x] = s;
  }

  // F_ae = (1-d) f_ae + t_m^f <am||ef> - 1/2 t_mn^af <mn||ef> - F_me t_m^a  (:1076, :1251, :1356, :1376, Fill_F3)
  std::vector<double> Fae((size_t)V * V);
#pragma omp parallel for schedule(dynamic)
  for (long x = 0; x < V * V; ++x) {
    const long a = x / V, e = x % V;
    double s = a == e ? 0.0 : fo(O + a, O + e);
    for (long m = 0; m < O; ++m)
      for (long f2 = 0; f2 < V; ++f2) s -= T1(m, f2) * B.g_ovvv(m, a, e, f2);  // <am||ef> = -<ma||ef>
    for (long m = 0; m < O; ++m)
      for (long n = 0; n < O; ++n)
        for (long f2 = 0; f2 < V; ++f2) s -= 0.5 * t2d(m, n, a, f2) * D[(size_t)(m * V + e) * ov + n * V + f2];
    for (long m = 0; m < O; ++m) s -= Fme[m * V + e] * T1(m, a);
    Fae[x] = s;
  }

  // W_mnij packed [mn][ij] = <mn||ij> + P(ij) t_j^e <mn||ie> + 1/2 tau_ij^ef <mn||ef>  (W_1, :659-1029)
  std::vector<double> W1((size_t)po * po);
#pragma omp parallel for schedule(dynamic)
  for (long mn = 0; mn < po; ++mn) {
    long n = 1;
    while (pidx(0, n + 1) <= mn) ++n;
    const long m = mn - pidx(0, n);
    for (long j = 1; j < O; ++j)
      for (long i = 0; i < j; ++i) {
        const long ij = pidx(i, j);
        double s = oooo[mn * po + ij];
        for (long e = 0; e < V; ++e) s += T1(j, e) * B.g_ooov(m, n, i, e) - T1(i, e) * B.g_ooov(m, n, j, e);
        s += dot(&taup[ij * pv], &oovv[mn * pv], pv);  // 1/2 sum_ef = sum_{e<f}
        W1[mn * po + ij] = s;
      }
  }

  // W_mnie [m][n][i][e] = <mn||ie> + t_i^f <mn||fe>  (W_3)
  std::vector<double> W3((size_t)O * O * O * V);
#pragma omp parallel for schedule(static)
  for (long x = 0; x < O * O * O * V; ++x) {
    const long e = x % V, i = (x / V) % O, n = (x / (V * O)) % O, m = x / (V * O * O);
    double s = B.g_ooov(m, n, i, e);
    for (long f2 = 0; f2 < V; ++f2) s += T1(i, f2) * B.g_oovv(m, n, f2, e);
    W3[x] = s;
  }

  // W_mbij [(m,b)][ij] (i<j) = <mb||ij> + P(ij) t_i^e <mb||ej> + 1/2 <mb||ef> tau_ij^ef  (W_2, :1047-1295)
  std::vector<double> W2((size_t)ov * po);
#pragma omp parallel for schedule(dynamic)
  for (long mb = 0; mb < ov; ++mb) {
    const long m = mb / V, b = mb % V;
    for (long j = 1; j < O; ++j)
      for (long i = 0; i < j; ++i) {
        const long ij = pidx(i, j);
        double s = ooov[ij * ov + mb];  // <mb||ij> = <ij||mb>
        for (long e = 0; e < V; ++e)   // <mb||ej> = -<mb||je>
          s += -T1(i, e) * B.g_ovov(m, b, j, e) + T1(j, e) * B.g_ovov(m, b, i, e);
        s += dot(&ovvv[mb * pv], &taup[ij * pv], pv);
        W2[mb * po + ij] = s;
      }
  }

  // W_mbej stored as W4t[(j,b)][(m,e)] (Fill_W4_MPI :4781-5089):
  //   <mb||ej> + t_j^f <mb||ef> - t_n^b W_mnej - 1/2 t_jn^fb <mn||ef>,  W_mnej = -W_mnje
  std::vector<double> Bt((size_t)ov * ov);  // Bt[(j,b)][(n,f)] = t_jn^fb
#pragma omp parallel for schedule(static)
  for (long jb = 0; jb < ov; ++jb)
    for (long n = 0; n < O; ++n)
      for (long f2 = 0; f2 < V; ++f2) Bt[(size_t)jb * ov + n * V + f2] = t2d(jb / V, n, f2, jb % V);
  std::vector<double> W4t((size_t)ov * ov);
#pragma omp parallel for schedule(dynamic)
  for (long jb = 0; jb < ov; ++jb) {
    const long j = jb / V, b = jb % V;
    for (long m = 0; m < O; ++m)
      for (long e = 0; e < V; ++e) {
        double s = -B.g_ovov(m, b, j, e);
        for (long f2 = 0; f2 < V; ++f2) s += T1(j, f2) * B.g_ovvv(m, b, e, f2);
        for (long n = 0; n < O; ++n) s += T1(n, b) * W3[((m * O + n) * O + j) * V + e];
        s -= 0.5 * dot(&D[(size_t)(m * V + e) * ov], &Bt[(size_t)jb * ov], ov);
        W4t[(size_t)jb * ov + m * V + e] = s;
      }
  }

  // ring[(i,a)][(j,b)] = t_im^ae W_mbej  (Fill_W5 :6201-6882)
  std::vector<double> Tr((size_t)ov * ov);  // Tr[(i,a)][(m,e)] = t_im^ae
#pragma omp parallel for schedule(static)
  for (long ia = 0; ia < ov; ++ia)
    for (long m = 0; m < O; ++m)
      for (long e = 0; e < V; ++e) Tr[(size_t)ia * ov + m * V + e] = t2d(ia / V, m, ia % V, e);
  std::vector<double> ring((size_t)ov * ov);
#pragma omp parallel for schedule(dynamic)
  for (long ia = 0; ia < ov; ++ia)
    for (long jb = 0; jb < ov; ++jb) ring[(size_t)ia * ov + jb] = dot(&Tr[(size_t)ia * ov], &W4t[(size_t)jb * ov], ov);
  auto RG = [&](long i, long a, long j, long b) { return ring[(size_t)(i * V + a) * ov + j * V + b]; };

  // R2 (Fill_t2_new :2001-2406)
#pragma omp parallel for schedule(dynamic)
  for (long ab = 0; ab < pv; ++ab) {
    long b = 1;
    while (pidx(0, b + 1) <= ab) ++b;
    const long a = ab - pidx(0, b);
    for (long j = 1; j < O; ++j)
      for (long i = 0; i < j; ++i) {
        const long ij = pidx(i, j);
        double s = oovv[ij * pv + ab];                                           // <ab||ij>
        s += RG(i, a, j, b) - RG(j, a, i, b) - RG(i, b, j, a) + RG(j, b, i, a);  // P(ij)P(ab) ring
        for (long m = 0; m < O; ++m)                                             // -P(ab) t_m^a W_mbij
          s -= T1(m, a) * W2[(m * V + b) * po + ij] - T1(m, b) * W2[(m * V + a) * po + ij];
        for (long mn = 0; mn < po; ++mn) s += taup[mn * pv + ab] * W1[mn * po + ij];  // 1/2 tau W_mnij
        for (long e = 0; e < V; ++e)                                             // P(ab) t_ij^ae F_be
          s += t2d(i, j, a, e) * Fae[b * V + e] - t2d(i, j, b, e) * Fae[a * V + e];
        for (long m = 0; m < O; ++m)                                             // -P(ij) t_im^ab F_mj
          s -= t2d(i, m, a, b) * Fmi[m * O + j] - t2d(j, m, a, b) * Fmi[m * O + i];
        for (long e = 0; e < V; ++e)  // P(ij) t_i^e <ab||ej>, <ab||ej> = -<je||ab>
          s += -T1(i, e) * ovvv[(j * V + e) * pv + ab] + T1(j, e) * ovvv[(i * V + e) * pv + ab];
        s += dot(&taup[ij * pv], &vvvv[ab * pv], pv);                            // 1/2 tau_ij^ef <ab||ef>
        r2p[ij * pv + ab] = s;
      }
  }

  // R1 (Fill_t2_new :2410-2457, part 1 :1400, :1409)
#pragma omp parallel for schedule(dynamic)
  for (long ia = 0; ia < ov; ++ia) {
    const long i = ia / V, a = ia % V;
    double s = fo(i, O + a);
    for (long e = 0; e < V; ++e)
      if (e != a) s += T1(i, e) * fo(O + a, O + e);
    for (long m = 0; m < O; ++m) {  // -1/2 tau_im^ef <ma||ef>
      double sim;
      const long im = ppair(i, m, sim);
      if (sim != 0.0) s -= sim * dot(&taup[im * pv], &ovvv[(m * V + a) * pv], pv);
    }
    for (long m = 0; m < O; ++m)  // -1/2 t_mn^ae W_mnie
      for (long n = 0; n < O; ++n)
        for (long e = 0; e < V; ++e) s -= 0.5 * t2d(m, n, a, e) * W3[((m * O + n) * O + i) * V + e];
    for (long m = 0; m < O; ++m)  // t_im^ae F_me
      for (long e = 0; e < V; ++e) s += t2d(i, m, a, e) * Fme[m * V + e];
    for (long m = 0; m < O; ++m) s -= T1(m, a) * Fmi[m * O + i];  // -t_m^a F_mi
    for (long n = 0; n < O; ++n)                                    // -t_n^f <na||if>
      for (long f2 = 0; f2 < V; ++f2) s -= T1(n, f2) * B.g_ovov(n, a, i, f2);
    r1[ia] = s;
  }
  return 0;
}

}  // extern "C"
