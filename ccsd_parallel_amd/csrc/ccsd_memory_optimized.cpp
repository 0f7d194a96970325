// Drop-in CCSD_Memory_optimized on top of the C ABI (see include/ccsd_memory_optimized.h).
//
// Spin-orbital convention of the reference (ccsd_memory_optimized.cpp:34-54, :483-485):
// occupied spin orbitals first, then virtual; spin orbital P has spatial
// orbital P/2 and spin P%2.  <PQ||RS> = <PQ|RS> - <PQ|SR> with
// <PQ|RS> = (pr|qs) when spin(P) = spin(R) and spin(Q) = spin(S), else 0,
// which is what the reference's Fill_integ*_2D slices compute
// (e.g. :3728, :3870).
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../include/ccsd_memory_optimized.h"

using ccsd_compat::mat;
using ccsd_compat::vec;

namespace {

long p2(long n) { return n * (n - 1) / 2; }

int frozen_core_orbitals(const vec& Z) {
  int n = 0;
  for (int i = 0; i < Z.n_elem(); ++i) {
    int z = (int)std::lround(Z(i));
    if (z > 54) n += 27;
    else if (z > 36) n += 18;
    else if (z > 18) n += 9;
    else if (z > 10) n += 5;
    else if (z > 2) n += 1;
  }
  return n;
}

}  // namespace

CCSD_Memory_optimized::CCSD_Memory_optimized(int n_N, vec zz, mat rr, std::string B_s, int n_Elec, int ran, int siz,
                                             Hartree_Fock_Solver* Hartfock, bool frez)
    : n_Nuclei(n_N), n_Electrons(n_Elec), rank(ran), size(siz), Z(zz), R(rr), Basis_Set(B_s), HartFock(Hartfock),
      freeze_core(frez) {}

CCSD_Memory_optimized::CCSD_Memory_optimized(int n_spatial, int n_occ_spatial, const double* eps,
                                             const double* mo_eri, double e_hf)
    : n_Electrons(2 * n_occ_spatial), n_(n_spatial), nocc_(n_occ_spatial), e_hf_(e_hf) {
  eps_.assign(eps, eps + n_spatial);
  mo_eri_.assign(mo_eri, mo_eri + (size_t)n_spatial * n_spatial * n_spatial * n_spatial);
}

CCSD_Memory_optimized::~CCSD_Memory_optimized() = default;

// AO -> MO: (pq|rs) = sum C(mu,p) C(nu,q) C(la,r) C(si,s) (mu nu|la si), four quarter transforms.
int CCSD_Memory_optimized::mo_from_provider(double toler) {
  e_hf_ = HartFock->get_Energy(toler, 0);
  const int n = HartFock->ReturnMatrixSize();
  if (n <= 0) return CCSD_ERR_ARG;
  mat C = HartFock->ReturnC();
  mat ev = HartFock->return_eigval_F();
  if (C.n_rows != n || C.n_cols != n || (int)ev.v.size() < n) return CCSD_ERR_ARG;
  const size_t n2 = (size_t)n * n, n4 = n2 * n2;
  std::vector<double> ao(n4);
  for (int j = 0; j < n; ++j)
    for (int l = 0; l <= j; ++l) {
      mat M = HartFock->Return_Field_Q(j, l);
      if (M.n_rows != n || M.n_cols != n) return CCSD_ERR_ARG;
      for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) {
          double v = M(p, q);
          ao[((size_t)p * n + q) * n2 + (size_t)j * n + l] = v;
          ao[((size_t)p * n + q) * n2 + (size_t)l * n + j] = v;
          ao[((size_t)j * n + l) * n2 + (size_t)p * n + q] = v;
          ao[((size_t)l * n + j) * n2 + (size_t)p * n + q] = v;
        }
    }
  // transform the last index, then rotate the index order; four times
  std::vector<double> tmp(n4);
  for (int pass = 0; pass < 4; ++pass) {
    // tmp[s][a][b][c] = sum_x ao[a][b][c][x] C(x, s)
    for (size_t abc = 0; abc < n2 * n; ++abc)
      for (int s = 0; s < n; ++s) {
        double acc = 0.0;
        const double* row = &ao[abc * n];
        for (int x = 0; x < n; ++x) acc += row[x] * C(x, s);
        tmp[(size_t)s * n2 * n + abc] = acc;
      }
    ao.swap(tmp);
  }
  mo_eri_.swap(ao);
  n_ = n;
  eps_.resize(n);
  for (int p = 0; p < n; ++p) eps_[p] = ev.v[p];
  nocc_ = n_Electrons / 2;
  return CCSD_OK;
}

int CCSD_Memory_optimized::build_spin_orbital_problem() {
  const int nf = freeze_core ? frozen_core_orbitals(Z) : 0;
  const int n = n_, nocc = nocc_ - nf;
  if (nocc < 1 || n - nocc_ < 1) return CCSD_ERR_ARG;
  const int O = 2 * nocc, V = 2 * (n - nocc_), N = O + V;
  // spin orbital -> (spatial, spin)
  auto spatial = [&](int P) { return P < O ? nf + P / 2 : nocc_ + (P - O) / 2; };
  auto spin = [&](int P) { return P < O ? P % 2 : (P - O) % 2; };
  const size_t nn = (size_t)n;
  auto chem = [&](int p, int q, int r, int s) { return mo_eri_[((p * nn + q) * nn + r) * nn + s]; };
  auto phys = [&](int P, int Q, int Rr, int S) {  // <PQ|RS>
    if (spin(P) != spin(Rr) || spin(Q) != spin(S)) return 0.0;
    return chem(spatial(P), spatial(Rr), spatial(Q), spatial(S));
  };
  auto anti = [&](int P, int Q, int Rr, int S) { return phys(P, Q, Rr, S) - phys(P, Q, S, Rr); };
  const long po = p2(O), pv = p2(V), ov = (long)O * V;
  std::vector<int> oi(po), oj(po), va(pv), vb(pv);
  for (int j = 0, k = 0; j < O; ++j)
    for (int i = 0; i < j; ++i, ++k) { oi[k] = i; oj[k] = j; }
  for (int b = 0, k = 0; b < V; ++b)
    for (int a = 0; a < b; ++a, ++k) { va[k] = O + a; vb[k] = O + b; }
  f_.assign((size_t)N * N, 0.0);
  for (int P = 0; P < N; ++P) f_[(size_t)P * N + P] = eps_[spatial(P)];
  oooo_.resize(po * po); ooov_.resize(po * ov); oovv_.resize(po * pv);
  ovov_.resize(ov * ov); ovvv_.resize(ov * pv); vvvv_.resize(pv * pv);
  for (long r = 0; r < po; ++r) {
    for (long c = 0; c < po; ++c) oooo_[r * po + c] = anti(oi[r], oj[r], oi[c], oj[c]);
    for (long c = 0; c < ov; ++c) ooov_[r * ov + c] = anti(oi[r], oj[r], (int)(c / V), O + (int)(c % V));
    for (long c = 0; c < pv; ++c) oovv_[r * pv + c] = anti(oi[r], oj[r], va[c], vb[c]);
  }
  for (long r = 0; r < ov; ++r) {
    const int i = (int)(r / V), a = O + (int)(r % V);
    for (long c = 0; c < ov; ++c) ovov_[r * ov + c] = anti(i, a, (int)(c / V), O + (int)(c % V));
    for (long c = 0; c < pv; ++c) ovvv_[r * pv + c] = anti(i, a, va[c], vb[c]);
  }
  for (long r = 0; r < pv; ++r)
    for (long c = 0; c < pv; ++c) vvvv_[r * pv + c] = anti(va[r], vb[r], va[c], vb[c]);
  return CCSD_OK;
}

double CCSD_Memory_optimized::CCSD(double toler, bool print_stuff) {
  status_ = CCSD_OK;
  if (HartFock) {
    status_ = mo_from_provider(toler);
    HartFock = nullptr;  // not needed any more (:75)
    if (status_ != CCSD_OK) return std::nan("");
  }
  status_ = build_spin_orbital_problem();
  if (status_ != CCSD_OK) return std::nan("");
  ccsd_options o;
  if (have_opt_) o = opt_;
  else ccsd_default_options(&o);
  o.verbose = print_stuff ? 1 : 0;
  const int O = 2 * (nocc_ - (freeze_core ? frozen_core_orbitals(Z) : 0)), V = 2 * (n_ - nocc_);
  if (print_stuff) {
    printf("Number of Electrons for CCSD: %d\n", O);
    printf("Number of Virtuals for CCSD: %d\n", V);
  }
  ccsd_problem p{O, V, f_.data(), oooo_.data(), ooov_.data(), oovv_.data(), ovov_.data(), ovvv_.data(),
                 vvvv_.data()};
  ccsd_ctx* ctx = nullptr;
  status_ = ccsd_create(&p, &o, &ctx);
  if (status_ != CCSD_OK) {
    fprintf(stderr, "CCSD_Memory_optimized: %s\n", ccsd_strerror(status_));
    return std::nan("");
  }
  const size_t n1 = (size_t)O * V, n2 = (size_t)p2(O) * p2(V);
  restarted_ = false;
  if (!checkpoint_.empty()) {
    if (FILE* fp = fopen(checkpoint_.c_str(), "rb")) {
      int hdr[2] = {0, 0};
      std::vector<double> t1(n1), t2(n2);
      if (fread(hdr, 4, 2, fp) == 2 && hdr[0] == O && hdr[1] == V && fread(t1.data(), 8, n1, fp) == n1 &&
          fread(t2.data(), 8, n2, fp) == n2 && ccsd_set_amplitudes(ctx, t1.data(), t2.data()) == CCSD_OK)
        restarted_ = true;
      fclose(fp);
      if (print_stuff) printf("Checkpoint %s: %s\n", checkpoint_.c_str(), restarted_ ? "restart" : "not used");
    }
  }
  status_ = ccsd_solve(ctx, toler, 1000, &result_);  // Itermax = 1000 (:42)
  if (!checkpoint_.empty() && (status_ == CCSD_OK || status_ == CCSD_ERR_NOT_CONVERGED)) {
    std::vector<double> t1(n1), t2(n2);
    if (ccsd_get_amplitudes(ctx, t1.data(), t2.data()) == CCSD_OK) {
      const std::string tmp = checkpoint_ + ".tmp";
      if (FILE* fp = fopen(tmp.c_str(), "wb")) {
        int hdr[2] = {O, V};
        bool ok = fwrite(hdr, 4, 2, fp) == 2 && fwrite(t1.data(), 8, n1, fp) == n1 &&
                  fwrite(t2.data(), 8, n2, fp) == n2;
        ok = (fclose(fp) == 0) && ok;
        if (ok) std::rename(tmp.c_str(), checkpoint_.c_str());
        else std::remove(tmp.c_str());
      }
    }
  }
  ccsd_destroy(ctx);
  if (status_ != CCSD_OK && status_ != CCSD_ERR_NOT_CONVERGED) return std::nan("");
  e_corr_ = result_.e_corr;
  return e_corr_ + e_hf_;
}
