// Drop-in CCSD_Memory_optimized on top of the C ABI (see include/ccsd_memory_optimized.h).
//
// The class gathers the RHF data (from the HF object or the direct
// constructor) and hands it to the device: by default (FP64) to
// ccsd_rhf_create, which iterates the closed-shell spin-adapted amplitudes
// (every spin block of the reference's RHF storage follows from them), and for
// the mixed precisions to ccsd_create_rhf, which transforms the integrals to
// the MO basis and packs the antisymmetrised spin-orbital blocks on the GPU
// with the reference's convention (ccsd_memory_optimized.cpp:34-54,
// :483-485): occupied spin orbitals first, spin orbital P has spatial orbital
// P/2 and spin P%2, <PQ||RS> = <PQ|RS> - <PQ|SR> with <PQ|RS> = (pr|qs) when
// the spins match pairwise (the Fill_integ*_2D slices, e.g. :3728, :3870).
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../include/ccsd_memory_optimized.h"

using ccsd_compat::mat;
using ccsd_compat::vec;

namespace {

long p2(long n) { return n * (n - 1) / 2; }

}  // namespace

CCSD_Memory_optimized::CCSD_Memory_optimized(int n_N, vec zz, mat rr, std::string B_s, int n_Elec, int ran, int siz,
                                             Hartree_Fock_Solver* Hartfock, bool frez)
    : n_Nuclei(n_N), n_Electrons(n_Elec), rank(ran), size(siz), Z(zz), R(rr), Basis_Set(B_s), HartFock(Hartfock),
      freeze_core(frez) {}

CCSD_Memory_optimized::CCSD_Memory_optimized(int n_spatial, int n_occ_spatial, const double* eps,
                                             const double* mo_eri, double e_hf)
    : n_Electrons(2 * n_occ_spatial), n_(n_spatial), nocc_(n_occ_spatial), e_hf_(e_hf) {
  eps_.assign(eps, eps + n_spatial);
  eri_.assign(mo_eri, mo_eri + (size_t)n_spatial * n_spatial * n_spatial * n_spatial);
}

CCSD_Memory_optimized::~CCSD_Memory_optimized() = default;

// Freeze_How_Meny_Orbitals (ccsd_memory_optimized.cpp:5203-5231): with
// freeze_core, one spatial orbital per nucleus with 2 < Z < 11 (Li..Ne) and
// none for H, He or heavier atoms.  The reference runs its loop to
// n_Electrons and so reads Z past its last nucleus (unchecked, ARMA_NO_DEBUG);
// this loop stops at the nuclei.
int CCSD_Memory_optimized::Freeze_How_Meny_Orbitals() const {
  int val = 0;
  if (!freeze_core) return 0;
  for (int i = 0; i < Z.n_elem(); ++i) {
    const long z = std::lround(Z(i));
    if (z > 2 && z < 11) val += 1;
  }
  return val;
}

// Inputs from the HF object (:49, :50, :68, :74, :2987): energy, size, MO
// coefficients, orbital energies and the AO integral slices
// Return_Field_Q(j, l)(p, q) = (j l | p q), j >= l -- the reference object's
// layout after the re-key at the end of its SCF (hartree_fock_solver.cpp:1045-1210)
// -- unfolded into the full n^4 array.  The transform itself runs on the GPU.
int CCSD_Memory_optimized::gather_from_provider(double toler) {
  e_hf_ = HartFock->get_Energy(toler, 0);
  const int n = HartFock->ReturnMatrixSize();
  if (n <= 0) return CCSD_ERR_ARG;
  mat C = HartFock->ReturnC();
  mat ev = HartFock->return_eigval_F();
  if (C.n_rows != n || C.n_cols != n || (int)ev.v.size() < n) return CCSD_ERR_ARG;
  const size_t n2 = (size_t)n * n;
  eri_.assign(n2 * n2, 0.0);
  for (int j = 0; j < n; ++j)
    for (int l = 0; l <= j; ++l) {
      mat M = HartFock->Return_Field_Q(j, l);
      if (M.n_rows != n || M.n_cols != n) return CCSD_ERR_ARG;
      for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) {
          const double v = M(p, q);
          eri_[((size_t)p * n + q) * n2 + (size_t)j * n + l] = v;
          eri_[((size_t)p * n + q) * n2 + (size_t)l * n + j] = v;
          eri_[((size_t)j * n + l) * n2 + (size_t)p * n + q] = v;
          eri_[((size_t)l * n + j) * n2 + (size_t)p * n + q] = v;
        }
    }
  C_.resize(n2);
  for (int mu = 0; mu < n; ++mu)
    for (int p = 0; p < n; ++p) C_[(size_t)mu * n + p] = C(mu, p);
  n_ = n;
  eps_.resize(n);
  for (int p = 0; p < n; ++p) eps_[p] = ev.v[p];
  nocc_ = n_Electrons / 2;
  return CCSD_OK;
}

double CCSD_Memory_optimized::CCSD(double toler, bool print_stuff) {
  status_ = CCSD_OK;
  if (HartFock) {
    status_ = gather_from_provider(toler);
    HartFock = nullptr;  // not needed any more (:75)
    if (status_ != CCSD_OK) return std::nan("");
  }
  ccsd_options o;
  if (have_opt_) o = opt_;
  else ccsd_default_options(&o);
  o.verbose = print_stuff ? 1 : 0;
  const int nf = Freeze_How_Meny_Orbitals();
  const int O = 2 * (nocc_ - nf), V = 2 * (n_ - nocc_);
  if (print_stuff) {
    printf("Number of Electrons for CCSD: %d\n", O);
    printf("Number of Virtuals for CCSD: %d\n", V);
  }
  ccsd_rhf_problem p{n_, nocc_, nf, eps_.data(), C_.empty() ? nullptr : C_.data(), eri_.data()};
  cs_used_ = spin_mode_ == 1 || (spin_mode_ < 0 && o.precision == CCSD_PREC_FP64);
  // amplitude shapes: closed-shell t1[o*v], t2[o*o*v*v]; spin-orbital t1[O*V], t2[P2(O)*P2(V)]
  const int o_s = nocc_ - nf, v_s = n_ - nocc_;
  const size_t n1 = cs_used_ ? (size_t)o_s * v_s : (size_t)O * V;
  const size_t n2 = cs_used_ ? (size_t)o_s * o_s * v_s * v_s : (size_t)p2(O) * p2(V);
  const int hdr_now[2] = {cs_used_ ? -o_s : O, cs_used_ ? -v_s : V};
  ccsd_ctx* ctx = nullptr;
  ccsd_rhf_ctx* rctx = nullptr;
  status_ = cs_used_ ? ccsd_rhf_create(&p, &o, &rctx) : ccsd_create_rhf(&p, &o, &ctx);
  if (status_ != CCSD_OK) {
    fprintf(stderr, "CCSD_Memory_optimized: %s\n", ccsd_strerror(status_));
    return std::nan("");
  }
  auto set_amps = [&](const double* t1, const double* t2) {
    return cs_used_ ? ccsd_rhf_set_amplitudes(rctx, t1, t2) : ccsd_set_amplitudes(ctx, t1, t2);
  };
  auto get_amps = [&](double* t1, double* t2) {
    return cs_used_ ? ccsd_rhf_get_amplitudes(rctx, t1, t2) : ccsd_get_amplitudes(ctx, t1, t2);
  };
  restarted_ = false;
  if (!checkpoint_.empty()) {
    if (FILE* fp = fopen(checkpoint_.c_str(), "rb")) {
      int hdr[2] = {0, 0};
      std::vector<double> t1(n1), t2(n2);
      if (fread(hdr, 4, 2, fp) == 2 && hdr[0] == hdr_now[0] && hdr[1] == hdr_now[1] &&
          fread(t1.data(), 8, n1, fp) == n1 && fread(t2.data(), 8, n2, fp) == n2 &&
          set_amps(t1.data(), t2.data()) == CCSD_OK)
        restarted_ = true;
      fclose(fp);
      if (print_stuff) printf("Checkpoint %s: %s\n", checkpoint_.c_str(), restarted_ ? "restart" : "not used");
    }
  }
  // Itermax = 1000 (:42)
  status_ = cs_used_ ? ccsd_rhf_solve(rctx, toler, 1000, &result_) : ccsd_solve(ctx, toler, 1000, &result_);
  if (!checkpoint_.empty() && (status_ == CCSD_OK || status_ == CCSD_ERR_NOT_CONVERGED)) {
    std::vector<double> t1(n1), t2(n2);
    if (get_amps(t1.data(), t2.data()) == CCSD_OK) {
      const std::string tmp = checkpoint_ + ".tmp";
      if (FILE* fp = fopen(tmp.c_str(), "wb")) {
        bool ok = fwrite(hdr_now, 4, 2, fp) == 2 && fwrite(t1.data(), 8, n1, fp) == n1 &&
                  fwrite(t2.data(), 8, n2, fp) == n2;
        ok = (fclose(fp) == 0) && ok;
        if (ok) std::rename(tmp.c_str(), checkpoint_.c_str());
        else std::remove(tmp.c_str());
      }
    }
  }
  if (cs_used_) ccsd_rhf_destroy(rctx);
  else ccsd_destroy(ctx);
  if (status_ != CCSD_OK && status_ != CCSD_ERR_NOT_CONVERGED) return std::nan("");
  e_corr_ = result_.e_corr;
  return e_corr_ + e_hf_;
}
