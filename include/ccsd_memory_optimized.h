// Drop-in replacement for the reference's CCSD_Memory_optimized
// (ccsd_memory_optimized.h:15-149, ccsd_memory_optimized.cpp:3-14 and :16-567).
//
// Same constructor arguments and the same entry point
//     double CCSD(double toler, bool print_stuff)
// returning E_HF + E_corr, with the amplitude iteration running on the GPU
// through the C ABI of ccsd_amd.h.  Differences from the reference:
//  * Armadillo is not used: vec/mat below are minimal column-major containers
//    with the accessors the interface needs.
//  * No MPI: `rank` and `size` are accepted and ignored (one process, one GPU).
//  * The Hartree-Fock object is any implementation of the abstract
//    Hartree_Fock_Solver below (the five calls the reference makes,
//    ccsd_memory_optimized.cpp:49, 50, 68, 74, 2987).
//  * A second constructor takes RHF MO data directly (orbital energies and
//    spatial MO integrals (pq|rs)), for callers without an HF object.
//  * The AO->MO transform and the spin-orbital integral blocks are built on
//    the GPU (ccsd_create_rhf) instead of in Prepear_AOs / Fill_integ*_2D on
//    the host.
#ifndef CCSD_MEMORY_OPTIMIZED_DROPIN_H
#define CCSD_MEMORY_OPTIMIZED_DROPIN_H

#include <string>
#include <vector>

#include "ccsd_amd.h"

namespace ccsd_compat {

// Column vector (Armadillo `vec` subset).
struct vec {
  std::vector<double> v;
  vec() = default;
  explicit vec(int n) : v(n, 0.0) {}
  int n_elem() const { return (int)v.size(); }
  double& operator()(int i) { return v[i]; }
  double operator()(int i) const { return v[i]; }
};

// Column-major matrix (Armadillo `mat` subset).
struct mat {
  int n_rows = 0, n_cols = 0;
  std::vector<double> v;
  mat() = default;
  mat(int r, int c) : n_rows(r), n_cols(c), v((size_t)r * c, 0.0) {}
  double& operator()(int i, int j) { return v[(size_t)j * n_rows + i]; }
  double operator()(int i, int j) const { return v[(size_t)j * n_rows + i]; }
};

}  // namespace ccsd_compat

// Input provider with the calls CCSD_Memory_optimized::CCSD makes on the
// reference's Hartree_Fock_Solver (hartree_fock_solver.h:61-71, 89).
class Hartree_Fock_Solver {
 public:
  virtual ~Hartree_Fock_Solver() = default;
  // Converged RHF total energy (the reference calls get_Energy(toler, 0), :49).
  virtual double get_Energy(double toler, int do_uhf) = 0;
  // Number of spatial basis functions n (:50).
  virtual int ReturnMatrixSize() = 0;
  // MO coefficients C(mu, p): AO mu (row), MO p (column), n x n (:68).
  virtual ccsd_compat::mat ReturnC() = 0;
  // Orbital energies, ascending, n entries (:74; a column vector in a mat).
  virtual ccsd_compat::mat return_eigval_F() = 0;
  // AO two-electron slice for j >= l: M(p, q) = (j l | p q) = (p q | j l),
  // chemists' notation (:2987).  This is the slice the reference's object
  // returns after get_Energy: the SCF fills field_Q(o1,o3)(o2,o4) =
  // (o1 o2|o3 o4) (hartree_fock_solver.cpp:798-802, used so by the Fock build
  // at :528), then re-keys it before returning (:1045-1210:
  // temp_Q(o1,o2)(o3,o4) = field_Q(o1,o3)(o2,o4); field_Q = temp_Q), and
  // Prepear_AOs contracts the slice over (p, q) and fills the result
  // symmetrically in (j, l) (ccsd_memory_optimized.cpp:2987-2991, :3025-3026).
  virtual ccsd_compat::mat Return_Field_Q(int j, int l) = 0;
};

class CCSD_Memory_optimized {
 public:
  // Reference constructor (ccsd_memory_optimized.cpp:3-14): nuclei count, charges,
  // positions, basis name, electrons, MPI rank/size (ignored), HF object (not
  // owned), frozen core flag.
  CCSD_Memory_optimized(int n_N, ccsd_compat::vec Z, ccsd_compat::mat R, std::string B_s, int n_Elec, int ran,
                        int siz, Hartree_Fock_Solver* Hartfock, bool frez);

  // Direct RHF MO input: n spatial orbitals, n_occ doubly occupied, orbital
  // energies eps[n], spatial MO integrals mo_eri[((p*n+q)*n+r)*n+s] = (pq|rs),
  // and the HF energy added to the returned total.
  CCSD_Memory_optimized(int n_spatial, int n_occ_spatial, const double* eps, const double* mo_eri, double e_hf);

  ~CCSD_Memory_optimized();

  // Iterate to |E_new - E_old| < toler; returns E_HF + E_corr (:16-567).
  double CCSD(double toler, bool print_stuff);

  double correlation_energy() const { return e_corr_; }
  const ccsd_result& last_result() const { return result_; }
  int last_status() const { return status_; }
  void set_options(const ccsd_options& o) { opt_ = o; have_opt_ = true; }
  // Which amplitudes CCSD() iterates.  -1 (default): the closed-shell
  // spin-adapted amplitudes (ccsd_rhf_create; the reference's RHF spin
  // blocks, about 10x fewer flops) when the precision is FP64, the
  // spin-orbital amplitudes (ccsd_create_rhf) for the mixed precisions;
  // 1: always closed-shell (FP64 only); 0: always spin-orbital.  Both give the
  // same energy.
  void set_spin_adapted(int mode) { spin_mode_ = mode; }
  bool closed_shell_used() const { return cs_used_; }
  // Amplitude checkpoint: when the file exists and holds amplitudes of the
  // same shape and kind, CCSD() starts from them instead of the MP2 guess;
  // after the iteration the final amplitudes are written back.  Format (little
  // endian), spin-orbital: int32 O, V; double t1[O*V]; double t2[P2(O)*P2(V)];
  // closed-shell: int32 -o, -v; double t1[o*v]; double t2[o*o*v*v] (the C
  // ABI's layouts).
  void set_checkpoint(const std::string& path) { checkpoint_ = path; }
  bool restarted() const { return restarted_; }
  // Spatial orbitals left out of the correlation when freeze_core is set
  // (ccsd_memory_optimized.cpp:5203-5231): one per nucleus with 2 < Z < 11.
  int Freeze_How_Meny_Orbitals() const;
  int nuclei() const { return n_Nuclei; }
  int mpi_rank() const { return rank; }  // kept for the reference's signature; one process here
  int mpi_size() const { return size; }

 private:
  int gather_from_provider(double toler);

  int n_Nuclei = 0, n_Electrons = 0, rank = 0, size = 1;
  ccsd_compat::vec Z;
  ccsd_compat::mat R;
  std::string Basis_Set;
  Hartree_Fock_Solver* HartFock = nullptr;
  bool freeze_core = false;

  int n_ = 0, nocc_ = 0;
  // spatial integrals (pq|rs): AO basis with coefficients C_[mu*n + p] (HF
  // provider), or MO basis with C_ empty (direct constructor)
  std::vector<double> eps_, eri_, C_;
  double e_hf_ = 0.0, e_corr_ = 0.0;
  ccsd_options opt_{};
  bool have_opt_ = false;
  std::string checkpoint_;
  bool restarted_ = false;
  ccsd_result result_{};
  int status_ = 0;
  int spin_mode_ = -1;
  bool cs_used_ = false;
};

#endif
