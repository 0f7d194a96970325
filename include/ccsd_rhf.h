/* Closed-shell Hartree-Fock input producer for the CCSD drop-in.
 *
 * Replaces the reference's Hartree_Fock_Solver (hartree_fock_solver.h,
 * hartree_fock_solver.cpp:578-1228) for the calls CCSD makes, with its own
 * integral engine (McMurchie-Davidson over Cartesian Gaussians, the role of
 * hartree_integrals.cpp:118-796) and the reference's basis data for the
 * elements and sets listed below (fill_alpha.cpp:16403-16424, 16919-17010,
 * 23296-).  Host code, FP64; no device work.
 *
 * Supported: basis "STO-3G" and "6-31G" for H (Z=1) and O (Z=8).
 * Geometry in bohr (the reference's INCAR default; use_angstrom multiplies
 * by 1.889725989, initializer.cpp:56-59).
 */
#ifndef CCSD_RHF_H
#define CCSD_RHF_H

#include "ccsd_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ccsd_rhf_result {
  double e_hf;        /* total RHF energy (electronic + nuclear repulsion) */
  double e_nuc;       /* nuclear repulsion */
  double commutator;  /* max |F P S - S P F| at exit (orthogonal basis) */
  int n_basis;        /* spatial basis functions n */
  int n_occ;          /* doubly occupied orbitals */
  int iterations;
  int converged;
} ccsd_rhf_result;

/* Number of basis functions for the molecule, or < 0 (unknown element/basis). */
int ccsd_basis_size(int n_atoms, const int* Z, const char* basis);

/* Converged closed-shell RHF.  R_bohr[3*a + x]; n_electrons even.  tol bounds
 * |dE| and the commutator.  Optional outputs (may be NULL), n = n_basis:
 * C[mu*n + p] MO coefficients, eps[p] orbital energies (ascending),
 * mo_eri[((p*n+q)*n+r)*n+s] = (pq|rs) in the MO basis, ao_eri likewise in the
 * AO basis.  cap_n must be >= n when any array is given. */
int ccsd_rhf(int n_atoms, const int* Z, const double* R_bohr, const char* basis, int n_electrons, double tol,
             ccsd_rhf_result* out, double* C, double* eps, double* mo_eri, double* ao_eri, int cap_n);

/* AO one-electron matrices (n x n, row-major) for tests: overlap S, kinetic T
 * and nuclear attraction V (H_core = T + V). */
int ccsd_ao_one_electron(int n_atoms, const int* Z, const double* R_bohr, const char* basis, double* S, double* T,
                         double* V, int cap_n);

#ifdef __cplusplus
}

#include <string>
#include <vector>

#include "ccsd_memory_optimized.h"

/* Concrete Hartree_Fock_Solver with the reference's constructor
 * (hartree_fock_solver.cpp:3): nuclei, charges Z, positions R (bohr, one row
 * per atom), basis name, electrons, print flag, MPI rank/size (one process
 * here), frozen-core flag (the CCSD class applies the freeze). */
class RHF_Hartree_Fock : public Hartree_Fock_Solver {
 public:
  RHF_Hartree_Fock(int n_N, ccsd_compat::vec Z, ccsd_compat::mat R, std::string basis, int n_elec, bool print_stuff,
                   int rank, int size, bool frozen_core);
  double get_Energy(double toler, int do_uhf) override;
  int ReturnMatrixSize() override;
  ccsd_compat::mat ReturnC() override;
  ccsd_compat::mat return_eigval_F() override;
  ccsd_compat::mat Return_Field_Q(int j, int l) override;
  const ccsd_rhf_result& result() const { return res_; }
  int status() const { return status_; }

 private:
  int n_atoms_, n_elec_;
  std::vector<int> Z_;
  std::vector<double> R_;
  std::string basis_;
  bool print_;
  ccsd_rhf_result res_{};
  int status_ = 0;
  std::vector<double> C_, eps_, ao_eri_;
};
#endif

#endif
