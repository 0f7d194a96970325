// Drives the drop-in CCSD_Memory_optimized class through both constructors.
//
//   dropin_main <input.bin> [rekeyed|prekey] [auto|spin|closed]
//     Input (binary, little endian): int32 n, nocc; double e_hf; double eps[n];
//     double C[n*n] (C[mu*n+p], AO mu, MO p); double ao[n^4] ((pq|rs) AO,
//     [((p*n+q)*n+r)*n+s]); double mo[n^4] (same integrals in the MO basis).
//     The provider holds its AO integrals the way the reference's
//     Hartree_Fock_Solver does: the SCF fill field_Q(o1,o3)(o2,o4) =
//     (o1 o2|o3 o4) (hartree_fock_solver.cpp:798-802), then the re-key
//     temp_Q(o1,o2)(o3,o4) = field_Q(o1,o3)(o2,o4) (:1045-1210) that get_Energy
//     applies before the CCSD object asks for slices.  "prekey" skips the
//     re-key (a provider handing out the SCF's own layout).  The last
//     argument selects the amplitudes the class iterates (set_spin_adapted).
//   dropin_main --frozen Z1 Z2 ...
//     Prints Freeze_How_Meny_Orbitals() for those nuclear charges with
//     freeze_core on and off (no device needed).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ccsd_memory_optimized.h"

using ccsd_compat::mat;
using ccsd_compat::vec;

struct FileHF : public Hartree_Fock_Solver {
  int n = 0;
  double ehf = 0.0;
  std::vector<double> eps, C;
  std::vector<mat> field_Q;  // [o1 * n + o3] -> (o2, o4)

  void fill_like_reference(const std::vector<double>& ao, bool rekey) {
    const size_t N = n;
    std::vector<mat> q(N * N, mat(n, n));
    for (size_t o1 = 0; o1 < N; ++o1)
      for (size_t o2 = 0; o2 < N; ++o2)
        for (size_t o3 = 0; o3 < N; ++o3)
          for (size_t o4 = 0; o4 < N; ++o4)
            q[o1 * N + o3](o2, o4) = ao[((o1 * N + o2) * N + o3) * N + o4];
    if (!rekey) {
      field_Q = q;
      return;
    }
    field_Q.assign(N * N, mat(n, n));
    for (size_t o1 = 0; o1 < N; ++o1)
      for (size_t o3 = 0; o3 < N; ++o3)
        for (size_t o2 = 0; o2 < N; ++o2)
          for (size_t o4 = 0; o4 < N; ++o4) field_Q[o1 * N + o2](o3, o4) = q[o1 * N + o3](o2, o4);
  }
  double get_Energy(double, int) override { return ehf; }
  int ReturnMatrixSize() override { return n; }
  mat ReturnC() override {
    mat m(n, n);
    for (int mu = 0; mu < n; ++mu)
      for (int p = 0; p < n; ++p) m(mu, p) = C[(size_t)mu * n + p];
    return m;
  }
  mat return_eigval_F() override {
    mat m(n, 1);
    for (int p = 0; p < n; ++p) m(p, 0) = eps[p];
    return m;
  }
  mat Return_Field_Q(int j, int l) override { return field_Q[(size_t)j * n + l]; }
};

static int frozen_counts(int argc, char** argv) {
  const int nz = argc - 2;
  vec Z(nz);
  for (int i = 0; i < nz; ++i) Z(i) = atof(argv[i + 2]);
  mat R(nz, 3);
  int nel = 0;
  for (int i = 0; i < nz; ++i) nel += (int)Z(i);
  CCSD_Memory_optimized on(nz, Z, R, "6-31G", nel, 0, 1, nullptr, true);
  CCSD_Memory_optimized off(nz, Z, R, "6-31G", nel, 0, 1, nullptr, false);
  printf("frozen_on %d frozen_off %d\n", on.Freeze_How_Meny_Orbitals(), off.Freeze_How_Meny_Orbitals());
  return 0;
}

int main(int argc, char** argv) {
  if (argc >= 2 && strcmp(argv[1], "--frozen") == 0) return frozen_counts(argc, argv);
  if (argc < 2) return 2;
  const bool rekey = !(argc >= 3 && strcmp(argv[2], "prekey") == 0);
  FILE* fp = fopen(argv[1], "rb");
  if (!fp) return 2;
  int hdr[2];
  FileHF hf;
  if (fread(hdr, 4, 2, fp) != 2 || fread(&hf.ehf, 8, 1, fp) != 1) return 2;
  hf.n = hdr[0];
  const int nocc = hdr[1];
  const size_t n = hf.n, n4 = n * n * n * n;
  hf.eps.resize(n); hf.C.resize(n * n);
  std::vector<double> ao(n4), mo(n4);
  if (fread(hf.eps.data(), 8, n, fp) != n || fread(hf.C.data(), 8, n * n, fp) != n * n ||
      fread(ao.data(), 8, n4, fp) != n4 || fread(mo.data(), 8, n4, fp) != n4)
    return 2;
  fclose(fp);
  hf.fill_like_reference(ao, rekey);
  vec Z(1);
  Z(0) = 2.0 * nocc;
  mat R(1, 3);
  int mode = -1;
  if (argc >= 4 && strcmp(argv[3], "spin") == 0) mode = 0;
  if (argc >= 4 && strcmp(argv[3], "closed") == 0) mode = 1;
  CCSD_Memory_optimized a(1, Z, R, "synthetic", 2 * nocc, 0, 1, &hf, false);
  a.set_spin_adapted(mode);
  double e1 = a.CCSD(1e-11, false);
  CCSD_Memory_optimized b(hf.n, nocc, hf.eps.data(), mo.data(), hf.ehf);
  b.set_spin_adapted(mode);
  double e2 = b.CCSD(1e-11, false);
  printf("provider_total %.15f provider_corr %.15f status %d iterations %d closed_shell %d\n", e1,
         a.correlation_energy(), a.last_status(), a.last_result().iterations, a.closed_shell_used() ? 1 : 0);
  printf("direct_total %.15f direct_corr %.15f status %d\n", e2, b.correlation_energy(), b.last_status());
  return (a.last_status() == 0 && b.last_status() == 0) ? 0 : 1;
}
