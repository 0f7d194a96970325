// Drives the drop-in CCSD_Memory_optimized class through both constructors.
// Input (binary, little endian): int32 n, nocc; double e_hf; double eps[n];
// double C[n*n] (C[mu*n+p], AO mu, MO p); double ao[n^4] ((pq|rs) AO, [((p*n+q)*n+r)*n+s]);
// double mo[n^4] (same integrals in the MO basis).
#include <cstdio>
#include <vector>

#include "ccsd_memory_optimized.h"

using ccsd_compat::mat;
using ccsd_compat::vec;

struct FileHF : public Hartree_Fock_Solver {
  int n = 0;
  double ehf = 0.0;
  std::vector<double> eps, C, ao;
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
  mat Return_Field_Q(int j, int l) override {
    mat m(n, n);
    for (int p = 0; p < n; ++p)
      for (int q = 0; q < n; ++q) m(p, q) = ao[(((size_t)p * n + q) * n + j) * n + l];
    return m;
  }
};

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  FILE* fp = fopen(argv[1], "rb");
  if (!fp) return 2;
  int hdr[2];
  FileHF hf;
  if (fread(hdr, 4, 2, fp) != 2 || fread(&hf.ehf, 8, 1, fp) != 1) return 2;
  hf.n = hdr[0];
  const int nocc = hdr[1];
  const size_t n = hf.n, n4 = n * n * n * n;
  hf.eps.resize(n); hf.C.resize(n * n); hf.ao.resize(n4);
  std::vector<double> mo(n4);
  if (fread(hf.eps.data(), 8, n, fp) != n || fread(hf.C.data(), 8, n * n, fp) != n * n ||
      fread(hf.ao.data(), 8, n4, fp) != n4 || fread(mo.data(), 8, n4, fp) != n4)
    return 2;
  fclose(fp);
  vec Z(1);
  Z(0) = 2.0 * nocc;
  mat R(1, 3);
  CCSD_Memory_optimized a(1, Z, R, "synthetic", 2 * nocc, 0, 1, &hf, false);
  double e1 = a.CCSD(1e-11, false);
  CCSD_Memory_optimized b(hf.n, nocc, hf.eps.data(), mo.data(), hf.ehf);
  double e2 = b.CCSD(1e-11, false);
  printf("provider_total %.15f provider_corr %.15f status %d iterations %d\n", e1, a.correlation_energy(),
         a.last_status(), a.last_result().iterations);
  printf("direct_total %.15f direct_corr %.15f status %d\n", e2, b.correlation_energy(), b.last_status());
  return (a.last_status() == 0 && b.last_status() == 0) ? 0 : 1;
}
