// AO -> MO four-index transform in HBM, the RHF input path of the engine.
//
// The reference transforms its AO integrals slice by slice on the host with
// Armadillo products and MPI all-to-all exchanges between the quarter steps
// (Prepear_AOs, ccsd_memory_optimized.cpp:2698-3522), then builds the
// antisymmetrised spin-orbital slices on the fly inside every iteration
// (Fill_integ*_2D, :3574-4779).  Here the four quarter transforms are four
// MFMA GEMMs over the whole n^4 block, and the six packed spin-orbital blocks
// are written once, in HBM, straight from the spatial MO integrals.
//
// Each pass contracts the last (contiguous) index and writes the new index
// first:  out[s][a][b][c] = sum_x C[x][s] in[a][b][c][x],
// one GEMM with M = n (s), N = n^3 (abc), K = n (x); after four passes the
// index order is back to [p][q][r][s].  2 n^5 flops per pass.
#include "engine.hpp"

namespace ccsd {

int ao_to_mo_device(double* eri, double* tmp, const double* C, int n, const GemmWorkspace& ws, hipStream_t s) {
  const long n1 = n, n3 = n1 * n1 * n1;
  double* in = eri;
  double* out = tmp;
  for (int pass = 0; pass < 4; ++pass) {
    GemmDesc<double> g;
    g.A = C; g.sAm = 1; g.sAk = n1;    // A(s, x) = C[x][s]
    g.B = in; g.sBk = 1; g.sBn = n1;   // B(x, abc) = in[abc][x]
    g.C = out; g.sCm = n3; g.sCn = 1;  // out[s][abc]
    g.M = n; g.N = (int)n3; g.K = n;
    g.alpha = 1.0; g.beta = 0.0;
    const hipError_t e = gemm<double>(g, ws, s);
    if (e != hipSuccess) return CCSD_ERR_HIP;
    std::swap(in, out);
  }
  // four swaps: the result is back in eri
  return CCSD_OK;
}

}  // namespace ccsd
