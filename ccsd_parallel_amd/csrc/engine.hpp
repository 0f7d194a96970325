// Engine context shared by the forward driver, the adjoint and the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>

#include "../../include/ccsd_amd.h"
#include "gemm.hpp"
#include "kernels.hpp"
#include "strassen.hpp"

namespace ccsd {

// The mixed Lambda schedule switches to FP64 when the BF16x3 iteration has
// set no new smallest max|dLambda| in this many iterations: it then sits at
// its rounding floor, above the switch threshold.  (The forward needs no such
// test: the low-precision map settles on its own fixed point, where |dE|
// falls below any threshold.)
constexpr int kLowpStall = 4;

struct LowPrec;   // FP32 / BF16 operand copies for the mixed-precision GEMMs
struct AdjState;  // adjoint (Lambda) buffers, allocated by ccsd_backward

struct Ctx {
  Dims d{};
  ccsd_options opt{};
  int device = 0;
  hipStream_t stream = nullptr;
  size_t bytes = 0;
  std::vector<void*> allocs;

  // static data (FP64)
  double *f = nullptr, *fov = nullptr;
  double *oooo = nullptr, *ooov = nullptr, *oovv = nullptr, *ovvv = nullptr, *vvvv = nullptr;
  double *G = nullptr, *Doovv = nullptr, *ovvv_u = nullptr, *oovv_ou = nullptr, *oovv_vu = nullptr;
  std::vector<double> h_f;  // host copy of the Fock matrix

  // amplitudes: [t1 (ov) | t2 packed (po*pv)]
  long nvec = 0;
  double *amp = nullptr, *amp_new = nullptr, *err = nullptr;
  double *tau = nullptr, *T2u = nullptr, *tauu = nullptr, *Tv = nullptr, *Tring = nullptr;

  // intermediates
  double *Fme = nullptr, *Fmi = nullptr, *Fae = nullptr, *Fae_t1 = nullptr;
  double *W1 = nullptr, *Q = nullptr, *W3 = nullptr, *W3u = nullptr, *W2 = nullptr, *U = nullptr;
  double *WT = nullptr, *X = nullptr, *Y = nullptr, *Z = nullptr, *Rp = nullptr, *R1 = nullptr;
  // temporaries of the one-level Strassen (OV)^3 products (FP64), or null;
  // with them the operand sums S1..S4 of Tring (formed once per iteration for
  // the W_mbej and ring products) and T1..T4 of Doovv (formed with Doovv)
  double* strassen_tmp = nullptr;
  double *strassen_tring_s = nullptr, *strassen_doovv_t = nullptr;

  // DIIS
  int diis_m = 0, diis_count = 0, diis_next = 0;
  double *dvec = nullptr, *derr = nullptr, *dcoef = nullptr;
  std::vector<double> B;  // diis_m x diis_m Gram matrix (host, FP64)

  // reductions / scratch
  double* partial = nullptr;
  double* scal = nullptr;  // small device scalars
  GemmWorkspace ws;

  LowPrec* lp = nullptr;
  AdjState* adj = nullptr;

  // state
  bool solved = false;
  double e_mp2 = 0.0, e_last = 0.0;
  double last_tol = 1e-10;  // tol of the last ccsd_solve (sets the default Lambda tolerance)
  std::vector<double> energies;
  double timings_ms[7] = {0, 0, 0, 0, 0, 0, 0};  // [5], [6]: RHF set-up (transform, packing)
  hipEvent_t ev[12] = {};
  std::string last_error;
};

int setup_from_host(Ctx* c, const ccsd_problem* p);
// RHF input: spatial integrals (pq|rs) [n^4] (AO basis with coefficients C
// [n][n], C[mu*n + p], or MO basis with C null) -> packed spin-orbital blocks
// in HBM.  Temporary device buffers of 1-2 n^4 doubles live only during the call.
int setup_from_rhf(Ctx* c, int n, int nocc, int nfrozen, const double* C, const double* eri);
// Four quarter transforms in place: eri [n^4] (AO) -> (pq|rs) (MO); tmp [n^4] scratch.
int ao_to_mo_device(double* eri, double* tmp, const double* C, int n, const GemmWorkspace& ws, hipStream_t s);
int setup_synthetic(Ctx* c, uint64_t seed, double scale);
int plan_and_allocate(Ctx* c, int precision);
int compute_rhs(Ctx* c, const double* t1, const double* t2, int precision, bool time_it);
// C = alpha A B + beta C for the square (OV)^3 FP64 products (beta 0 or 1):
// one level of Strassen when the context holds its temporaries and the
// product qualifies (with beta 1 through `scratch`, M x N), else the engine.
// sums: operand sums formed beforehand for g's A / B views (strassen.hpp).
int square_product(Ctx* c, const GemmDesc<double>& g, double* scratch, const GemmWorkspace* ws = nullptr,
                   const StrassenSums* sums = nullptr);
// T1..T4 of Doovv as the W_mbej product's B operand, after Doovv is (re)built.
int strassen_doovv_sums(Ctx* c);
int device_energy(Ctx* c, const double* t1, const double* t2, double* e);
// strassen_min_order: as ccsd_options (0 = the default order)
size_t plan_bytes(int O, int V, int precision, int diis_m, int strassen_min_order = 0);
// Bordered Pulay system (B_ij = <e_i|e_j>, -1 border, rhs -1) with diagonal
// scaling; false when singular or the coefficients are not finite / huge.
bool diis_solve_bordered(const std::vector<double>& B, int m, std::vector<double>& coef);

// Row-major strided GEMM descriptor C(MxN) = alpha A(MxK) B(KxN) + beta C.
inline GemmDesc<double> G64(const double* A, long sAm, long sAk, const double* B, long sBk, long sBn, double* C,
                            long sCm, long sCn, long M, long N, long K, double alpha, double beta) {
  GemmDesc<double> g;
  g.A = A; g.B = B; g.C = C;
  g.M = (int)M; g.N = (int)N; g.K = (int)K;
  g.sAm = sAm; g.sAk = sAk; g.sBk = sBk; g.sBn = sBn; g.sCm = sCm; g.sCn = sCn;
  g.alpha = alpha; g.beta = beta;
  return g;
}

// GEMM helper with error propagation
template <typename T>
inline int run_gemm(Ctx* c, const GemmDesc<T>& g, const GemmWorkspace* ws = nullptr, int force_split = 0) {
  hipError_t e = gemm<T>(g, ws ? *ws : c->ws, c->stream, nullptr, -1, force_split);
  if (e != hipSuccess) {
    c->last_error = hipGetErrorString(e);
    return CCSD_ERR_HIP;
  }
  return CCSD_OK;
}

}  // namespace ccsd
