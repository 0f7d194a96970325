// Mixed-precision operands and contractions of the low-precision iterations.
//
// The four dominant contractions (pp ladder, ring, W_mbej quadratic term,
// W_mbij tau term) run as FP32 MFMA GEMMs or BF16 hi/lo x3 GEMMs on operands
// stored k-contiguous with a leading dimension padded to a multiple of 64
// (zero filled), so those kernels need no K bounds checks; BF16 operands keep
// hi and lo interleaved per 8-element chunk (one 128-byte line per row per
// 32-element k-tile).  Every other
// contraction of the residual runs on FP32 copies of its operands (same
// layouts as the FP64 arrays) through the generic MFMA engine.  All products
// accumulate in FP32 and land in the FP64 intermediates directly (fused
// epilogues); energy, denominators, DIIS and the convergence test stay FP64.
#pragma once
#include <vector>

#include "engine.hpp"

namespace ccsd {

struct LowPrec {
  int have_fp32 = 0, have_bf16 = 0, have_sec = 0;
  long ld_pv = 0, ld_ov = 0;  // padded leading dims for K = pv and K = ov
  long ld_po = 0, ld_v = 0;   // ... and for K = po and K = V
  // FP32 operands of the dominant contractions (k-contiguous, padded); ovvv32
  // and tau32 also serve the secondary contractions in both modes
  float *vvvv32 = nullptr, *ovvv32 = nullptr, *Doovv32 = nullptr;
  float *tau32 = nullptr, *Tring32 = nullptr, *WT32 = nullptr;
  // BF16 hi/lo operands, interleaved per 8-element chunk (row r at r * 2 * ld)
  uint16_t *vvvv_b = nullptr, *ovvv_b = nullptr, *Doovv_b = nullptr;
  uint16_t *tau_b = nullptr, *Tring_b = nullptr, *WT_b = nullptr;
  // BF16 hi/lo operands of three secondary products of the BF16x3 iterations:
  // <mn||ef> (static), W1^T and tau^T (hh ladder), Tv and F_ae (Y)
  uint16_t *oovv_b = nullptr, *W1T_b = nullptr, *tauT_b = nullptr, *Tv_b = nullptr, *Fae_b = nullptr;
  // ... and of W_mbej's t1 term: the unpacked <mb||ef> (static) and t1
  uint16_t *ovvv_u_b = nullptr, *t1_b = nullptr;
  // FP32 copies for the secondary contractions, same layouts as the FP64 arrays:
  // static integrals ...
  float *oovv_f = nullptr, *oovv_vu_f = nullptr, *oovv_ou_f = nullptr, *G_f = nullptr, *ovvv_u_f = nullptr;
  // ... and per-iteration amplitudes / intermediates
  float *t1_f = nullptr, *T2u_f = nullptr, *Tv_f = nullptr;
  float *W1_f = nullptr, *W3_f = nullptr, *W3u_f = nullptr, *W2_f = nullptr, *Fae_f = nullptr, *Fmi_f = nullptr;
  // BF16 operands of the low-precision Lambda iterations: W_mbej^T and
  // <ia||bc>^T (fixed during a backward) and the per-iteration W_mbej and
  // W_mbij cotangents
  uint16_t *WTt_b = nullptr, *ovvvT_b = nullptr, *WTbar_b = nullptr, *W2bar_b = nullptr;
  int have_adj = 0, have_ovvvT = 0;
  // FP32 cotangents of the low-precision Lambda iterations' secondary products
  // (the forward intermediates they pair with reuse the FP32 copies above,
  // refreshed at the converged amplitudes by lowprec_adj_setup): lam1, lam2
  // (rows padded to ld_pv, as tau32), the unpacked lam2 (Yb, Zb), and the
  // small cotangents converted once complete
  float *lam1_f = nullptr, *lam2_f = nullptr, *Yb_f = nullptr, *Zb_f = nullptr, *Faeb_f = nullptr;
  float *W1b_f = nullptr, *Sb_f = nullptr, *Ub_f = nullptr, *W3b_f = nullptr, *Fmib_f = nullptr;
  std::vector<void*> allocs;
};

size_t lowprec_plan_bytes(const Dims& d, int precision);
int lowprec_allocate(Ctx* c, int precision);
void lowprec_free(Ctx* c);
// Low-precision amplitude operands (tau, Tring, t1) of the current iteration.
int lowprec_amplitudes(Ctx* c, int precision);
// Destination of the ring-layout unpack (Tring as the mode's operand, padded rows).
RingOut lowprec_ring_out(Ctx* c, int precision);
// F_me += <mn||ef> t_n^f and R1 += t_im^ae F_me on the low-precision copies.
int lowprec_fme(Ctx* c, int precision, const double* t1);
int lowprec_r1_ring(Ctx* c, int precision);
int lowprec_w2(Ctx* c, int precision);
// FP32 accumulator of W_mbej (stored [(j,b),(m,e)], leading dimension *ld) in
// the low-precision modes: the FP32 ring operand itself, or in BF16 mode the
// upper half of X (free until the ring GEMM writes X).
float* lowprec_wmbej_acc(Ctx* c, int precision, long* ld);
// W_mbej += 1/2 Tring Doovv on the accumulator; BF16 mode emits the ring operand
int lowprec_wmbej(Ctx* c, int precision);
int lowprec_ladder(Ctx* c, int precision);
int lowprec_ring(Ctx* c, int precision);
// BF16x3 forms of three secondary products (BF16x3 iterations only):
//   W1 += <mn||ef> tau^T;  Rp += W1^T tau (hh ladder, W1 complete);
//   Y = Tv F_ae^T (F_ae complete, Tv_f current)
int lowprec_w1_bf16(Ctx* c);
int lowprec_hh_bf16(Ctx* c);
int lowprec_y_bf16(Ctx* c);
// S[(m,b,e)][j] = sum_f <mb||ef> t_j^f into FP32 S32 [(m,b,e)][j] (BF16x3)
int lowprec_wmbej_s_bf16(Ctx* c, const double* t1, float* S32);

// Low-precision (BF16x3) pieces of the adjoint's reverse pass, used by the
// Lambda iterations before their FP64 sweeps.  lowprec_adj_setup converts the
// operands fixed at the converged amplitudes (call after the FP64 forward
// intermediates are in place); the others add one O(N^6) product each:
//   ring:   Tringb += Xb WT,  WTb += Xb Tring   (Xb = ring unpack of lam2)
//   ladder: taub += lam2 <ab||cd>
//   wmbej:  Tringb += 1/2 WTb Doovv
//   w2:     taub += W2b <mb||ef>
int lowprec_adj_setup(Ctx* c);
// lam -> lam1_f, lam2_f and the FP32 unpacks Yb_f (virtual pairs) and Zb_f
// (occupied pairs) for one Lambda iteration's secondary products.
int lowprec_adj_lambda(Ctx* c, const double* lam);
int lowprec_adj_ring(Ctx* c, double* Tringb, double* WTb, const double* lam2);
int lowprec_adj_ladder(Ctx* c, double* taub, const double* lam2);
int lowprec_adj_wmbej(Ctx* c, double* Tringb, const double* WTb);
int lowprec_adj_w2(Ctx* c, double* taub, const double* W2b);

// FP32 GEMM with the operands A32 / B32 (strides of g unless overridden by the
// caller in a copy of g) writing g's FP64 destination.
GemmDesc<float> f32_desc(const GemmDesc<double>& g, const float* A32, const float* B32);

// Self-test entry used by ccsd_selftest_gemm: C(MxN) (FP64) = alpha * A B + beta C
// with A given k-contiguous [M][K] and B given k-contiguous [N][K] (device FP64).
int lowprec_gemm_nt(const double* A, const double* B, double* C, int M, int N, int K, double alpha, double beta,
                    int precision, hipStream_t s);

}  // namespace ccsd
