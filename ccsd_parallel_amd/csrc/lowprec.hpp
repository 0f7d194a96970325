// Mixed-precision operands for the four dominant contractions
// (pp ladder, ring, W_mbej quadratic term, W_mbij tau term).
//
// Operands are stored k-contiguous with a leading dimension padded to a
// multiple of 64 (zero filled), so the low-precision GEMMs run without K
// bounds checks and with 16-byte vector loads.  Results are produced in FP32
// and added into the FP64 destination.
#pragma once
#include "engine.hpp"

namespace ccsd {

size_t lowprec_plan_bytes(const Dims& d, int precision);
int lowprec_allocate(Ctx* c, int precision);
void lowprec_free(Ctx* c);
int lowprec_amplitudes(Ctx* c, int precision);
int lowprec_w2(Ctx* c, int precision);
int lowprec_wmbej(Ctx* c, int precision);
int lowprec_ladder(Ctx* c, int precision);
int lowprec_ring(Ctx* c, int precision);

// Self-test entry used by ccsd_selftest_gemm: C(MxN) (FP64) = alpha * A B + beta C
// with A given k-contiguous [M][K] and B given k-contiguous [N][K] (device FP64).
int lowprec_gemm_nt(const double* A, const double* B, double* C, int M, int N, int K, double alpha, double beta,
                    int precision, hipStream_t s);

}  // namespace ccsd
