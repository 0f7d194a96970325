// One level of Strassen's method over the GEMM engine (strassen.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "gemm.hpp"

namespace ccsd {

// Smallest n for which the half-size products still fill the device (4096^3
// products of 128 x 128 tiles are 1024 tiles, two rounds of 512 slots).
constexpr int kStrassenMinN = 8192;

// Square products C = alpha A B (beta 0, no batches), n even and at least
// min_n, A and B with a unit stride.  FP64: no FP64 side output.  FP32
// operands: the result in C (FP32) or in Cd (FP64), one of them.
bool strassen_applies(const GemmDesc<double>& d, int min_n = kStrassenMinN);
bool strassen_applies(const GemmDesc<float>& d, int min_n = kStrassenMinN);

// Doubles of the temporaries gemm_strassen needs for an n x n product (either
// precision).
size_t strassen_temp_elems(long n);

// C = alpha A B through seven (n/2)^3 engine calls, for any even n; tmp holds
// strassen_temp_elems(n) doubles.  hipErrorInvalidValue if the product does
// not qualify (strassen_applies with min_n = 2) or tmp is null.
hipError_t gemm_strassen(const GemmDesc<double>& d, double* tmp, const GemmWorkspace& ws, hipStream_t s);
hipError_t gemm_strassen(const GemmDesc<float>& d, double* tmp, const GemmWorkspace& ws, hipStream_t s);

// C += P over rows x cols (row strides ldc, ldp): for callers that need
// beta = 1 and run the product into P first.
hipError_t strassen_accumulate(float* C, long ldc, const float* P, long ldp, long rows, long cols, hipStream_t s);

}  // namespace ccsd
