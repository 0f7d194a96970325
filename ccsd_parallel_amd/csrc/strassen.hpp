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
// FP64: `sums` may supply the operand sums formed beforehand by
// strassen_sums (for an operand shared by several products, or constant).
struct StrassenSums {
  const double* S = nullptr;  // S1..S4 of A (4 (n/2)^2 doubles), or null
  const double* T = nullptr;  // T1..T4 of B, or null
};
hipError_t gemm_strassen(const GemmDesc<double>& d, double* tmp, const GemmWorkspace& ws, hipStream_t s,
                         const StrassenSums* sums = nullptr);
hipError_t gemm_strassen(const GemmDesc<float>& d, double* tmp, const GemmWorkspace& ws, hipStream_t s);

// The operand sums of d's A (into S, 4 (n/2)^2 doubles: S1 = A21 + A22, S2 =
// S1 - A11, S3 = A11 - A21, S4 = A12 - S2) and of its B (into T: T1 = B12 - B11,
// T2 = B22 - T1, T3 = B22 - B12, T4 = T2 - B21), either may be null.  Each
// block keeps its operand's orientation (row stride n/2), so a product that
// uses them must view A / B with the same strides as d.
hipError_t strassen_sums(const GemmDesc<double>& d, double* S, double* T, hipStream_t s);

// C += P over rows x cols (row strides ldc, ldp): for callers that need
// beta = 1 and run the product into P first.
hipError_t strassen_accumulate(float* C, long ldc, const float* P, long ldp, long rows, long cols, hipStream_t s);

}  // namespace ccsd
