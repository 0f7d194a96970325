// The strided, batched GEMM descriptor (no HIP dependencies, so host-only
// code such as the contraction planner can build it).
//
//   C[b](m,n) = alpha * sum_k A[b](m,k) * B[b](k,n) + beta * C[b](m,n)
//
// Element addresses: A + b1*sA1 + b2*sA2 + m*sAm + k*sAk (likewise B, C), with
// the batch index z = b1*nb2 + b2.  With ``reduce_batch`` all batches sum into
// one C (the batch strides of C are ignored).
#pragma once

namespace ccsd {

template <typename T>
struct GemmDesc {
  const T* A = nullptr;
  const T* B = nullptr;
  T* C = nullptr;
  int M = 0, N = 0, K = 0;
  long sAm = 0, sAk = 0, sBk = 0, sBn = 0, sCm = 0, sCn = 0;
  int nb1 = 1, nb2 = 1;
  long sA1 = 0, sA2 = 0, sB1 = 0, sB2 = 0, sC1 = 0, sC2 = 0;
  T alpha = 1, beta = 0;
  bool reduce_batch = false;
  // When set, the result goes to this FP64 array (same offsets as C, which is
  // then unused): C_d = alpha * (A B accumulated in T) + beta * C_d.  Lets a
  // lower-precision product update an FP64 intermediate without a second pass.
  double* Cd = nullptr;
};

}  // namespace ccsd
