// CPU restatement of the reference's contraction loops (test and baseline
// infrastructure only: bench.py times it as the CPU baseline beside the GPU;
// nothing in the product links it).
//
// The reference evaluates every CCSD term as Armadillo element-wise products
// reduced with accu() over compact slices, distributed over MPI ranks by
// output index (ccsd_memory_optimized.cpp:2046-2048).  Two loop shapes carry
// almost all of its flops:
//  * ladder type: one output per (a<b, i<j) is a dot product over the packed
//    virtual pairs c<d, e.g. 1/2 sum_cd <ab||cd> tau_ij^cd (Fill_t2_new,
//    :2093, :2158, :2208, :2277, :2342, :2393); W_mnij, W_mbij and the hh
//    ladder have the same form;
//  * ring type: one output per ((i,a),(j,b)) is a dot product over (m,e),
//    W_5(a,b)(i,j) = sum_em t2(b,j)(e,m) W_4(a,i)(e,m) (Fill_W5, :6240), as is
//    the t2 part of W_mbej (Fill_W4_MPI, :4807-4809).
// Here OpenMP threads take the place of the MPI ranks and split the output
// rows the same way; each output is one dot product over contiguous rows, as
// accu(A % B) is.
#include <omp.h>

#include <cstddef>

extern "C" {

// out[r][q] = sum_k A[r][k] * B[q][k] for r < nr, q < nq (row-major, rows of length nk).
// Ladder type: A = <ab||cd> rows (a<b), B = tau rows (i<j), nk = P2(V).
// Ring type:   A = Tring rows (i,a), B = W_mbej^T rows (j,b), nk = O V.
int cpu_dot_rows(const double* A, const double* B, double* out, long nr, long nq, long nk, int threads) {
  if (!A || !B || !out || nr < 0 || nq < 0 || nk < 0) return -1;
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
  for (long r = 0; r < nr; ++r) {
    const double* a = A + r * nk;
    for (long q = 0; q < nq; ++q) {
      const double* b = B + q * nk;
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      long k = 0;
      for (; k + 4 <= nk; k += 4) {
        s0 += a[k] * b[k];
        s1 += a[k + 1] * b[k + 1];
        s2 += a[k + 2] * b[k + 2];
        s3 += a[k + 3] * b[k + 3];
      }
      for (; k < nk; ++k) s0 += a[k] * b[k];
      out[r * nq + q] = (s0 + s1) + (s2 + s3);
    }
  }
  return 0;
}

int cpu_max_threads() { return omp_get_max_threads(); }

}  // extern "C"
