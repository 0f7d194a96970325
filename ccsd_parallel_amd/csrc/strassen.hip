// One level of Strassen's method (Winograd's form: 7 half-size products, 15
// block additions) over the MFMA GEMM engine, for the square products of the
// spin-orbital iteration (the ring Tring . W_mbej^T and the W_mbej quadratic
// term, both (OV)^3), in FP64 and on FP32 operands (with an FP32 or an FP64
// result).
//
// With A, B, C split into h x h blocks (h = n / 2):
//   S1 = A21 + A22, S2 = S1 - A11, S3 = A11 - A21, S4 = A12 - S2
//   T1 = B12 - B11, T2 = B22 - T1, T3 = B22 - B12, T4 = T2 - B21
//   M1 = A11 B11, M2 = A12 B21, M3 = S4 B22, M4 = A22 T4,
//   M5 = S1 T1, M6 = S2 T2, M7 = S3 T3
//   C11 = M1 + M2, C12 = M1 + M6 + M5 + M3, C21 = M1 + M6 + M7 - M4,
//   C22 = M1 + M6 + M7 + M5
// (V. Strassen, Numer. Math. 13, 354 (1969); S. Winograd's variant with 15
// additions).  The schedule below accumulates in the C blocks (12 block
// passes, three of them fused pairs; 8 of them form the operand sums, which a
// caller may form once with strassen_sums and pass in) and needs three
// h x h temporaries (S and T of the operand type, one product of the result
// type).  Each product is one engine call at
// the half size: at n = 12000 the seven 6000^3 products and the block passes
// take less time than one 12000^3 product.  The error bound grows by a small
// constant factor over the classical product (one level only).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "gemm.hpp"
#include "strassen.hpp"

namespace ccsd {
namespace {

// D = a X + b Y over a rows x cols block; each view is (pointer, row stride)
// with unit stride along its columns.  Y may be null (D = a X).  D may alias X
// or Y.
template <typename T>
__global__ void block_axpby_kernel(T* D, long ldd, const T* X, long ldx, T a, const T* Y, long ldy, T b, long rows,
                                   long cols) {
  for (long r = blockIdx.y; r < rows; r += gridDim.y) {
    T* d = D + r * ldd;
    const T* x = X + r * ldx;
    const T* y = Y ? Y + r * ldy : nullptr;
    for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < cols; c += (long)gridDim.x * blockDim.x)
      d[c] = y ? a * x[c] + b * y[c] : a * x[c];
  }
}

// D = X + Y and E = D in one pass (E may be null); or, with Y null, D += X and
// E += X (the two accumulations of one block).
template <typename T>
__global__ void block_sum2_kernel(T* D, long ldd, T* E, long lde, const T* X, long ldx, const T* Y, long ldy,
                                  long h) {
  for (long r = blockIdx.y; r < h; r += gridDim.y) {
    T* d = D + r * ldd;
    T* e = E + r * lde;
    const T* x = X + r * ldx;
    const T* y = Y ? Y + r * ldy : nullptr;
    for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < h; c += (long)gridDim.x * blockDim.x) {
      if (y) {
        const T v = x[c] + y[c];
        d[c] = v;
        e[c] = v;
      } else {
        const T v = x[c];
        d[c] += v;
        e[c] += v;
      }
    }
  }
}

template <typename T>
struct View {
  const T* p;
  long ld;
};

dim3 block_grid(long rows, long cols) {
  return dim3((unsigned)std::min<long>((cols + 255) / 256, 8), (unsigned)std::min<long>(rows, 65535));
}

// (D, E) = (X + Y, X + Y), or with Y null (D + X, E + X)
template <typename T>
hipError_t sum2(View<T> D, View<T> E, View<T> X, View<T> Y, long h, hipStream_t s) {
  hipLaunchKernelGGL(block_sum2_kernel<T>, block_grid(h, h), dim3(256), 0, s, const_cast<T*>(D.p), D.ld,
                     const_cast<T*>(E.p), E.ld, X.p, X.ld, Y.p, Y.ld, h);
  return hipGetLastError();
}

template <typename T>
hipError_t axpby(View<T> D, View<T> X, T a, View<T> Y, T b, long h, hipStream_t s) {
  hipLaunchKernelGGL(block_axpby_kernel<T>, block_grid(h, h), dim3(256), 0, s, const_cast<T*>(D.p), D.ld, X.p, X.ld,
                     a, Y.p, Y.ld, b, h, h);
  return hipGetLastError();
}

template <typename T>
bool applies(const GemmDesc<T>& d, int min_n) {
  return d.M == d.N && d.N == d.K && d.M % 2 == 0 && d.M >= min_n && d.nb1 == 1 && d.nb2 == 1 &&
         !d.reduce_batch && d.beta == T(0) && d.sCn == 1 && (d.sAk == 1 || d.sAm == 1) && (d.sBk == 1 || d.sBn == 1);
}

// C = alpha A B on operands of type E; the result (type R) is d.C (R = E) or
// d.Cd (R = double, E = float).  tmp: S and T (h^2 of E each), then the
// product temporary (h^2 of R).  preS / preT: the four operand sums of A / B
// formed beforehand (strassen_sums), or null to form them here.
template <typename E, typename R>
hipError_t strassen_run(const GemmDesc<E>& d, R* Cr, void* tmp, const GemmWorkspace& ws, hipStream_t s,
                        const E* preS = nullptr, const E* preT = nullptr) {
  const long h = d.M / 2;
  // Operand blocks are handled in the orientation that keeps their unit stride
  // along the columns of a view: A as [m][k] (k-contiguous) or [k][m]; B as
  // [n][k] (k-contiguous) or [k][n].  The temporaries S (A-shaped) and T
  // (B-shaped) take the same orientation with row stride h.
  const bool ak = d.sAk == 1, bk = d.sBk == 1;
  const long lda = ak ? d.sAm : d.sAk, ldb = bk ? d.sBn : d.sBk;
  auto ablk = [&](int bi, int bj) {  // A block (bi, bj): rows m in bi, k in bj
    return View<E>{d.A + bi * h * d.sAm + bj * h * d.sAk, lda};
  };
  auto bblk = [&](int bi, int bj) {  // B block (bi, bj): k in bi, n in bj
    return View<E>{d.B + bi * h * d.sBk + bj * h * d.sBn, ldb};
  };
  auto cblk = [&](int bi, int bj) { return View<R>{Cr + bi * h * d.sCm + bj * h, d.sCm}; };
  E* Sp = static_cast<E*>(tmp);
  E* Tp = Sp + h * h;
  R* Mp = reinterpret_cast<R*>(Sp + 2 * h * h);
  const View<E> S{Sp, h}, T{Tp, h};
  const View<R> Tm{Mp, h};
  // operand sum q (0-3: S1..S4 / T1..T4): the precomputed block, or the temporary
  auto sv = [&](int q) { return preS ? View<E>{preS + q * h * h, h} : S; };
  auto tv = [&](int q) { return preT ? View<E>{preT + q * h * h, h} : T; };
  // C(view) = alpha a b + beta C, a an A-shaped view, b a B-shaped view
  auto mul = [&](View<R> c, View<E> a, View<E> b, E alpha, E beta) {
    GemmDesc<E> g;
    g.A = a.p; g.B = b.p;
    if constexpr (std::is_same<E, R>::value)
      g.C = const_cast<R*>(c.p);
    else
      g.Cd = const_cast<R*>(c.p);
    g.M = (int)h; g.N = (int)h; g.K = (int)h;
    if (ak) { g.sAm = a.ld; g.sAk = 1; } else { g.sAm = 1; g.sAk = a.ld; }
    if (bk) { g.sBk = 1; g.sBn = b.ld; } else { g.sBk = b.ld; g.sBn = 1; }
    g.sCm = c.ld; g.sCn = 1;
    g.alpha = alpha; g.beta = beta;
    return gemm<E>(g, ws, s);
  };
  const E al = d.alpha, one = 1, zero = 0;
  const View<E> A11 = ablk(0, 0), A12 = ablk(0, 1), A21 = ablk(1, 0), A22 = ablk(1, 1);
  const View<E> B11 = bblk(0, 0), B12 = bblk(0, 1), B21 = bblk(1, 0), B22 = bblk(1, 1);
  const View<R> C11 = cblk(0, 0), C12 = cblk(0, 1), C21 = cblk(1, 0), C22 = cblk(1, 1);
  hipError_t e;
#define ST(call) \
  if ((e = (call)) != hipSuccess) return e
  if (!preS) ST(axpby(S, A21, one, A22, one, h, s));   // S1
  if (!preT) ST(axpby(T, B12, one, B11, -one, h, s));  // T1
  ST(mul(Tm, sv(0), tv(0), al, zero));                 // M5
  if (!preS) ST(axpby(S, S, one, A11, -one, h, s));    // S2
  if (!preT) ST(axpby(T, B22, one, T, -one, h, s));    // T2
  ST(mul(C12, sv(1), tv(1), al, zero));                // M6
  ST(mul(C11, A11, B11, al, zero));                    // M1
  ST(sum2(C12, C21, C12, C11, h, s));                  // C12 = C21 = U2 = M1 + M6
  ST(mul(C11, A12, B21, al, one));                     // C11 = M1 + M2
  if (!preS) ST(axpby(S, A12, one, S, -one, h, s));    // S4
  if (!preT) ST(axpby(T, T, one, B21, -one, h, s));    // T4
  ST(sum2(C12, C22, C12, Tm, h, s));                   // C12 = C22 = U4 = U2 + M5
  ST(mul(C12, sv(3), B22, al, one));                   // C12 = U4 + M3
  ST(mul(C21, A22, tv(3), -al, one));                  // C21 = U2 - M4
  if (!preS) ST(axpby(S, A11, one, A21, -one, h, s));  // S3
  if (!preT) ST(axpby(T, B22, one, B12, -one, h, s));  // T3
  ST(mul(Tm, sv(2), tv(2), al, zero));                 // M7
  ST(sum2(C21, C22, Tm, View<R>{nullptr, 0}, h, s));  // C21 = U2 - M4 + M7, C22 = U4 + M7
#undef ST
  return hipSuccess;
}

}  // namespace

bool strassen_applies(const GemmDesc<double>& d, int min_n) { return applies(d, min_n) && d.Cd == nullptr; }

bool strassen_applies(const GemmDesc<float>& d, int min_n) {
  return applies(d, min_n) && ((d.C == nullptr) != (d.Cd == nullptr));
}

size_t strassen_temp_elems(long n) { return 3ull * (size_t)(n / 2) * (size_t)(n / 2); }

hipError_t gemm_strassen(const GemmDesc<double>& d, double* tmp, const GemmWorkspace& ws, hipStream_t s,
                         const StrassenSums* sums) {
  if (!strassen_applies(d, 2) || !tmp) return hipErrorInvalidValue;
  return strassen_run<double, double>(d, d.C, tmp, ws, s, sums ? sums->S : nullptr, sums ? sums->T : nullptr);
}

hipError_t strassen_sums(const GemmDesc<double>& d, double* S, double* T, hipStream_t s) {
  if (!strassen_applies(d, 2)) return hipErrorInvalidValue;
  const long h = d.M / 2, hh = h * h;
  const bool ak = d.sAk == 1, bk = d.sBk == 1;
  const long lda = ak ? d.sAm : d.sAk, ldb = bk ? d.sBn : d.sBk;
  auto ablk = [&](int bi, int bj) { return View<double>{d.A + bi * h * d.sAm + bj * h * d.sAk, lda}; };
  auto bblk = [&](int bi, int bj) { return View<double>{d.B + bi * h * d.sBk + bj * h * d.sBn, ldb}; };
  auto blk = [&](double* p, int q) { return View<double>{p + q * hh, h}; };
  hipError_t e;
#define ST(call) \
  if ((e = (call)) != hipSuccess) return e
  if (S) {
    ST(axpby(blk(S, 0), ablk(1, 0), 1.0, ablk(1, 1), 1.0, h, s));    // S1 = A21 + A22
    ST(axpby(blk(S, 1), blk(S, 0), 1.0, ablk(0, 0), -1.0, h, s));   // S2 = S1 - A11
    ST(axpby(blk(S, 2), ablk(0, 0), 1.0, ablk(1, 0), -1.0, h, s));  // S3 = A11 - A21
    ST(axpby(blk(S, 3), ablk(0, 1), 1.0, blk(S, 1), -1.0, h, s));   // S4 = A12 - S2
  }
  if (T) {
    ST(axpby(blk(T, 0), bblk(0, 1), 1.0, bblk(0, 0), -1.0, h, s));  // T1 = B12 - B11
    ST(axpby(blk(T, 1), bblk(1, 1), 1.0, blk(T, 0), -1.0, h, s));   // T2 = B22 - T1
    ST(axpby(blk(T, 2), bblk(1, 1), 1.0, bblk(0, 1), -1.0, h, s));  // T3 = B22 - B12
    ST(axpby(blk(T, 3), blk(T, 1), 1.0, bblk(1, 0), -1.0, h, s));   // T4 = T2 - B21
  }
#undef ST
  return hipSuccess;
}

hipError_t gemm_strassen(const GemmDesc<float>& d, double* tmp, const GemmWorkspace& ws, hipStream_t s) {
  if (!strassen_applies(d, 2) || !tmp) return hipErrorInvalidValue;
  if (d.Cd) return strassen_run<float, double>(d, d.Cd, tmp, ws, s);
  return strassen_run<float, float>(d, d.C, tmp, ws, s);
}

hipError_t strassen_accumulate(float* C, long ldc, const float* P, long ldp, long rows, long cols, hipStream_t s) {
  hipLaunchKernelGGL(block_axpby_kernel<float>, block_grid(rows, cols), dim3(256), 0, s, C, ldc, P, ldp, 1.0f, C,
                     ldc, 1.0f, rows, cols);
  return hipGetLastError();
}

}  // namespace ccsd
