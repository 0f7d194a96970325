// One level of Strassen's method (Winograd's form: 7 half-size products, 15
// block additions) over the MFMA GEMM engine, for the square FP64 products of
// the spin-orbital iteration (the ring Tring . W_mbej^T and the W_mbej
// quadratic term, both (OV)^3).
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
// passes, three of them fused pairs) and needs three
// h x h temporaries (S, T and one product).  Each product is one engine call at
// the half size: at n = 12000 the seven 6000^3 products and the block passes
// take less time than one 12000^3 product.  The error bound grows by a small
// constant factor over the classical product (one level only).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gemm.hpp"
#include "strassen.hpp"

namespace ccsd {
namespace {

// D = a X + b Y over an h x h block; each view is (pointer, row stride) with
// unit stride along its columns.  Y may be null (D = a X).  D may alias X or Y.
__global__ void block_axpby_kernel(double* D, long ldd, const double* X, long ldx, double a, const double* Y,
                                   long ldy, double b, long h) {
  for (long r = blockIdx.y; r < h; r += gridDim.y) {
    double* d = D + r * ldd;
    const double* x = X + r * ldx;
    const double* y = Y ? Y + r * ldy : nullptr;
    for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < h; c += (long)gridDim.x * blockDim.x)
      d[c] = y ? a * x[c] + b * y[c] : a * x[c];
  }
}

// D = X + Y and E = D in one pass (E may be null); or, with Y null, D += X and
// E += X (the two accumulations of one block).
__global__ void block_sum2_kernel(double* D, long ldd, double* E, long lde, const double* X, long ldx,
                                  const double* Y, long ldy, long h) {
  for (long r = blockIdx.y; r < h; r += gridDim.y) {
    double* d = D + r * ldd;
    double* e = E + r * lde;
    const double* x = X + r * ldx;
    const double* y = Y ? Y + r * ldy : nullptr;
    for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < h; c += (long)gridDim.x * blockDim.x) {
      if (y) {
        const double v = x[c] + y[c];
        d[c] = v;
        e[c] = v;
      } else {
        const double v = x[c];
        d[c] += v;
        e[c] += v;
      }
    }
  }
}

struct View {
  const double* p;
  long ld;
};

// (D, E) = (X + Y, X + Y), or with Y null (D + X, E + X)
hipError_t sum2(View D, View E, View X, View Y, long h, hipStream_t s) {
  const unsigned gx = (unsigned)std::min<long>((h + 255) / 256, 8);
  const unsigned gy = (unsigned)std::min<long>(h, 65535);
  hipLaunchKernelGGL(block_sum2_kernel, dim3(gx, gy), dim3(256), 0, s, const_cast<double*>(D.p), D.ld,
                     const_cast<double*>(E.p), E.ld, X.p, X.ld, Y.p, Y.ld, h);
  return hipGetLastError();
}

hipError_t axpby(View D, View X, double a, View Y, double b, long h, hipStream_t s) {
  const unsigned gx = (unsigned)std::min<long>((h + 255) / 256, 8);
  const unsigned gy = (unsigned)std::min<long>(h, 65535);
  hipLaunchKernelGGL(block_axpby_kernel, dim3(gx, gy), dim3(256), 0, s, const_cast<double*>(D.p), D.ld, X.p, X.ld, a,
                     Y.p, Y.ld, b, h);
  return hipGetLastError();
}

}  // namespace

bool strassen_applies(const GemmDesc<double>& d, int min_n) {
  return d.M == d.N && d.N == d.K && d.M % 2 == 0 && d.M >= min_n && d.nb1 == 1 && d.nb2 == 1 &&
         !d.reduce_batch && d.beta == 0.0 && d.sCn == 1 && d.Cd == nullptr && (d.sAk == 1 || d.sAm == 1) &&
         (d.sBk == 1 || d.sBn == 1);
}

size_t strassen_temp_elems(long n) { return 3ull * (size_t)(n / 2) * (size_t)(n / 2); }

hipError_t gemm_strassen(const GemmDesc<double>& d, double* tmp, const GemmWorkspace& ws, hipStream_t s) {
  if (!strassen_applies(d, 2) || !tmp) return hipErrorInvalidValue;
  const long h = d.M / 2;
  // Operand blocks are handled in the orientation that keeps their unit stride
  // along the columns of a view: A as [m][k] (k-contiguous) or [k][m]; B as
  // [n][k] (k-contiguous) or [k][n].  The temporaries S (A-shaped) and T
  // (B-shaped) take the same orientation with row stride h.
  const bool ak = d.sAk == 1, bk = d.sBk == 1;
  const long lda = ak ? d.sAm : d.sAk, ldb = bk ? d.sBn : d.sBk;
  auto ablk = [&](int bi, int bj) {  // A block (bi, bj): rows m in bi, k in bj
    return View{d.A + bi * h * d.sAm + bj * h * d.sAk, lda};
  };
  auto bblk = [&](int bi, int bj) {  // B block (bi, bj): k in bi, n in bj
    return View{d.B + bi * h * d.sBk + bj * h * d.sBn, ldb};
  };
  auto cblk = [&](int bi, int bj) { return View{d.C + bi * h * d.sCm + bj * h, d.sCm}; };
  double* Sp = tmp;
  double* Tp = tmp + h * h;
  double* Mp = tmp + 2 * h * h;
  const View S{Sp, h}, T{Tp, h}, Tm{Mp, h};
  // C(view) = alpha a b + beta C, a an A-shaped view, b a B-shaped view
  auto mul = [&](View c, View a, View b, double alpha, double beta) {
    GemmDesc<double> g;
    g.A = a.p; g.B = b.p; g.C = const_cast<double*>(c.p);
    g.M = (int)h; g.N = (int)h; g.K = (int)h;
    if (ak) { g.sAm = a.ld; g.sAk = 1; } else { g.sAm = 1; g.sAk = a.ld; }
    if (bk) { g.sBk = 1; g.sBn = b.ld; } else { g.sBk = b.ld; g.sBn = 1; }
    g.sCm = c.ld; g.sCn = 1;
    g.alpha = alpha; g.beta = beta;
    return gemm<double>(g, ws, s);
  };
  const double al = d.alpha;
  const View A11 = ablk(0, 0), A12 = ablk(0, 1), A21 = ablk(1, 0), A22 = ablk(1, 1);
  const View B11 = bblk(0, 0), B12 = bblk(0, 1), B21 = bblk(1, 0), B22 = bblk(1, 1);
  const View C11 = cblk(0, 0), C12 = cblk(0, 1), C21 = cblk(1, 0), C22 = cblk(1, 1);
  hipError_t e;
#define ST(call) \
  if ((e = (call)) != hipSuccess) return e
  ST(axpby(S, A21, 1.0, A22, 1.0, h, s));    // S1
  ST(axpby(T, B12, 1.0, B11, -1.0, h, s));   // T1
  ST(mul(Tm, S, T, al, 0.0));                // M5
  ST(axpby(S, S, 1.0, A11, -1.0, h, s));     // S2
  ST(axpby(T, B22, 1.0, T, -1.0, h, s));     // T2
  ST(mul(C12, S, T, al, 0.0));               // M6
  ST(mul(C11, A11, B11, al, 0.0));           // M1
  ST(sum2(C12, C21, C12, C11, h, s));        // C12 = C21 = U2 = M1 + M6
  ST(mul(C11, A12, B21, al, 1.0));           // C11 = M1 + M2
  ST(axpby(S, A12, 1.0, S, -1.0, h, s));     // S4
  ST(axpby(T, T, 1.0, B21, -1.0, h, s));     // T4
  ST(sum2(C12, C22, C12, Tm, h, s));         // C12 = C22 = U4 = U2 + M5
  ST(mul(C12, S, B22, al, 1.0));             // C12 = U4 + M3
  ST(mul(C21, A22, T, -al, 1.0));            // C21 = U2 - M4
  ST(axpby(S, A11, 1.0, A21, -1.0, h, s));   // S3
  ST(axpby(T, B22, 1.0, B12, -1.0, h, s));   // T3
  ST(mul(Tm, S, T, al, 0.0));                // M7
  ST(sum2(C21, C22, Tm, View{nullptr, 0}, h, s));  // C21 = U2 - M4 + M7, C22 = U4 + M7
#undef ST
  return hipSuccess;
}

}  // namespace ccsd
