// Adjoint of the converged CCSD energy on the GPU.
//
// Lagrangian  L = E(T) + Lambda . (G(T) - D o T),  G = right-hand side of compute_rhs.
//  * Lambda equations dL/dT = 0:  D o Lambda = dE/dT + (dG/dT)^T Lambda,
//    iterated like the amplitudes (Jacobi + DIIS).  (dG/dT)^T Lambda is the
//    reverse pass of compute_rhs written out by hand: every forward GEMM
//    C += a A B contributes A_bar += a C_bar B^T and/or B_bar += a A^T C_bar,
//    every layout gather contributes the signed scatter.
//  * Gradient dE/dtheta = dL/dtheta at (T*, Lambda*), theta = (f, packed
//    <pq||rs> blocks): the same reverse pass with the integral and Fock
//    cotangents switched on -- the unrelaxed one- and two-particle density
//    contractions -- plus dE/dtheta and the denominator term.
//  The four-virtual block gradient (P2(V) x P2(V), 16 GB at O=40 V=300) is
//  formed whole in HBM only where it fits (opt.recompute_v4): otherwise it is
//  recomputed in row slabs from tau and Lambda and streamed to the host.
//  * With opt.precision = BF16X3 the Lambda iterations run their five O(N^6)
//    reverse products (ring x2, W_mbej, pp ladder, W_mbij) as BF16x3 GEMMs
//    until max|dLambda| < lambda_switch_factor * lambda_tol (or until it stalls:
//    no new smallest max|dLambda| in kLowpStall iterations), then refine: FP64
//    passes, each followed (until one meets max|dLambda| < lambda_tol) by a
//    BF16x3 solve of the linear correction equation.  The gradient pass is
//    always FP64.
// The reference has no adjoint; its forward algebra is restated in engine.hip.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "adjoint.hpp"
#include "lowprec.hpp"

namespace ccsd {

struct AdjState {
  double *e = nullptr;                      // dE/dT  [t1 | t2 packed]
  double *lam = nullptr, *lnew = nullptr, *lerr = nullptr, *tb = nullptr;  // nvec each
  // iterative refinement of a mixed run: the correction and its constant part
  // (nvec each, allocated on first use)
  double *dcorr = nullptr, *rcorr = nullptr;
  double *taub = nullptr, *T2ub = nullptr, *tauub = nullptr, *Tvb = nullptr, *Tringb = nullptr;
  double *Fmeb = nullptr, *Fmib = nullptr, *Faeb = nullptr, *W1b = nullptr, *Qb = nullptr;
  double *W3b = nullptr, *W3ub = nullptr, *W2b = nullptr, *Ub = nullptr, *WTb = nullptr;
  double *Xb = nullptr, *Yb = nullptr, *Zb = nullptr;
  double *partial = nullptr;
  GemmWorkspace ws;
  std::vector<void*> allocs;
  size_t bytes = 0;
};

namespace {

constexpr int kThreads = 256;
inline int grid_for(long n) {
  long b = (n + kThreads - 1) / kThreads;
  return (int)std::max<long>(1, std::min<long>(b, 1L << 20));
}
#define GRID_LOOP(idx, n) \
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < (n); idx += (long)gridDim.x * blockDim.x)
#define LAUNCH(kern, n, ...) hipLaunchKernelGGL(kern, dim3(grid_for(n)), dim3(kThreads), 0, s, __VA_ARGS__)

#define TRY(x)                      \
  do {                              \
    int _rc = (x);                  \
    if (_rc != CCSD_OK) return _rc; \
  } while (0)
#define HIPTRY(x)                            \
  do {                                       \
    hipError_t _e = (x);                     \
    if (_e != hipSuccess) {                  \
      c->last_error = hipGetErrorString(_e); \
      return CCSD_ERR_HIP;                   \
    }                                        \
  } while (0)

__device__ __forceinline__ long pidx(long p, long q) { return q * (q - 1) / 2 + p; }
__device__ __forceinline__ void pair_decode(long k, int& p, int& q) {
  long qq = (long)((1.0 + sqrt(1.0 + 8.0 * (double)k)) * 0.5);
  while (qq * (qq - 1) / 2 > k) --qq;
  while ((qq + 1) * qq / 2 <= k) ++qq;
  q = (int)qq;
  p = (int)(k - qq * (qq - 1) / 2);
}
__device__ __forceinline__ double sgn_pair(int p, int q, long& k) {
  if (p == q) { k = 0; return 0.0; }
  if (p < q) { k = pidx(p, q); return 1.0; }
  k = pidx(q, p);
  return -1.0;
}

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[kThreads / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < kThreads / 64; ++k) s += sh[k];
  __syncthreads();
  return s;
}

__global__ void zero_kernel(double* x, long n) { GRID_LOOP(i, n) x[i] = 0.0; }
__global__ void scale_add_kernel(double* y, const double* x, double a, long n) { GRID_LOOP(i, n) y[i] += a * x[i]; }

// out[(i,j)][c] = sgn(i,j) in[pidx][c]   (leading pair axis over n items unpacked)
template <typename T>
__global__ void unpack_pair_rows_kernel(T* out, const double* in, int n, long cols) {
  const long tot = (long)n * n * cols;
  GRID_LOOP(idx, tot) {
    long ij = idx / cols, cc = idx % cols;
    int i = (int)(ij / n), j = (int)(ij % n);
    long k;
    double sg = sgn_pair(i, j, k);
    out[idx] = (T)((sg == 0.0) ? 0.0 : sg * in[k * cols + cc]);
  }
}
// out[r][(i,j)] = sgn(i,j) in[r][pidx]
__global__ void unpack_pair_cols_kernel(double* out, const double* in, long rows, int n) {
  const long nn = (long)n * n, pn = (long)n * (n - 1) / 2;
  const long tot = rows * nn;
  GRID_LOOP(idx, tot) {
    long r = idx / nn, ij = idx % nn;
    int i = (int)(ij / n), j = (int)(ij % n);
    long k;
    double sg = sgn_pair(i, j, k);
    out[idx] = (sg == 0.0) ? 0.0 : sg * in[r * pn + k];
  }
}
// out[b][pidx(m<n)][c] += in[b][m][n][c] - in[b][n][m][c]
__global__ void pack_pair_rows_adj_kernel(double* out, const double* in, int n, long cols, long nbatch) {
  const long pn = (long)n * (n - 1) / 2;
  const long tot = nbatch * pn * cols;
  GRID_LOOP(idx, tot) {
    long b = idx / (pn * cols), rem = idx % (pn * cols);
    long k = rem / cols, cc = rem % cols;
    int m, q;
    pair_decode(k, m, q);
    const double* src = in + b * (long)n * n * cols;
    out[idx] += src[((long)m * n + q) * cols + cc] - src[((long)q * n + m) * cols + cc];
  }
}
// out[r][pidx(a<b)] += in[r][a][b] - in[r][b][a]
__global__ void pack_pair_cols_adj_kernel(double* out, const double* in, long rows, int n) {
  const long pn = (long)n * (n - 1) / 2, nn = (long)n * n;
  const long tot = rows * pn;
  GRID_LOOP(idx, tot) {
    long r = idx / pn, k = idx % pn;
    int a, b;
    pair_decode(k, a, b);
    out[idx] += in[r * nn + (long)a * n + b] - in[r * nn + (long)b * n + a];
  }
}

// t1_bar[i,a] += sum_{j,b} taubar_full[i,j,a,b] t1[j,b]   (reverse of tau = t2 + t1 t1 - t1 t1)
__global__ void tau_vjp_t1_kernel(double* t1b, const double* taub, const double* t1, Dims d) {
  const int i = blockIdx.x / d.V, a = blockIdx.x % d.V;
  double acc = 0.0;
  for (long jb = threadIdx.x; jb < d.ov; jb += blockDim.x) {
    const int j = (int)(jb / d.V), b = (int)(jb % d.V);
    long k1, k2;
    const double s = sgn_pair(i, j, k1) * sgn_pair(a, b, k2);
    if (s != 0.0) acc += s * taub[k1 * d.pv + k2] * t1[jb];
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) t1b[blockIdx.x] += acc;
}

// C[r][c] += alpha x[r] y[c]
__global__ void outer_add_kernel(double* C, const double* x, const double* y, long nr, long nc, double alpha) {
  const long tot = nr * nc;
  GRID_LOOP(idx, tot) C[idx] += alpha * x[idx / nc] * y[idx % nc];
}

// S_bar[(m,b,e)][j] = WT_bar[(j,b),(m,e)];  S_bar[(m,a,e)][m] -= Fae_bar[a][e]
template <typename T>
__global__ void sbar_kernel(T* Sb, const double* WTb, const double* Faeb, Dims d) {
  const long n = (long)d.O * d.V * d.V * d.O;
  GRID_LOOP(idx, n) {
    const int j = (int)(idx % d.O);
    long t = idx / d.O;
    const int e = (int)(t % d.V);
    t /= d.V;
    const int b = (int)(t % d.V);
    const int m = (int)(t / d.V);
    double v = WTb[((long)j * d.V + b) * d.ov + (long)m * d.V + e];
    if (j == m) v -= Faeb[(long)b * d.V + e];
    Sb[idx] = (T)v;
  }
}

// t1_bar[i,e] += sum_{a != e} lam1[i,a] f[O+a][O+e]
__global__ void r1init_vjp_kernel(double* t1b, const double* lam1, const double* f, Dims d) {
  const long N = d.O + d.V;
  GRID_LOOP(idx, d.ov) {
    const int i = (int)(idx / d.V), e = (int)(idx % d.V);
    double s = 0.0;
    for (int a = 0; a < d.V; ++a)
      if (a != e) s += lam1[(long)i * d.V + a] * f[(long)(d.O + a) * N + d.O + e];
    t1b[idx] += s;
  }
}

// t1_bar[n,e] += sum_{m,i} Fmi_bar[m,i] <mn||ie>
__global__ void fmi_t1_vjp_kernel(double* t1b, const double* Fmib, const double* ooov, Dims d) {
  const int nn = blockIdx.x / d.V, e = blockIdx.x % d.V;
  double acc = 0.0;
  for (long mi = threadIdx.x; mi < (long)d.O * d.O; mi += blockDim.x) {
    const int m = (int)(mi / d.O), i = (int)(mi % d.O);
    long k;
    const double sg = sgn_pair(m, nn, k);
    if (sg != 0.0) acc += sg * Fmib[mi] * ooov[k * d.ov + (long)i * d.V + e];
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) t1b[blockIdx.x] += acc;
}

// lnew = (e + jt) / D ; err = lnew - lam    (packed T2 and T1)
// Orbital-energy denominator of element idx of [t1 | packed t2].
__device__ __forceinline__ double lambda_denominator(long idx, const double* f, const Dims& d) {
  const long N = d.O + d.V;
  if (idx < d.ov) {
    const int i = (int)(idx / d.V), a = (int)(idx % d.V);
    return f[(long)i * N + i] - f[(long)(d.O + a) * N + d.O + a];
  }
  const long q = idx - d.ov;
  int i, j, a, b;
  pair_decode(q / d.pv, i, j);
  pair_decode(q % d.pv, a, b);
  return f[(long)i * N + i] + f[(long)j * N + j] - f[(long)(d.O + a) * N + d.O + a] - f[(long)(d.O + b) * N + d.O + b];
}

__global__ void lambda_update_kernel(double* lnew, double* err, const double* e, const double* jt, const double* lam,
                                     const double* f, Dims d) {
  const long n = d.ov + d.po * d.pv;
  GRID_LOOP(idx, n) {
    const double v = (e[idx] + jt[idx]) / lambda_denominator(idx, f, d);
    lnew[idx] = v;
    err[idx] = v - lam[idx];
  }
}

// Correction equation of the refinement, dcorr = r0 + T(dcorr) / D:
// dnew = r0 + jt / D ; err = dnew - dcur   (jt = T(dcur) from the reverse pass)
__global__ void correction_update_kernel(double* dnew, double* err, const double* r0, const double* jt,
                                         const double* dcur, const double* f, Dims d) {
  const long n = d.ov + d.po * d.pv;
  GRID_LOOP(idx, n) {
    const double v = r0[idx] + jt[idx] / lambda_denominator(idx, f, d);
    dnew[idx] = v;
    err[idx] = v - dcur[idx];
  }
}

// ---- theta-only kernels
// fdiag[p] = -sum over amplitudes containing p (occupied) / + (virtual) of lam * t   (denominator term)
__global__ void fdiag_kernel(double* out, const double* lam, const double* amp, Dims d) {
  const int p = blockIdx.x;  // 0..O+V-1
  double acc = 0.0;
  const bool occ = p < d.O;
  const int q = occ ? p : p - d.O;
  // T1 part
  if (occ) {
    for (int a = threadIdx.x; a < d.V; a += blockDim.x) acc -= lam[(long)q * d.V + a] * amp[(long)q * d.V + a];
  } else {
    for (int i = threadIdx.x; i < d.O; i += blockDim.x) acc += lam[(long)i * d.V + q] * amp[(long)i * d.V + q];
  }
  // T2 part: pairs containing q on the relevant side
  const int n = occ ? d.O : d.V;
  const long other = occ ? d.pv : d.po;
  for (long t = threadIdx.x; t < (long)(n - 1) * other; t += blockDim.x) {
    const int r = (int)(t / other);
    const long w = t % other;
    const int x = r < q ? r : r + 1;
    const long pr = q < x ? pidx(q, x) : pidx(x, q);
    const long off = occ ? pr * d.pv + w : w * d.pv + pr;
    const double v = lam[d.ov + off] * amp[d.ov + off];
    acc += occ ? -v : v;
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) out[p] = acc;
}

// ooov_bar[mn][(i,e)] += Fmi_bar[m,i] t1[n,e] - Fmi_bar[n,i] t1[m,e]
__global__ void fmi_ooov_kernel(double* oob, const double* Fmib, const double* t1, Dims d) {
  const long n = d.po * d.ov;
  GRID_LOOP(idx, n) {
    const long mn = idx / d.ov, ie = idx % d.ov;
    int m, nn;
    pair_decode(mn, m, nn);
    const int i = (int)(ie / d.V), e = (int)(ie % d.V);
    oob[idx] += Fmib[(long)m * d.O + i] * t1[(long)nn * d.V + e] - Fmib[(long)nn * d.O + i] * t1[(long)m * d.V + e];
  }
}
// ooov_bar[mn][i*V+e] += W3_bar[i][mn][e]
__global__ void w3_ooov_kernel(double* oob, const double* W3b, Dims d) {
  const long n = d.po * d.ov;
  GRID_LOOP(idx, n) {
    const long mn = idx / d.ov, ie = idx % d.ov;
    const long i = ie / d.V, e = ie % d.V;
    oob[idx] += W3b[(i * d.po + mn) * d.V + e];
  }
}
// ovov_bar[(m,b),(j,e)] = -G_bar[(m,e),(j,b)]
__global__ void ovov_from_gbar_kernel(double* ovb, const double* Gb, Dims d) {
  const long n = d.ov * d.ov;
  GRID_LOOP(idx, n) {
    const long r = idx / d.ov, cc = idx % d.ov;
    const int m = (int)(r / d.V), b = (int)(r % d.V), j = (int)(cc / d.V), e = (int)(cc % d.V);
    ovb[idx] = -Gb[((long)m * d.V + e) * d.ov + (long)j * d.V + b];
  }
}
// x = 0.5 (x + x^T) for a square n x n matrix (upper triangle pass)
__global__ void symmetrize_kernel(double* x, long n) {
  const long tot = n * (n - 1) / 2;
  GRID_LOOP(k, tot) {
    int i, j;
    pair_decode(k, i, j);
    const double v = 0.5 * (x[(long)i * n + j] + x[(long)j * n + i]);
    x[(long)i * n + j] = v;
    x[(long)j * n + i] = v;
  }
}

}  // namespace

// ============================================================================
static int adj_alloc(Ctx* c, AdjState* A, double** p, long n) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, (size_t)std::max<long>(n, 1) * 8);
  if (e != hipSuccess) {
    c->last_error = hipGetErrorString(e);
    return CCSD_ERR_HIP;
  }
  A->allocs.push_back(q);
  A->bytes += (size_t)std::max<long>(n, 1) * 8;
  *p = (double*)q;
  return CCSD_OK;
}

void adjoint_free(Ctx* c) {
  if (!c->adj) return;
  for (void* p : c->adj->allocs) (void)hipFree(p);
  delete c->adj;
  c->adj = nullptr;
}

static int adjoint_prepare(Ctx* c) {
  if (c->adj) return CCSD_OK;
  c->adj = new AdjState();
  AdjState* A = c->adj;
  const Dims& d = c->d;
  const long O = d.O, V = d.V, VV = V * V;
  TRY(adj_alloc(c, A, &A->e, c->nvec));
  TRY(adj_alloc(c, A, &A->lam, c->nvec));
  TRY(adj_alloc(c, A, &A->lnew, c->nvec));
  TRY(adj_alloc(c, A, &A->lerr, c->nvec));
  TRY(adj_alloc(c, A, &A->tb, c->nvec));
  TRY(adj_alloc(c, A, &A->taub, d.po * d.pv));
  TRY(adj_alloc(c, A, &A->T2ub, O * O * d.pv));
  TRY(adj_alloc(c, A, &A->tauub, O * O * d.pv));
  TRY(adj_alloc(c, A, &A->Tvb, d.po * VV));
  TRY(adj_alloc(c, A, &A->Tringb, d.ov * d.ov));
  TRY(adj_alloc(c, A, &A->Fmeb, d.ov));
  TRY(adj_alloc(c, A, &A->Fmib, O * O));
  TRY(adj_alloc(c, A, &A->Faeb, VV));
  TRY(adj_alloc(c, A, &A->W1b, d.po * d.po));
  TRY(adj_alloc(c, A, &A->Qb, d.po * O * O));
  TRY(adj_alloc(c, A, &A->W3b, O * d.po * V));
  TRY(adj_alloc(c, A, &A->W3ub, O * O * O * V));
  TRY(adj_alloc(c, A, &A->W2b, d.po * d.ov));
  TRY(adj_alloc(c, A, &A->Ub, O * O * d.ov));
  TRY(adj_alloc(c, A, &A->WTb, d.ov * d.ov));
  TRY(adj_alloc(c, A, &A->Xb, std::max(d.ov * d.ov, O * VV * O)));
  TRY(adj_alloc(c, A, &A->Yb, d.po * VV));
  TRY(adj_alloc(c, A, &A->Zb, O * O * d.pv));
  TRY(adj_alloc(c, A, &A->partial, 1 << 16));
  // batch reductions over O*O (or P2(O)) slices of O x V outputs need more than the forward workspace
  const long wsd = std::max<long>(8L << 20, 2L * std::max(O * O, d.po) * d.ov);
  double* w = nullptr;
  TRY(adj_alloc(c, A, &w, wsd));
  A->ws = GemmWorkspace{w, (size_t)wsd * 8};
  return CCSD_OK;
}

// Integral / Fock cotangents (only for the final gradient pass).
struct ThetaBars {
  double *f = nullptr;      // N x N
  double *oooo = nullptr, *ooov = nullptr, *oovv = nullptr, *ovvv = nullptr;
  double *G = nullptr, *Doovv = nullptr, *oovv_ou = nullptr, *oovv_vu = nullptr, *ovvv_u = nullptr;
  double *fvv = nullptr, *fov = nullptr;  // V x V off-diagonal part, O x V
};

// Reverse pass of compute_rhs at the forward intermediates held in c (computed
// at the converged amplitudes) with cotangents (lam1, lam2).  Writes
// (dG/dT)^T lam into A->tb; with th != nullptr also accumulates the integral
// and Fock cotangents.  With lowp (Lambda iterations before the FP64 sweeps)
// the five O(N^6) products run as BF16x3 GEMMs and the secondary products
// with O(N^5) work run on FP32 copies of their operands (lowprec_adj_setup,
// lowprec_adj_lambda), each accumulating into the FP64 cotangents.
static int reverse_pass(Ctx* c, const double* lam, ThetaBars* th, bool lowp = false) {
  AdjState* A = c->adj;
  const Dims& d = c->d;
  const long O = d.O, V = d.V, VV = V * V, po = d.po, pv = d.pv, ov = d.ov;
  hipStream_t s = c->stream;
  const double* t1 = c->amp;
  const double* lam1 = lam;
  const double* lam2 = lam + ov;
  double* t1b = A->tb;
  double* t2b = A->tb + ov;
  const GemmWorkspace* W = &A->ws;
  LowPrec* L = lowp ? c->lp : nullptr;
  auto zero = [&](double* x, long n) { LAUNCH(zero_kernel, n, x, n); };
  // FP32 product into g's FP64 destination (the adjoint's workspace: batch reductions need it)
  auto f32 = [&](const GemmDesc<float>& f) -> int {
    hipError_t e = gemm<float>(f, *W, s);
    if (e != hipSuccess) {
      c->last_error = hipGetErrorString(e);
      return CCSD_ERR_HIP;
    }
    return CCSD_OK;
  };
  // g in FP64, or (lowp) on the FP32 operands A32 / B32 laid out as g's
  auto mm = [&](const GemmDesc<double>& g, const float* A32, const float* B32) -> int {
    return lowp ? f32(f32_desc(g, A32, B32)) : run_gemm(c, g, W);
  };
  // The amplitude cotangents (t1b, t2b and the layouts that fold into them:
  // taub, tauub, T2ub, Tvb, Tringb) are needed by the Lambda iterations only.
  // With th (the gradient pass) nothing reads them -- no intermediate or
  // integral cotangent is formed from them -- so every product that only
  // feeds them is skipped there (the ring, ladder, W_mbej and W_mbij ones
  // among them: four O(N^6) products).
  const bool need_t = th == nullptr;
  if (lowp) TRY(lowprec_adj_lambda(c, lam));
  if (need_t) {
    zero(A->tb, c->nvec);
    zero(A->taub, po * pv);
    zero(A->T2ub, O * O * pv);
    zero(A->tauub, O * O * pv);
    zero(A->Tvb, po * VV);
    zero(A->Tringb, ov * ov);
  }
  zero(A->Fmeb, ov);
  zero(A->Fmib, O * O);
  zero(A->Faeb, VV);
  zero(A->W1b, po * po);
  zero(A->W3b, O * po * V);
  zero(A->W3ub, O * O * O * V);
  zero(A->W2b, po * ov);
  zero(A->WTb, ov * ov);

  // -- assembly: R2 = Rp + P(ij)P(ab) X + P(ab) Y + P(ij) Z
  if (!lowp) {  // low precision: unpacked into the BF16 ring operand and the FP32 Yb / Zb
    k_unpack_ring(A->Xb, lam2, d, s);
    k_unpack_vir(A->Yb, lam2, d, s);
    k_unpack_occ(A->Zb, A->Zb, lam2, lam2, d, s);
  }

  // -- T1 right-hand side
  if (need_t) LAUNCH(r1init_vjp_kernel, ov, t1b, lam1, c->f, d);  // t1 (1-delta) f_ae
  {  // -1/2 tau_im^ef <ma||ef>, summed over m
    GemmDesc<double> g = G64(lam1, V, 1, c->ovvv, pv, 1, A->tauub, O * pv, 1, O, pv, V, -1.0, 1.0);
    g.nb1 = (int)O; g.sB1 = V * pv; g.sC1 = pv;
    if (need_t && lowp) {
      GemmDesc<float> f = f32_desc(g, L->lam1_f, L->ovvv32);
      f.sBk = L->ld_pv; f.sB1 = V * L->ld_pv;
      TRY(f32(f));
    } else if (need_t) {
      TRY(run_gemm(c, g, W));
    }
    if (th) {
      GemmDesc<double> h = G64(lam1, 1, V, c->tauu, O * pv, 1, th->ovvv, pv, 1, V, pv, O, -1.0, 1.0);
      h.nb1 = (int)O; h.sB1 = pv; h.sC1 = V * pv;
      TRY(run_gemm(c, h, W));
    }
  }
  // + W3 . Tv_flat
  TRY(mm(G64(lam1, V, 1, c->Tv, 1, V, A->W3b, po * V, 1, O, po * V, V, 1.0, 1.0), L ? L->lam1_f : nullptr,
         L ? L->Tv_f : nullptr));
  if (need_t)
    TRY(mm(G64(c->W3, 1, po * V, lam1, V, 1, A->Tvb, V, 1, po * V, V, O, 1.0, 1.0), L ? L->W3_f : nullptr,
           L ? L->lam1_f : nullptr));
  // + Tring . Fme
  if (need_t) LAUNCH(outer_add_kernel, ov * ov, A->Tringb, lam1, c->Fme, ov, ov, 1.0);
  TRY(run_gemm(c, G64(c->Tring, ov, 1, lam1, 1, 0, A->Fmeb, 1, 0, ov, 1, ov, 1.0, 1.0), W));
  // + G . t1   (G symmetric)
  if (need_t)
    TRY(mm(G64(c->G, ov, 1, lam1, 1, 0, t1b, 1, 0, ov, 1, ov, 1.0, 1.0), L ? L->G_f : nullptr,
           L ? L->lam1_f : nullptr));
  if (th) LAUNCH(outer_add_kernel, ov * ov, th->G, lam1, t1, ov, ov, 1.0);
  // - Fmi^T . t1
  TRY(run_gemm(c, G64(t1, V, 1, lam1, 1, V, A->Fmib, O, 1, O, O, V, -1.0, 1.0), W));
  if (need_t) TRY(run_gemm(c, G64(c->Fmi, O, 1, lam1, V, 1, t1b, V, 1, O, V, O, -1.0, 1.0), W));
  if (th) {
    LAUNCH(scale_add_kernel, ov, th->fov, lam1, 1.0, ov);
    // (1-delta) f_ae: fvv_bar[a,e] += sum_i lam1[i,a] t1[i,e]  (diagonal dropped on output)
    TRY(run_gemm(c, G64(lam1, 1, V, t1, V, 1, th->fvv, V, 1, V, V, O, 1.0, 1.0), W));
  }

  // -- Z[i][j][ab] = -Fmi[m,j] T2u[i][m][ab] - t1[i,e] ovvv[(j,e)][ab]
  const float* Zf = L ? L->Zb_f : nullptr;
  {
    GemmDesc<double> g = G64(c->T2u, pv, 1, A->Zb, 1, pv, A->Fmib, O, 1, O, O, pv, -1.0, 1.0);
    g.nb1 = (int)O; g.sA1 = O * pv; g.sB1 = O * pv; g.reduce_batch = true;
    TRY(mm(g, L ? L->T2u_f : nullptr, Zf));
  }
  if (need_t) {
    GemmDesc<double> g = G64(c->Fmi, O, 1, A->Zb, pv, 1, A->T2ub, pv, 1, O, pv, O, -1.0, 1.0);
    g.nb1 = (int)O; g.sB1 = O * pv; g.sC1 = O * pv;
    TRY(mm(g, L ? L->Fmi_f : nullptr, Zf));
  }
  if (need_t) {
    GemmDesc<double> g = G64(A->Zb, O * pv, 1, c->ovvv, 1, pv, t1b, V, 1, O, V, pv, -1.0, 1.0);
    g.nb1 = (int)O; g.sA1 = pv; g.sB1 = V * pv; g.reduce_batch = true;
    if (lowp) {
      GemmDesc<float> f = f32_desc(g, Zf, L->ovvv32);
      f.sBn = L->ld_pv; f.sB1 = V * L->ld_pv;
      TRY(f32(f));
    } else {
      TRY(run_gemm(c, g, W));
    }
  }
  if (th) {
    GemmDesc<double> g = G64(t1, 1, V, A->Zb, O * pv, 1, th->ovvv, pv, 1, V, pv, O, -1.0, 1.0);
    g.nb1 = (int)O; g.sB1 = pv; g.sC1 = V * pv;
    TRY(run_gemm(c, g, W));
  }

  // -- Y[ij][a][b] = Tv[ij][a][e] Fae[b][e] - t1[m,a] W2[ij][(m,b)]
  const float* Yf = L ? L->Yb_f : nullptr;
  if (need_t) TRY(mm(G64(A->Yb, V, 1, c->Fae, V, 1, A->Tvb, V, 1, po * V, V, V, 1.0, 1.0), Yf, L ? L->Fae_f : nullptr));
  TRY(mm(G64(A->Yb, 1, V, c->Tv, V, 1, A->Faeb, V, 1, V, V, po * V, 1.0, 1.0), Yf, L ? L->Tv_f : nullptr));
  if (need_t) {
    GemmDesc<double> g = G64(c->W2, V, 1, A->Yb, 1, V, t1b, V, 1, O, V, V, -1.0, 1.0);
    g.nb1 = (int)po; g.sA1 = ov; g.sB1 = VV; g.reduce_batch = true;
    TRY(mm(g, L ? L->W2_f : nullptr, Yf));
  }
  {
    GemmDesc<double> g = G64(t1, V, 1, A->Yb, V, 1, A->W2b, V, 1, O, V, V, -1.0, 1.0);
    g.nb1 = (int)po; g.sB1 = VV; g.sC1 = ov;
    TRY(mm(g, L ? L->t1_f : nullptr, Yf));
  }

  // -- ring X = Tring . WT^T
  if (lowp) {
    TRY(lowprec_adj_ring(c, A->Tringb, A->WTb, lam2));
  } else {
    // the (OV)^3 products go through square_product (one level of Strassen at
    // OV >= 8192) with the forward's X as scratch: the reverse pass never reads X
    if (need_t) TRY(square_product(c, G64(A->Xb, ov, 1, c->WT, ov, 1, A->Tringb, ov, 1, ov, ov, ov, 1.0, 1.0), c->X, W));
    TRY(square_product(c, G64(A->Xb, 1, ov, c->Tring, ov, 1, A->WTb, ov, 1, ov, ov, ov, 1.0, 1.0), c->X, W));
  }

  // -- Rp = <ab||ij> + tau . vvvv + W1^T . tau     (cotangent lam2)
  if (lowp) TRY(lowprec_adj_ladder(c, A->taub, lam2));
  else if (need_t) TRY(run_gemm(c, G64(lam2, pv, 1, c->vvvv, pv, 1, A->taub, pv, 1, po, pv, pv, 1.0, 1.0), W));
  {
    GemmDesc<double> g = G64(c->tau, pv, 1, lam2, 1, pv, A->W1b, po, 1, po, po, pv, 1.0, 1.0);
    if (lowp) {
      GemmDesc<float> f = f32_desc(g, L->tau32, L->lam2_f);
      f.sAm = L->ld_pv; f.sBn = L->ld_pv;
      TRY(f32(f));
    } else {
      TRY(run_gemm(c, g, W));
    }
  }
  if (need_t) {
    GemmDesc<double> g = G64(c->W1, po, 1, lam2, pv, 1, A->taub, pv, 1, po, pv, po, 1.0, 1.0);
    if (lowp) {
      GemmDesc<float> f = f32_desc(g, L->W1_f, L->lam2_f);
      f.sBk = L->ld_pv;
      TRY(f32(f));
    } else {
      TRY(run_gemm(c, g, W));
    }
  }
  if (th) LAUNCH(scale_add_kernel, po * pv, th->oovv, lam2, 1.0, po * pv);

  // -- F_ae = (1-d) f + Fae_t1 - Fme^T t1 - Tv_flat^T oovv_vu_flat
  TRY(run_gemm(c, G64(t1, V, 1, A->Faeb, V, 1, A->Fmeb, V, 1, O, V, V, -1.0, 1.0), W));
  if (need_t) TRY(run_gemm(c, G64(c->Fme, V, 1, A->Faeb, 1, V, t1b, V, 1, O, V, V, -1.0, 1.0), W));
  if (lowp) k_to_f32(L->Faeb_f, A->Faeb, VV, s);  // F_ae's cotangent is complete here
  if (need_t)
    TRY(mm(G64(c->oovv_vu, V, 1, A->Faeb, 1, V, A->Tvb, V, 1, po * V, V, V, -1.0, 1.0), L ? L->oovv_vu_f : nullptr,
           L ? L->Faeb_f : nullptr));
  if (th) {
    TRY(run_gemm(c, G64(c->Tv, V, 1, A->Faeb, V, 1, th->oovv_vu, V, 1, po * V, V, V, -1.0, 1.0), W));
    LAUNCH(scale_add_kernel, VV, th->fvv, A->Faeb, 1.0, VV);
  }

  // -- F_mi = (1-d) f + t1 <mn||ie> + oovv_ou . T2u^T + Fme . t1^T
  TRY(run_gemm(c, G64(A->Fmib, O, 1, t1, V, 1, A->Fmeb, V, 1, O, V, O, 1.0, 1.0), W));
  if (need_t) TRY(run_gemm(c, G64(A->Fmib, 1, O, c->Fme, V, 1, t1b, V, 1, O, V, O, 1.0, 1.0), W));
  if (lowp) k_to_f32(L->Fmib_f, A->Fmib, O * O, s);  // complete here
  if (need_t) {
    TRY(mm(G64(A->Fmib, 1, O, c->oovv_ou, O * pv, 1, A->T2ub, O * pv, 1, O, O * pv, O, 1.0, 1.0),
           L ? L->Fmib_f : nullptr, L ? L->oovv_ou_f : nullptr));
    hipLaunchKernelGGL(fmi_t1_vjp_kernel, dim3(ov), dim3(kThreads), 0, s, t1b, A->Fmib, c->ooov, d);
  }
  if (th) {
    TRY(run_gemm(c, G64(A->Fmib, O, 1, c->T2u, O * pv, 1, th->oovv_ou, O * pv, 1, O, O * pv, O, 1.0, 1.0), W));
    LAUNCH(fmi_ooov_kernel, po * ov, th->ooov, A->Fmib, t1, d);
    // (1-delta) f_mi: copied into the Fock cotangent at the end (oo block)
  }

  // -- W_mbej (WT): (c) + 1/2 Tring . Doovv
  if (lowp) TRY(lowprec_adj_wmbej(c, A->Tringb, A->WTb));
  else if (need_t) {
    StrassenSums sd;  // B = Doovv with the forward W_mbej product's strides: its operand sums
    sd.T = c->strassen_doovv_t;
    TRY(square_product(c, G64(A->WTb, ov, 1, c->Doovv, ov, 1, A->Tringb, ov, 1, ov, ov, ov, 0.5, 1.0), c->X, W, &sd));
  }
  if (th) TRY(square_product(c, G64(c->Tring, ov, 1, A->WTb, ov, 1, th->Doovv, ov, 1, ov, ov, ov, 0.5, 1.0), c->X, W));
  // (b) + t1[n,b] W3u[j][m][n][e], batched over (j, m)
  if (need_t) {
    GemmDesc<double> g = G64(c->W3u, V, 1, A->WTb, 1, ov, t1b, V, 1, O, V, V, 1.0, 1.0);
    g.nb1 = (int)O; g.nb2 = (int)O; g.sA1 = O * O * V; g.sA2 = O * V; g.sB1 = V * ov; g.sB2 = V;
    g.reduce_batch = true;
    TRY(run_gemm(c, g, W));
  }
  {
    GemmDesc<double> g = G64(t1, V, 1, A->WTb, ov, 1, A->W3ub, V, 1, O, V, V, 1.0, 1.0);
    g.nb1 = (int)O; g.nb2 = (int)O; g.sB1 = V * ov; g.sB2 = V; g.sC1 = O * O * V; g.sC2 = O * V;
    TRY(run_gemm(c, g, W));
  }
  if (th) LAUNCH(scale_add_kernel, ov * ov, th->G, A->WTb, 1.0, ov * ov);  // + <mb||ej>
  // (a) S[(m,b,e)][j] = ovvv_u[(m,b,e)][f] t1[j][f]  (S_bar in Xb: Xb is free after the ring)
  double* Sb = A->Xb;
  if (lowp) {
    LAUNCH(sbar_kernel<float>, O * VV * O, L->Sb_f, A->WTb, A->Faeb, d);
    TRY(f32(f32_desc(G64(Sb, 1, O, c->ovvv_u, V, 1, t1b, V, 1, O, V, O * VV, 1.0, 1.0), L->Sb_f, L->ovvv_u_f)));
  } else {
    LAUNCH(sbar_kernel<double>, O * VV * O, Sb, A->WTb, A->Faeb, d);
    if (need_t) TRY(run_gemm(c, G64(Sb, 1, O, c->ovvv_u, V, 1, t1b, V, 1, O, V, O * VV, 1.0, 1.0), W));
  }
  if (th) TRY(run_gemm(c, G64(Sb, O, 1, t1, V, 1, th->ovvv_u, V, 1, O * VV, V, O, 1.0, 1.0), W));

  // -- W_mbij: W2 = ooov + tau . ovvv^T + P(ij) U,  U[i][j][(m,b)] = t1[i,e] G[(m,e),(j,b)]
  if (lowp)
    hipLaunchKernelGGL(unpack_pair_rows_kernel<float>, dim3(grid_for(O * O * ov)), dim3(kThreads), 0, s, L->Ub_f,
                       A->W2b, (int)O, ov);
  else
    hipLaunchKernelGGL(unpack_pair_rows_kernel<double>, dim3(grid_for(O * O * ov)), dim3(kThreads), 0, s, A->Ub,
                       A->W2b, (int)O, ov);
  if (need_t) {
    GemmDesc<double> g = G64(A->Ub, O * ov, 1, c->G, 1, ov, t1b, V, 1, O, V, V, 1.0, 1.0);
    g.nb1 = (int)O; g.nb2 = (int)O;  // b1 = m, b2 = j
    g.sA1 = V; g.sA2 = ov; g.sB1 = V * ov; g.sB2 = V; g.reduce_batch = true;
    TRY(mm(g, L ? L->Ub_f : nullptr, L ? L->G_f : nullptr));
  }
  if (th) {
    GemmDesc<double> g = G64(t1, 1, V, A->Ub, O * ov, 1, th->G, ov, 1, V, V, O, 1.0, 1.0);
    g.nb1 = (int)O; g.nb2 = (int)O;
    g.sB1 = V; g.sB2 = ov; g.sC1 = V * ov; g.sC2 = V;
    TRY(run_gemm(c, g, W));
  }
  if (lowp) TRY(lowprec_adj_w2(c, A->taub, A->W2b));
  else if (need_t) TRY(run_gemm(c, G64(A->W2b, ov, 1, c->ovvv, pv, 1, A->taub, pv, 1, po, pv, ov, 1.0, 1.0), W));
  if (th) {
    TRY(run_gemm(c, G64(A->W2b, 1, ov, c->tau, pv, 1, th->ovvv, pv, 1, ov, pv, po, 1.0, 1.0), W));
    LAUNCH(scale_add_kernel, po * ov, th->ooov, A->W2b, 1.0, po * ov);
  }

  // -- W_mnie: W3[i][mn][e] = ooov[mn][i,e] + t1[i,f] oovv_vu[mn][f][e];  W3u = unpack(W3)
  hipLaunchKernelGGL(pack_pair_rows_adj_kernel, dim3(grid_for(O * po * V)), dim3(kThreads), 0, s, A->W3b, A->W3ub,
                     (int)O, V, O);
  if (lowp) k_to_f32(L->W3b_f, A->W3b, O * po * V, s);  // complete here
  if (need_t)
    TRY(mm(G64(A->W3b, po * V, 1, c->oovv_vu, V, 1, t1b, V, 1, O, V, po * V, -1.0, 1.0), L ? L->W3b_f : nullptr,
           L ? L->oovv_vu_f : nullptr));
  if (th) {
    GemmDesc<double> g = G64(t1, 1, V, A->W3b, po * V, 1, th->oovv_vu, V, 1, V, V, O, 1.0, 1.0);
    g.nb1 = (int)po; g.sB1 = V; g.sC1 = VV;
    TRY(run_gemm(c, g, W));
    LAUNCH(w3_ooov_kernel, po * ov, th->ooov, A->W3b, d);
  }

  // -- W_mnij: W1 = oooo + oovv . tau^T + P(ij) Q,  Q[(mn,i)][j] = ooov[(mn,i)][e] t1[j][e]
  hipLaunchKernelGGL(unpack_pair_cols_kernel, dim3(grid_for(po * O * O)), dim3(kThreads), 0, s, A->Qb, A->W1b, po,
                     (int)O);
  if (need_t) TRY(run_gemm(c, G64(A->Qb, 1, O, c->ooov, V, 1, t1b, V, 1, O, V, po * O, 1.0, 1.0), W));
  if (lowp) k_to_f32(L->W1b_f, A->W1b, po * po, s);  // complete here
  if (need_t)
    TRY(mm(G64(A->W1b, 1, po, c->oovv, pv, 1, A->taub, pv, 1, po, pv, po, 1.0, 1.0), L ? L->W1b_f : nullptr,
           L ? L->oovv_f : nullptr));
  if (th) {
    TRY(run_gemm(c, G64(A->Qb, O, 1, t1, V, 1, th->ooov, V, 1, po * O, V, O, 1.0, 1.0), W));
    TRY(run_gemm(c, G64(A->W1b, po, 1, c->tau, pv, 1, th->oovv, pv, 1, po, pv, po, 1.0, 1.0), W));
    LAUNCH(scale_add_kernel, po * po, th->oooo, A->W1b, 1.0, po * po);
  }

  // -- F_me = f_me + Doovv . t1
  if (need_t) TRY(run_gemm(c, G64(c->Doovv, ov, 1, A->Fmeb, 1, 0, t1b, 1, 0, ov, 1, ov, 1.0, 1.0), W));
  if (th) {
    LAUNCH(outer_add_kernel, ov * ov, th->Doovv, A->Fmeb, t1, ov, ov, 1.0);
    LAUNCH(scale_add_kernel, ov, th->fov, A->Fmeb, 1.0, ov);
  }

  // -- amplitude layouts: t2_bar = pack^T(Tring_bar) + pack^T(Tv_bar) + pack^T(T2u_bar); tau_bar += pack^T(tauu_bar)
  if (!need_t) {
    HIPTRY(hipGetLastError());
    return CCSD_OK;
  }
  k_assemble_t2(t2b, nullptr, t2b, A->Tringb, A->Tvb, A->T2ub, c->f, nullptr, 0, d, s);
  hipLaunchKernelGGL(pack_pair_rows_adj_kernel, dim3(grid_for(po * pv)), dim3(kThreads), 0, s, A->taub, A->tauub,
                     (int)O, pv, 1L);
  // -- tau = t2 + t1 t1 - t1 t1
  LAUNCH(scale_add_kernel, po * pv, t2b, A->taub, 1.0, po * pv);
  hipLaunchKernelGGL(tau_vjp_t1_kernel, dim3(ov), dim3(kThreads), 0, s, t1b, A->taub, t1, d);
  HIPTRY(hipGetLastError());
  return CCSD_OK;
}

static int maxabs(Ctx* c, const double* x, long n, double* out) {
  const int nb = reduce_blocks_for(n);
  k_maxabs_partials(c->adj->partial, x, n, nb, c->stream);
  std::vector<double> h(nb);
  HIPTRY(hipMemcpyAsync(h.data(), c->adj->partial, nb * 8, hipMemcpyDeviceToHost, c->stream));
  HIPTRY(hipStreamSynchronize(c->stream));
  double m = 0.0;
  for (double v : h) m = std::max(m, v);
  *out = m;
  return CCSD_OK;
}

int adjoint_run(Ctx* c, const GradOut& g, ccsd_grad_info* info) {
  auto t0 = std::chrono::steady_clock::now();
  const Dims& d = c->d;
  const long O = d.O, V = d.V, N = O + V, VV = V * V, po = d.po, pv = d.pv, ov = d.ov;
  hipStream_t s = c->stream;
  TRY(adjoint_prepare(c));
  AdjState* A = c->adj;
  const double* t1 = c->amp;
  const double* t2 = c->amp + ov;

  // forward intermediates at the converged amplitudes (FP64)
  TRY(compute_rhs(c, t1, t2, CCSD_PREC_FP64, false));

  // dE/dT: t2 part <ij||ab>; t1 part f_ia + tau-VJP of <ij||ab>
  HIPTRY(hipMemcpyAsync(A->e + ov, c->oovv, po * pv * 8, hipMemcpyDeviceToDevice, s));
  HIPTRY(hipMemcpyAsync(A->e, c->fov, ov * 8, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(tau_vjp_t1_kernel, dim3(ov), dim3(kThreads), 0, s, A->e, c->oovv, t1, d);

  // Lambda iterations (Jacobi + DIIS on the forward's DIIS buffers).  FP64
  // runs: FP64 sweeps until max|dLambda| < tol.  BF16x3 runs: BF16x3 reverse
  // products until the switch, then iterative refinement -- the equations are
  // linear, Lambda = (e + T(Lambda)) / D, so at the current Lambda one FP64
  // pass gives the exact update difference r0 = (e + T(Lambda)) / D - Lambda,
  // and the correction dLambda = r0 + T(dLambda) / D is solved with BF16x3
  // reverse passes (their rounding scales with |dLambda|, orders of magnitude
  // below the tolerance); Lambda += dLambda, and the next FP64 pass either
  // meets max|dLambda| < tol (converged: the FP64 equations hold to tol, as in
  // an FP64 run) or starts another correction.  At O=40/V=300 this takes two
  // FP64 passes where FP64 sweeps with the BF16x3 phase's DIIS subspace took four.
  LAUNCH(zero_kernel, c->nvec, A->lam, c->nvec);
  const int m = c->diis_m;
  struct Diis {
    int count = 0, next = 0;
    std::vector<double> B;
  };
  Diis dl;
  dl.B.assign((size_t)std::max(m, 1) * std::max(m, 1), 0.0);
  // Pulay step on the forward's DIIS buffers: stores (vnew, verr) and writes
  // the extrapolation into *out, or swaps vnew into *out without one
  auto diis_step = [&](Diis& st, double*& vnew, const double* verr, double*& out) -> int {
    bool extrap = false;
    if (m >= 2) {
      const int slot = st.next;
      HIPTRY(hipMemcpyAsync(c->dvec + (long)slot * c->nvec, vnew, c->nvec * 8, hipMemcpyDeviceToDevice, s));
      HIPTRY(hipMemcpyAsync(c->derr + (long)slot * c->nvec, verr, c->nvec * 8, hipMemcpyDeviceToDevice, s));
      st.count = std::min(st.count + 1, m);
      st.next = (slot + 1) % m;
      const int nbd = reduce_blocks_for(c->nvec);
      k_multidot_partials(c->partial, c->derr, c->nvec, st.count, verr, c->nvec, nbd, s);
      k_sum_partials(c->scal, c->partial, nbd, st.count, s);
      std::vector<double> row(st.count);
      HIPTRY(hipMemcpyAsync(row.data(), c->scal, st.count * 8, hipMemcpyDeviceToHost, s));
      HIPTRY(hipStreamSynchronize(s));
      for (int k = 0; k < st.count; ++k) { st.B[(size_t)slot * m + k] = row[k]; st.B[(size_t)k * m + slot] = row[k]; }
      if (st.count >= 2) {
        std::vector<double> Bc((size_t)st.count * st.count), coef;
        for (int i = 0; i < st.count; ++i)
          for (int j = 0; j < st.count; ++j) Bc[(size_t)i * st.count + j] = st.B[(size_t)i * m + j];
        if (diis_solve_bordered(Bc, st.count, coef)) {
          HIPTRY(hipMemcpyAsync(c->dcoef, coef.data(), st.count * 8, hipMemcpyHostToDevice, s));
          k_lincomb(out, c->dvec, c->nvec, c->dcoef, st.count, c->nvec, s);
          extrap = true;
        }
      }
    }
    if (!extrap) std::swap(out, vnew);
    return CCSD_OK;
  };
  int it = 0, lowp_its = 0, corr_its = 0, fp64_after = c->opt.final_fp64_sweeps;
  double delta = 1e300;
  // Lambda tolerance: the option, else 1e-2 x the forward's (clamped): the
  // gradient's error is then set by where the amplitudes stopped
  const double tol = c->opt.lambda_tol > 0.0 ? c->opt.lambda_tol
                                              : std::min(1e-8, std::max(1e-11, 1e-2 * c->last_tol));
  bool conv = false, lowp = c->opt.precision == CCSD_PREC_BF16X3, refine = false;
  double best_low = HUGE_VAL;  // smallest max|dLambda| of the BF16x3 phase
  int since_best = 0;
  if (lowp) {
    TRY(lowprec_adj_setup(c));
    fp64_after = 0;
  }
  // One correction solve (BF16x3, fresh DIIS subspace) from r0 = lerr, then
  // lam += dcorr.  Stops once max|d(dcorr)| < kCorrTol * tol, or when it stalls.
  // (O=40/V=300, tol 1e-9: 0.3 takes 5 passes and the verifying FP64 pass
  // lands at 1.1e-10; 0.1 and 0.05 take 7, at 9e-12 -- same gradient.)
  constexpr double kCorrTol = 0.3;
  constexpr int kCorrMax = 20;
  auto correct = [&]() -> int {
    if (!A->dcorr) {
      TRY(adj_alloc(c, A, &A->dcorr, c->nvec));
      TRY(adj_alloc(c, A, &A->rcorr, c->nvec));
    }
    HIPTRY(hipMemcpyAsync(A->rcorr, A->lerr, c->nvec * 8, hipMemcpyDeviceToDevice, s));
    HIPTRY(hipMemcpyAsync(A->dcorr, A->lerr, c->nvec * 8, hipMemcpyDeviceToDevice, s));
    Diis dc;
    dc.B.assign((size_t)std::max(m, 1) * std::max(m, 1), 0.0);
    double best = HUGE_VAL, dd = HUGE_VAL;
    int since = 0;
    for (int k = 0; k < kCorrMax; ++k) {
      TRY(reverse_pass(c, A->dcorr, nullptr, true));
      ++lowp_its;
      ++corr_its;
      LAUNCH(correction_update_kernel, c->nvec, A->lnew, A->lerr, A->rcorr, A->tb, A->dcorr, c->f, d);
      TRY(maxabs(c, A->lerr, c->nvec, &dd));
      if (!std::isfinite(dd)) return CCSD_ERR_NONFINITE;
      TRY(diis_step(dc, A->lnew, A->lerr, A->dcorr));
      if (dd < kCorrTol * tol) break;
      if (dd < best) {
        best = dd;
        since = 0;
      } else if (++since >= kLowpStall) {
        break;
      }
    }
    LAUNCH(scale_add_kernel, c->nvec, A->lam, A->dcorr, 1.0, c->nvec);
    return CCSD_OK;
  };
  for (it = 1; it <= 500; ++it) {
    TRY(reverse_pass(c, A->lam, nullptr, lowp));
    if (lowp) ++lowp_its;
    else ++fp64_after;
    LAUNCH(lambda_update_kernel, c->nvec, A->lnew, A->lerr, A->e, A->tb, A->lam, c->f, d);
    TRY(maxabs(c, A->lerr, c->nvec, &delta));
    if (!std::isfinite(delta)) return CCSD_ERR_NONFINITE;
    if (refine) {
      if (delta < tol) {
        std::swap(A->lam, A->lnew);
        conv = true;
        break;
      }
      TRY(correct());
      continue;
    }
    TRY(diis_step(dl, A->lnew, A->lerr, A->lam));
    if (lowp) {
      const double f = c->opt.lambda_switch_factor > 0 ? c->opt.lambda_switch_factor : c->opt.switch_factor;
      if (delta < best_low) {
        best_low = delta;
        since_best = 0;
      } else {
        ++since_best;
      }
      // switch once max|dLambda| is near the tolerance, or once the BF16x3
      // phase stalls at its rounding floor
      if (delta < std::max(f * tol, tol) || since_best >= kLowpStall) {
        lowp = false;
        refine = true;
      }
      continue;
    }
    if (delta < tol && fp64_after >= c->opt.final_fp64_sweeps) { conv = true; break; }
  }

  // gradient pass.  The four-virtual block dE/d<ef||ab> = 1/2 (tau^T lam2 +
  // lam2^T tau) needs only tau and the converged Lambda, so its slabs are formed
  // on a stream of their own and copied to the host on a second one while the
  // main stream runs the reverse pass with the integral cotangents; the other
  // blocks go out on a third stream once that pass and its folds are done.
  // (All streams at the default priority; page-locked destinations make the
  // copies DMA transfers.)
  ThetaBars th;
  std::vector<void*> tmp;
  hipStream_t vs = nullptr, cs = nullptr, cs2 = nullptr;
  hipEvent_t ev_lam = nullptr, ev_fold = nullptr, ev_comp[2] = {nullptr, nullptr}, ev_copy[2] = {nullptr, nullptr};
  double* slab[2] = {nullptr, nullptr};
  auto teardown = [&]() {
    for (hipStream_t q : {vs, cs, cs2})
      if (q) (void)hipStreamSynchronize(q);
    (void)hipStreamSynchronize(s);
    for (double* p : slab)
      if (p) (void)hipFree(p);
    for (hipEvent_t e : {ev_lam, ev_fold, ev_comp[0], ev_comp[1], ev_copy[0], ev_copy[1]})
      if (e) (void)hipEventDestroy(e);
    for (hipStream_t q : {vs, cs, cs2})
      if (q) (void)hipStreamDestroy(q);
    for (void* p : tmp) (void)hipFree(p);
  };
  auto talloc = [&](double** p, long n) -> int {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, (size_t)std::max<long>(n, 1) * 8);
    if (e != hipSuccess) {
      c->last_error = hipGetErrorString(e);
      return CCSD_ERR_HIP;
    }
    tmp.push_back(q);
    *p = (double*)q;
    LAUNCH(zero_kernel, n, *p, n);
    return CCSD_OK;
  };
  int rc = CCSD_OK;
  if ((rc = talloc(&th.oooo, po * po)) || (rc = talloc(&th.ooov, po * ov)) || (rc = talloc(&th.oovv, po * pv)) ||
      (rc = talloc(&th.ovvv, ov * pv)) || (rc = talloc(&th.G, ov * ov)) || (rc = talloc(&th.Doovv, ov * ov)) ||
      (rc = talloc(&th.oovv_ou, O * O * pv)) || (rc = talloc(&th.oovv_vu, po * VV)) ||
      (rc = talloc(&th.ovvv_u, ov * VV)) || (rc = talloc(&th.fvv, VV)) || (rc = talloc(&th.fov, ov)) ||
      (rc = talloc(&th.f, N + 8))) {
    teardown();
    return rc;
  }
  // Four-virtual block policy (recompute_v4): 0 forms the whole P2(V)^2 block in
  // HBM, 1 streams it through two 512 MB row-slab buffers, < 0 (default) forms it
  // whole when free memory holds it with 4 GB to spare and streams it otherwise.
  // Whole block: its row slabs are formed on the main stream before the reverse
  // pass (at full speed) and each is copied out on cs while that pass runs.
  // Slabs: they are formed on vs, each waiting for the copy of the slab two
  // before it, so the reverse pass on the main stream is never held back.
  const long slab_rows = std::max<long>(1, std::min<long>(pv, (512L << 20) / 8 / pv));
  const size_t whole_bytes = (size_t)pv * pv * 8;
  bool whole = false;
  if (c->opt.recompute_v4 != 1) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      teardown();
      return CCSD_ERR_HIP;
    }
    whole = c->opt.recompute_v4 == 0 ? whole_bytes <= fr : whole_bytes + (4ull << 30) <= fr;
    if (!whole && c->opt.recompute_v4 == 0) {
      c->last_error = "four-virtual gradient block does not fit; set recompute_v4";
      teardown();
      return CCSD_ERR_BUDGET;
    }
  }
  const int nbuf = whole ? 1 : (slab_rows >= pv ? 1 : 2);
  for (int b = 0; b < nbuf; ++b)
    if (hipMalloc(&slab[b], whole ? whole_bytes : (size_t)slab_rows * pv * 8) != hipSuccess) {
      teardown();
      return whole ? CCSD_ERR_BUDGET : CCSD_ERR_HIP;
    }
  bool ok = hipStreamCreateWithFlags(&vs, hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&cs2, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&ev_lam, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&ev_fold, hipEventDisableTiming) == hipSuccess;
  for (int b = 0; b < 2 && ok; ++b)
    ok = hipEventCreateWithFlags(&ev_comp[b], hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&ev_copy[b], hipEventDisableTiming) == hipSuccess;
  ok = ok && hipEventRecord(ev_lam, s) == hipSuccess && hipStreamWaitEvent(vs, ev_lam, 0) == hipSuccess;
  if (!ok) {
    teardown();
    return CCSD_ERR_HIP;
  }
  {
    const double* lam2 = A->lam + ov;
    const GemmWorkspace none{};  // vs runs beside the main stream's GEMMs, which hold A->ws
    hipStream_t gs = whole ? s : vs;
    const GemmWorkspace& wsl = whole ? A->ws : none;
    long k = 0;
    for (long r0 = 0; r0 < pv && ok; r0 += slab_rows, ++k) {
      const long nr = std::min(slab_rows, pv - r0);
      const int b = (int)(k % nbuf);
      double* dst = whole ? slab[0] + r0 * pv : slab[b];
      if (!whole && k >= nbuf && hipStreamWaitEvent(vs, ev_copy[b], 0) != hipSuccess) ok = false;
      ok = ok && gemm<double>(G64(c->tau + r0, 1, pv, lam2, pv, 1, dst, pv, 1, nr, pv, po, 0.5, 0.0), wsl, gs) ==
                     hipSuccess;
      ok = ok && gemm<double>(G64(lam2 + r0, 1, pv, c->tau, pv, 1, dst, pv, 1, nr, pv, po, 0.5, 1.0), wsl, gs) ==
                     hipSuccess;
      ok = ok && hipEventRecord(ev_comp[b], gs) == hipSuccess && hipStreamWaitEvent(cs, ev_comp[b], 0) == hipSuccess &&
           hipMemcpyAsync(g.vvvv + r0 * pv, dst, (size_t)nr * pv * 8, hipMemcpyDeviceToHost, cs) == hipSuccess &&
           hipEventRecord(ev_copy[b], cs) == hipSuccess;
    }
    if (!ok) {
      teardown();
      return CCSD_ERR_HIP;
    }
  }
  rc = reverse_pass(c, A->lam, &th);
  if (rc != CCSD_OK) { teardown(); return rc; }
  // dE/dtheta: <ij||ab> pairs with tau, f_ov with t1
  LAUNCH(scale_add_kernel, po * pv, th.oovv, c->tau, 1.0, po * pv);
  LAUNCH(scale_add_kernel, ov, th.fov, t1, 1.0, ov);
  // folds of the derived layouts back onto the packed blocks
  k_assemble_t2(th.oovv, nullptr, th.oovv, th.Doovv, th.oovv_vu, th.oovv_ou, c->f, nullptr, 0, d, s);
  hipLaunchKernelGGL(pack_pair_cols_adj_kernel, dim3(grid_for(ov * pv)), dim3(kThreads), 0, s, th.ovvv, th.ovvv_u, ov,
                     (int)V);
  double* ovovb = th.Doovv;  // reuse: Doovv cotangent already folded
  LAUNCH(ovov_from_gbar_kernel, ov * ov, ovovb, th.G, d);
  hipLaunchKernelGGL(symmetrize_kernel, dim3(grid_for(ov * (ov - 1) / 2)), dim3(kThreads), 0, s, ovovb, ov);
  hipLaunchKernelGGL(symmetrize_kernel, dim3(grid_for(po * (po - 1) / 2)), dim3(kThreads), 0, s, th.oooo, po);
  // denominator term on the Fock diagonal
  hipLaunchKernelGGL(fdiag_kernel, dim3(N), dim3(kThreads), 0, s, th.f, A->lam, c->amp, d);
  if (hipGetLastError() != hipSuccess) { teardown(); return CCSD_ERR_HIP; }
  std::vector<double> fov(ov), fvv(VV), fmi(O * O), fd(N);
  auto d2h = [&](double* h, const double* dv, long n) {
    return hipMemcpyAsync(h, dv, (size_t)n * 8, hipMemcpyDeviceToHost, cs2) == hipSuccess;
  };
  ok = hipEventRecord(ev_fold, s) == hipSuccess && hipStreamWaitEvent(cs2, ev_fold, 0) == hipSuccess;
  ok = ok && d2h(g.oooo, th.oooo, po * po) && d2h(g.ooov, th.ooov, po * ov) && d2h(g.oovv, th.oovv, po * pv) &&
       d2h(g.ovov, ovovb, ov * ov) && d2h(g.ovvv, th.ovvv, ov * pv) && d2h(fov.data(), th.fov, ov) &&
       d2h(fvv.data(), th.fvv, VV) && d2h(fmi.data(), A->Fmib, O * O) && d2h(fd.data(), th.f, N);
  for (hipStream_t q : {cs2, cs, vs, s})
    ok = ok && hipStreamSynchronize(q) == hipSuccess;
  teardown();
  if (!ok) return CCSD_ERR_HIP;

  // Fock cotangent: ov block (used as f[i][O+a]), oo / vv off-diagonal, diagonal from D; symmetrised
  std::vector<double> G((size_t)N * N, 0.0);
  for (long i = 0; i < O; ++i)
    for (long a = 0; a < V; ++a) G[i * N + O + a] = fov[i * V + a];
  for (long m2 = 0; m2 < O; ++m2)
    for (long i = 0; i < O; ++i)
      if (m2 != i) G[m2 * N + i] = fmi[m2 * O + i];
  for (long a = 0; a < V; ++a)
    for (long e = 0; e < V; ++e)
      if (a != e) G[(O + a) * N + O + e] = fvv[a * V + e];
  for (long p = 0; p < N; ++p) G[p * N + p] = fd[p];
  for (long p = 0; p < N; ++p)
    for (long q = 0; q < N; ++q) g.fock[p * N + q] = 0.5 * (G[p * N + q] + G[q * N + p]);
  if (info) {
    info->lambda_iterations = std::min(it, 500) + corr_its;
    info->lambda_lowp_iterations = lowp_its;
    info->converged = conv ? 1 : 0;
    info->lambda_residual = delta;
    info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return conv ? CCSD_OK : CCSD_ERR_NOT_CONVERGED;
}

}  // namespace ccsd
