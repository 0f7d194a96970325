// Closed-shell (spin-adapted) CCSD for RHF inputs on the device: the spin-block
// mode of the reference's RHF storage.
//
// The reference iterates spin-orbital amplitudes of an RHF reference in compact
// storage and skips the spin-forbidden blocks by spin conservation
// (ccsd_memory_optimized.cpp:483-485, the even/odd loop pairs); for a
// closed-shell reference every spin block follows from one spatial amplitude
// set t1[i][a], t2[i][j][a][b] (the alpha-beta block), so this path iterates
// those directly: o^2 v^4 / 4 ladder flops against the spin-orbital path's
// (2o)^2 (2v)^4 / 4, and 1/16 of the ring flops.  Every contraction of the term
// sequence (rplan.hpp) runs on the strided MFMA GEMM engine; integrals are kept
// in HBM in the layouts their contractions read.  Driver semantics as in
// capi.hip: MP2 start, Jacobi updates with FP64 DIIS, |dE| < tol.  FP64 only.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "engine.hpp"
#include "rplan.hpp"

namespace ccsd {

struct RCtx {
  int device = 0;
  hipStream_t stream = nullptr;
  ccsd_options opt{};
  long o = 0, v = 0;
  size_t bytes = 0;
  std::vector<void*> allocs;
  std::vector<double> h_eo, h_ev;
  double *eo = nullptr, *ev = nullptr;
  rplan::RStatic S;
  rplan::RWork W;
  long nvec = 0;  // o v + o^2 v^2
  double *amp = nullptr, *amp_new = nullptr, *err = nullptr, *R = nullptr;
  int diis_m = 0, diis_count = 0, diis_next = 0;
  double *dvec = nullptr, *derr = nullptr, *dcoef = nullptr;
  std::vector<double> B;
  double *partial = nullptr, *scal = nullptr;
  GemmWorkspace ws;
  rplan::Scratch sc;
  double e_mp2 = 0.0;
  double timings_ms[1] = {0};  // device time of the last residual
  hipEvent_t ev_[4] = {};
  std::string last_error;
};

}  // namespace ccsd

struct ccsd_rhf_ctx {
  ccsd::RCtx c;
};

namespace ccsd {
namespace {

constexpr int kThreadsR = 256;
inline int rgrid(long n) { return (int)std::max<long>(1, std::min<long>((n + kThreadsR - 1) / kThreadsR, 1L << 20)); }
#define RGRID_LOOP(idx, n) \
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < (n); idx += (long)gridDim.x * blockDim.x)

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

struct CopyShape {
  int nd;
  long ext[rplan::kMaxDims], sstr[rplan::kMaxDims], dstr[rplan::kMaxDims];
};

// dst = alpha src + beta dst over one shape; element e is decomposed row-major
// over the shape (the last dimension fastest), so a destination with unit
// last stride is written coalesced.  I: index type (32-bit when every offset fits).
template <typename I>
__global__ void strided_copy_kernel(double* dst, const double* src, CopyShape sh, long n, double alpha, double beta) {
  RGRID_LOOP(e, n) {
    I r = (I)e, so = 0, dof = 0;
#pragma unroll
    for (int t = rplan::kMaxDims - 1; t >= 0; --t) {
      if (t >= sh.nd) continue;
      const I ext = (I)sh.ext[t];
      const I q = r / ext, i = r - q * ext;
      r = q;
      so += i * (I)sh.sstr[t];
      dof += i * (I)sh.dstr[t];
    }
    const double x = alpha * src[so];
    dst[dof] = (beta != 0.0) ? x + beta * dst[dof] : x;
  }
}

// Transposing copy: the source is contiguous along dimension ds and the
// destination along dimension dd (ds != dd).  A workgroup moves 32 x 32 tiles
// of the (ds, dd) plane through LDS so that reads and writes are both
// coalesced; grid z (and a loop) covers the remaining dimensions, whose
// extents and strides are in sh (with the tile dimensions' extents set to 1).
__global__ __launch_bounds__(256) void tiled_copy_kernel(double* dst, const double* src, CopyShape sh, long ext_s,
                                                         long ext_d, long sstr_d, long dstr_s, long nrest,
                                                         double alpha, double beta) {
  __shared__ double tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const long s0 = (long)blockIdx.x * 32, d0 = (long)blockIdx.y * 32;
  for (long z = blockIdx.z; z < nrest; z += gridDim.z) {
    long r = z, so = 0, dof = 0;
#pragma unroll
    for (int t = rplan::kMaxDims - 1; t >= 0; --t) {
      if (t >= sh.nd) continue;
      const long q = r / sh.ext[t], i = r - q * sh.ext[t];
      r = q;
      so += i * sh.sstr[t];
      dof += i * sh.dstr[t];
    }
#pragma unroll
    for (int y = 0; y < 32; y += 8) {
      const long is = s0 + tx, id = d0 + y + ty;
      if (is < ext_s && id < ext_d) tile[y + ty][tx] = src[so + is + id * sstr_d];
    }
    __syncthreads();
#pragma unroll
    for (int y = 0; y < 32; y += 8) {
      const long id = d0 + tx, is = s0 + y + ty;
      if (is < ext_s && id < ext_d) {
        double* q = dst + dof + id + is * dstr_s;
        const double x = alpha * tile[tx][y + ty];
        *q = (beta != 0.0) ? x + beta * *q : x;
      }
    }
    __syncthreads();
  }
}

// tau[i][j][a][b] = t2 + t1[i][a] t1[j][b]
__global__ void rtau_kernel(double* tau, const double* t1, const double* t2, long o, long v) {
  const long n = o * o * v * v;
  RGRID_LOOP(e, n) {
    const long b = e % v, a = (e / v) % v, j = (e / (v * v)) % o, i = e / (o * v * v);
    tau[e] = t2[e] + t1[i * v + a] * t1[j * v + b];
  }
}

// Denominator of amplitude e of the vector [t1 (o v) | t2 (o o v v)].
__device__ __forceinline__ double rden(long e, const double* eo, const double* ev, long o, long v) {
  if (e < o * v) return eo[e / v] - ev[e % v];
  const long f = e - o * v;
  const long b = f % v, a = (f / v) % v, j = (f / (v * v)) % o, i = f / (o * v * v);
  return eo[i] + eo[j] - ev[a] - ev[b];
}

// new = R / D, err = new - old
__global__ void rupdate_kernel(double* out, double* err, const double* R, const double* old, const double* eo,
                               const double* ev, long o, long v, long n) {
  RGRID_LOOP(e, n) {
    const double x = R[e] / rden(e, eo, ev, o, v);
    out[e] = x;
    err[e] = x - old[e];
  }
}

// MP2 start: t1 = 0, t2[i][j][a][b] = (ia|jb) / D
__global__ void rmp2_kernel(double* amp, const double* ovovT, const double* eo, const double* ev, long o, long v,
                            long n) {
  RGRID_LOOP(e, n) amp[e] = (e < o * v) ? 0.0 : ovovT[e - o * v] / rden(e, eo, ev, o, v);
}

// Pair decode: p1 index k = q(q+1)/2 + p (p <= q), p0 index k = q(q-1)/2 + p (p < q).
__device__ __forceinline__ void dec1(long k, long& p, long& q) {
  q = (long)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
  while ((q + 1) * (q + 2) / 2 <= k) ++q;
  while (q * (q + 1) / 2 > k) --q;
  p = k - q * (q + 1) / 2;
}
__device__ __forceinline__ void dec0(long k, long& p, long& q) {
  q = (long)((sqrt(8.0 * (double)k + 1.0) + 1.0) * 0.5);
  while (q * (q - 1) / 2 > k) --q;
  while ((q + 1) * q / 2 <= k) ++q;
  p = k - q * (q - 1) / 2;
}
__device__ __forceinline__ long pk1(long p, long q) { return p <= q ? q * (q + 1) / 2 + p : p * (p + 1) / 2 + q; }
__device__ __forceinline__ long pk0(long p, long q) { return p < q ? q * (q - 1) / 2 + p : p * (p - 1) / 2 + q; }

struct View4 {
  const double* p;
  long s0, s1, s2, s3;
};

// Vp[(ab)][(cd)] = (ac|bd) + (ad|bc) over a <= b, c <= d; Vm = (ac|bd) - (ad|bc)
// over a < b, c < d; rows of pitch l1 / l0
__global__ void rvpm_kernel(double* Vp, double* Vm, View4 w, long v, long l1, long l0) {
  const long n1 = v * (v + 1) / 2;
  RGRID_LOOP(e, n1 * n1) {
    long a, b, c, d;
    dec1(e / n1, a, b);
    dec1(e % n1, c, d);
    const double x = w.p[a * w.s0 + c * w.s1 + b * w.s2 + d * w.s3];
    const double y = w.p[a * w.s0 + d * w.s1 + b * w.s2 + c * w.s3];
    Vp[(e / n1) * l1 + e % n1] = x + y;
    if (a < b && c < d) Vm[pk0(a, b) * l0 + pk0(c, d)] = x - y;
  }
}

// tp[(ij)][(cd)] = (tau_ij^cd + tau_ij^dc) / 2, halved again on c = d (i <= j, c <= d);
// tm[(ij)][(cd)] = (tau_ij^cd - tau_ij^dc) / 2 (i < j, c < d); rows of pitch l1 / l0
__global__ void rladder_pack_kernel(double* tp, double* tm, const double* tau, long o, long v, long l1, long l0) {
  const long n1 = v * (v + 1) / 2, m1 = o * (o + 1) / 2;
  RGRID_LOOP(e, m1 * n1) {
    long i, j, c, d;
    dec1(e / n1, i, j);
    dec1(e % n1, c, d);
    const double* t = tau + (i * o + j) * v * v;
    const double x = t[c * v + d], y = t[d * v + c];
    tp[(e / n1) * l1 + e % n1] = 0.5 * (x + y) * (c == d ? 0.5 : 1.0);
    if (i < j && c < d) tm[pk0(i, j) * l0 + pk0(c, d)] = 0.5 * (x - y);
  }
}

// R2[i][j][a][b] += S[(ij)][(ab)] + sgn(i,j) sgn(a,b) A[(ij)][(ab)]
__global__ void rladder_unpack_kernel(double* R2, const double* Sl, const double* Al, long o, long v) {
  const long n1 = v * (v + 1) / 2, n0 = v * (v - 1) / 2;
  RGRID_LOOP(e, o * o * v * v) {
    const long b = e % v, a = (e / v) % v, j = (e / (v * v)) % o, i = e / (o * v * v);
    double x = Sl[pk1(i, j) * n1 + pk1(a, b)];
    if (i != j && a != b) x += ((i < j) == (a < b) ? 1.0 : -1.0) * Al[pk0(i, j) * n0 + pk0(a, b)];
    R2[e] += x;
  }
}

// Xp[(ka)][(cd)] = (kd|ac) + (kc|ad) over c <= d, Xm = (kd|ac) - (kc|ad) over c < d,
// from ovvv[k][d][a][c] = (kd|ac); rows of pitch l1 / l0
__global__ void rxpm_kernel(double* Xp, double* Xm, const double* ovvv, long o, long v, long l1, long l0) {
  const long n1 = v * (v + 1) / 2;
  RGRID_LOOP(e, o * v * n1) {
    const long ka = e / n1, k = ka / v, a = ka % v;
    long c, d;
    dec1(e % n1, c, d);
    const double* w = ovvv + k * v * v * v;
    const double x = w[(d * v + a) * v + c], y = w[(c * v + a) * v + d];
    Xp[ka * l1 + e % n1] = x + y;
    if (c < d) Xm[ka * l0 + pk0(c, d)] = x - y;
  }
}

// Xk[k][i][j][a] = XS[(ka)][(ij)] + sgn(i,j) XA[(ka)][(ij)]
__global__ void rxk_unpack_kernel(double* Xk, const double* XS, const double* XA, long o, long v) {
  const long m1 = o * (o + 1) / 2, m0 = o * (o - 1) / 2;
  RGRID_LOOP(e, o * o * o * v) {
    const long a = e % v, j = (e / v) % o, i = (e / (v * o)) % o, k = e / (v * o * o);
    const long ka = k * v + a;
    double x = XS[ka * m1 + pk1(i, j)];
    if (i != j) x += (i < j ? 1.0 : -1.0) * XA[ka * m0 + pk0(i, j)];
    Xk[e] = x;
  }
}

struct GpuExec {
  RCtx* c;
  int copy(const rplan::Copy& cp) {
    CopyShape sh;
    sh.nd = 0;
    long n = 1, smax = 0, dmax = 0;
    // dimensions of extent 1 dropped; the rest ordered by decreasing
    // destination stride, so consecutive threads write consecutive elements
    int ord[rplan::kMaxDims], m = 0;
    for (int t = 0; t < cp.dst.nd; ++t)
      if (cp.dst.ext[t] != 1) ord[m++] = t;
    std::sort(ord, ord + m, [&](int a, int b) { return std::labs(cp.dst.str[a]) > std::labs(cp.dst.str[b]); });
    for (int k = 0; k < m; ++k) {
      const int t = ord[k];
      sh.ext[k] = cp.dst.ext[t];
      sh.sstr[k] = cp.src.str[t];
      sh.dstr[k] = cp.dst.str[t];
      n *= sh.ext[k];
      smax += (sh.ext[k] - 1) * std::labs(sh.sstr[k]);
      dmax += (sh.ext[k] - 1) * std::labs(sh.dstr[k]);
    }
    sh.nd = m;
    for (int t = 0; t < cp.dst.nd; ++t)
      if (cp.dst.ext[t] == 0) return CCSD_OK;
    // a transpose in the two unit-stride dimensions: tiled through LDS
    int ds = -1, dd = -1;
    for (int k = 0; k < m; ++k) {
      if (sh.sstr[k] == 1 && ds < 0) ds = k;
      if (sh.dstr[k] == 1 && dd < 0) dd = k;
    }
    if (m >= 2 && ds >= 0 && dd >= 0 && ds != dd && sh.ext[ds] >= 8 && sh.ext[dd] >= 8) {
      CopyShape rest = sh;
      const long es = sh.ext[ds], ed = sh.ext[dd], sd = sh.sstr[dd], dsd = sh.dstr[ds];
      rest.ext[ds] = rest.ext[dd] = 1;
      const long nrest = n / (es * ed);
      dim3 grid((unsigned)((es + 31) / 32), (unsigned)((ed + 31) / 32), (unsigned)std::min<long>(nrest, 65535));
      hipLaunchKernelGGL(tiled_copy_kernel, grid, dim3(256), 0, c->stream, cp.dst.p, cp.src.p, rest, es, ed, sd, dsd,
                         nrest, cp.alpha, cp.beta);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) {
        c->last_error = hipGetErrorString(e);
        return CCSD_ERR_HIP;
      }
      return CCSD_OK;
    }
    if (n == 0) return CCSD_OK;
    const bool narrow = n < (1L << 31) && smax < (1L << 31) && dmax < (1L << 31);
    if (narrow)
      hipLaunchKernelGGL(strided_copy_kernel<unsigned>, dim3(rgrid(n)), dim3(kThreadsR), 0, c->stream, cp.dst.p,
                         cp.src.p, sh, n, cp.alpha, cp.beta);
    else
      hipLaunchKernelGGL(strided_copy_kernel<unsigned long>, dim3(rgrid(n)), dim3(kThreadsR), 0, c->stream, cp.dst.p,
                         cp.src.p, sh, n, cp.alpha, cp.beta);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      c->last_error = hipGetErrorString(e);
      return CCSD_ERR_HIP;
    }
    return CCSD_OK;
  }
  int ein(const char* spec, double alpha, const rplan::View& A, const rplan::View& B, double beta,
          const rplan::View& C) {
    rplan::Plan p;
    std::string err;
    if (!rplan::make_plan(spec, alpha, A, B, beta, C, c->sc, &p, &err)) {
      c->last_error = err;
      return CCSD_ERR_STATE;
    }
    for (const rplan::Copy& cp : p.pre) TRY(copy(cp));
    hipError_t e = gemm<double>(p.g, c->ws, c->stream);
    if (e != hipSuccess) {
      c->last_error = std::string(spec) + ": " + hipGetErrorString(e);
      return CCSD_ERR_HIP;
    }
    for (const rplan::Copy& cp : p.post) TRY(copy(cp));
    return CCSD_OK;
  }
  int axpy(const char* spec, double alpha, const rplan::View& A, double beta, const rplan::View& C) {
    rplan::Copy cp;
    std::string err;
    if (!rplan::make_axpy(spec, alpha, A, beta, C, &cp, &err)) {
      c->last_error = err;
      return CCSD_ERR_STATE;
    }
    return copy(cp);
  }
  int tau(const double* t1, const double* t2, double* tau, long o, long v) {
    hipLaunchKernelGGL(rtau_kernel, dim3(rgrid(o * o * v * v)), dim3(kThreadsR), 0, c->stream, tau, t1, t2, o, v);
    return hipGetLastError() == hipSuccess ? CCSD_OK : CCSD_ERR_HIP;
  }
  int vpm(const rplan::View& w, double* Vp, double* Vm, long v) {
    const View4 w4{w.p, w.str[0], w.str[1], w.str[2], w.str[3]};
    const long n1 = v * (v + 1) / 2;
    hipLaunchKernelGGL(rvpm_kernel, dim3(rgrid(n1 * n1)), dim3(kThreadsR), 0, c->stream, Vp, Vm, w4, v,
                       rplan::ldpair(n1), rplan::ldpair(rplan::npair0(v)));
    return hipGetLastError() == hipSuccess ? CCSD_OK : CCSD_ERR_HIP;
  }
  int ladder_pack(const double* tau, double* tp, double* tm, long o, long v) {
    const long n = o * (o + 1) / 2 * (v * (v + 1) / 2);
    hipLaunchKernelGGL(rladder_pack_kernel, dim3(rgrid(n)), dim3(kThreadsR), 0, c->stream, tp, tm, tau, o, v,
                       rplan::ldpair(rplan::npair1(v)), rplan::ldpair(rplan::npair0(v)));
    return hipGetLastError() == hipSuccess ? CCSD_OK : CCSD_ERR_HIP;
  }
  int xpm(const rplan::View& w, double* Xp, double* Xm, long o, long v) {
    const long n = o * v * rplan::npair1(v);
    hipLaunchKernelGGL(rxpm_kernel, dim3(rgrid(n)), dim3(kThreadsR), 0, c->stream, Xp, Xm, w.p, o, v,
                       rplan::ldpair(rplan::npair1(v)), rplan::ldpair(rplan::npair0(v)));
    return hipGetLastError() == hipSuccess ? CCSD_OK : CCSD_ERR_HIP;
  }
  int xk_unpack(const double* XS, const double* XA, double* Xk, long o, long v) {
    hipLaunchKernelGGL(rxk_unpack_kernel, dim3(rgrid(o * o * o * v)), dim3(kThreadsR), 0, c->stream, Xk, XS, XA, o,
                       v);
    return hipGetLastError() == hipSuccess ? CCSD_OK : CCSD_ERR_HIP;
  }
  int ladder_unpack(const double* Sl, const double* Al, double* R2, long o, long v) {
    hipLaunchKernelGGL(rladder_unpack_kernel, dim3(rgrid(o * o * v * v)), dim3(kThreadsR), 0, c->stream, R2, Sl, Al,
                       o, v);
    return hipGetLastError() == hipSuccess ? CCSD_OK : CCSD_ERR_HIP;
  }
};

// Device bytes of a context (statics, work, amplitude vectors, DIIS, scratch,
// workspace), and the transient n^4 buffers of the set-up.
size_t rplan_bytes(long n, long o, long v, int diis_m, bool with_C, size_t* transient) {
  const long t2 = o * o * v * v, ov3 = o * v * v * v, nvec = o * v + t2;
  const long n1 = v * (v + 1) / 2, n0 = v * (v - 1) / 2;
  long el = 10 * t2 + 2 * o * o * o * v + o * o * o * o + 5 * ov3 + (n1 + o * v) * rplan::ldpair(n1) +
            (n0 + o * v) * rplan::ldpair(n0);  // statics
  std::vector<long> wn;
  rplan::rccsd_work_sizes(o, v, &wn);
  for (long x : wn) el += x;
  el += 4 * nvec + 2L * diis_m * nvec + diis_m;  // amp, amp_new, err, R; DIIS
  el += 3 * (t2 + o * o * o * v + o * o * o * o);  // planner scratch
  el += o + v + std::max(1L << 16, 1024L * diis_m) + 4096;  // eo, ev, reduction partials and sums
  *transient = (size_t)(with_C ? 2 : 1) * (size_t)n * n * n * n * 8 + (with_C ? (size_t)n * n * 8 : 0);
  return (size_t)el * 8 + kGemmWorkspaceBytes;
}

int ralloc(RCtx* c, double** p, long n) {
  void* q = nullptr;
  HIPTRY(hipMalloc(&q, (size_t)std::max<long>(n, 1) * sizeof(double)));
  c->allocs.push_back(q);
  c->bytes += (size_t)std::max<long>(n, 1) * sizeof(double);
  *p = (double*)q;
  return CCSD_OK;
}

void rrelease(RCtx* c) {
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void* p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  if (c->ws.ptr) (void)hipFree(c->ws.ptr);
  c->ws.ptr = nullptr;
  for (auto& e : c->ev_)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  c->stream = nullptr;
}

// E = sum_ijab L[i][j][a][b] tau[i][j][a][b], L = 2 (ia|jb) - (ib|ja)
int renergy(RCtx* c, const double* amp, double* E) {
  const long o = c->o, v = c->v, t2 = o * o * v * v;
  hipLaunchKernelGGL(rtau_kernel, dim3(rgrid(t2)), dim3(kThreadsR), 0, c->stream, c->W.tau, amp, amp + o * v, o, v);
  const int nb = reduce_blocks_for(t2);
  k_dot_partials(c->partial, c->S.Lt, c->W.tau, t2, nb, c->stream);
  k_sum_partials(c->scal, c->partial, nb, 1, c->stream);
  double h = 0.0;
  HIPTRY(hipMemcpyAsync(&h, c->scal, 8, hipMemcpyDeviceToHost, c->stream));
  HIPTRY(hipStreamSynchronize(c->stream));
  *E = h;
  return CCSD_OK;
}

int rrhs(RCtx* c, const double* amp, bool time_it) {
  GpuExec x{c};
  const long o = c->o, v = c->v;
  if (time_it) HIPTRY(hipEventRecord(c->ev_[0], c->stream));
  TRY(rplan::rccsd_rhs(x, o, v, c->S, c->W, const_cast<double*>(amp), const_cast<double*>(amp) + o * v, c->R,
                       c->R + o * v));
  if (time_it) {
    HIPTRY(hipEventRecord(c->ev_[1], c->stream));
    HIPTRY(hipEventSynchronize(c->ev_[1]));
    float ms = 0.f;
    HIPTRY(hipEventElapsedTime(&ms, c->ev_[0], c->ev_[1]));
    c->timings_ms[0] = ms;
  }
  return CCSD_OK;
}

void rdiis_reset(RCtx* c) {
  c->diis_count = 0;
  c->diis_next = 0;
  std::fill(c->B.begin(), c->B.end(), 0.0);
}

// One Jacobi update (+ DIIS): c->amp holds the new amplitudes; *dt2 = max |dt2|.
int rupdate_once(RCtx* c, bool allow_diis, bool* fell, double* dt2) {
  hipStream_t s = c->stream;
  const long ov = c->o * c->v, t2 = c->nvec - ov;
  TRY(rrhs(c, c->amp, true));
  hipLaunchKernelGGL(rupdate_kernel, dim3(rgrid(c->nvec)), dim3(kThreadsR), 0, s, c->amp_new, c->err, c->R, c->amp,
                     c->eo, c->ev, c->o, c->v, c->nvec);
  const int nb = reduce_blocks_for(t2);
  k_maxabs_partials(c->partial, c->err + ov, t2, nb, s);
  std::vector<double> hp(nb);
  HIPTRY(hipMemcpyAsync(hp.data(), c->partial, nb * 8, hipMemcpyDeviceToHost, s));
  HIPTRY(hipStreamSynchronize(s));
  double mx = 0.0;
  for (double q : hp) mx = std::max(mx, q);
  *dt2 = mx;
  *fell = false;
  if (c->diis_m >= 2 && allow_diis) {
    const int m = c->diis_m, slot = c->diis_next;
    HIPTRY(hipMemcpyAsync(c->dvec + (long)slot * c->nvec, c->amp_new, c->nvec * 8, hipMemcpyDeviceToDevice, s));
    HIPTRY(hipMemcpyAsync(c->derr + (long)slot * c->nvec, c->err, c->nvec * 8, hipMemcpyDeviceToDevice, s));
    c->diis_count = std::min(c->diis_count + 1, m);
    c->diis_next = (slot + 1) % m;
    const int cnt = c->diis_count;
    const int nbd = reduce_blocks_for(c->nvec);
    k_multidot_partials(c->partial, c->derr, c->nvec, cnt, c->err, c->nvec, nbd, s);
    k_sum_partials(c->scal, c->partial, nbd, cnt, s);
    std::vector<double> row(cnt);
    HIPTRY(hipMemcpyAsync(row.data(), c->scal, cnt * 8, hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    for (int k = 0; k < cnt; ++k) {
      c->B[(size_t)slot * m + k] = row[k];
      c->B[(size_t)k * m + slot] = row[k];
    }
    if (cnt >= 2) {
      std::vector<double> Bc((size_t)cnt * cnt), coef;
      for (int i = 0; i < cnt; ++i)
        for (int j = 0; j < cnt; ++j) Bc[(size_t)i * cnt + j] = c->B[(size_t)i * m + j];
      if (diis_solve_bordered(Bc, cnt, coef)) {
        HIPTRY(hipMemcpyAsync(c->dcoef, coef.data(), cnt * 8, hipMemcpyHostToDevice, s));
        k_lincomb(c->amp, c->dvec, c->nvec, c->dcoef, cnt, c->nvec, s);
        HIPTRY(hipGetLastError());
        return CCSD_OK;
      }
      *fell = true;
    }
  }
  std::swap(c->amp, c->amp_new);
  return CCSD_OK;
}

int rmp2_start(RCtx* c) {
  hipLaunchKernelGGL(rmp2_kernel, dim3(rgrid(c->nvec)), dim3(kThreadsR), 0, c->stream, c->amp, c->S.ovovT, c->eo,
                     c->ev, c->o, c->v, c->nvec);
  HIPTRY(hipGetLastError());
  return renergy(c, c->amp, &c->e_mp2);
}

int rsetup(RCtx* c, const ccsd_rhf_problem* p) {
  const int n = p->n_basis, nocc = p->n_occ, nf = p->n_frozen;
  const long o = c->o, v = c->v, t2 = o * o * v * v, ov3 = o * v * v * v;
  size_t transient = 0;
  const size_t need = rplan_bytes(n, o, v, c->opt.diis_dim, p->C != nullptr, &transient);
  size_t free_b = 0, total_b = 0;
  HIPTRY(hipMemGetInfo(&free_b, &total_b));
  if ((c->opt.hbm_budget_bytes > 0 && need + transient > c->opt.hbm_budget_bytes) || need + transient > free_b) {
    c->last_error = "closed-shell CCSD needs more device memory than the budget allows";
    return CCSD_ERR_BUDGET;
  }
  HIPTRY(hipMalloc(&c->ws.ptr, kGemmWorkspaceBytes));
  c->ws.bytes = kGemmWorkspaceBytes;
  rplan::RStatic& S = c->S;
  double** st[] = {&S.ovov, &S.Lov, &S.Lt, &S.oovv, &S.ovovP, &S.M1, &S.ovovx, &S.ovovT, &S.oovvI, &S.ovovK};
  for (double** q : st) TRY(ralloc(c, q, t2));
  TRY(ralloc(c, &S.ovoo, o * v * o * o));
  TRY(ralloc(c, &S.Y2, o * o * o * v));
  TRY(ralloc(c, &S.oooo, o * o * o * o));
  double** s3[] = {&S.ovvv, &S.ovvvR, &S.ovvvR2, &S.ovvvY, &S.ovvvL};
  for (double** q : s3) TRY(ralloc(c, q, ov3));
  TRY(ralloc(c, &S.Vp, rplan::npair1(v) * rplan::ldpair(rplan::npair1(v))));
  TRY(ralloc(c, &S.Vm, rplan::npair0(v) * rplan::ldpair(rplan::npair0(v))));
  TRY(ralloc(c, &S.Xp, o * v * rplan::ldpair(rplan::npair1(v))));
  TRY(ralloc(c, &S.Xm, o * v * rplan::ldpair(rplan::npair0(v))));
  std::vector<long> wn;
  rplan::rccsd_work_sizes(o, v, &wn);
  rplan::RWork& W = c->W;
  double** wp[] = {&W.tau, &W.Tr, &W.Tx, &W.Z, &W.Wr1, &W.Wr2, &W.W12, &W.tA, &W.tB, &W.Y, &W.H, &W.Xk, &W.Woooo,
                   &W.Fov, &W.Foo, &W.Fvv, &W.Loo, &W.Lvv, &W.u, &W.tp, &W.tm, &W.Sl, &W.Al, &W.XS, &W.XA};
  for (size_t k = 0; k < wn.size(); ++k) TRY(ralloc(c, wp[k], wn[k]));
  c->nvec = o * v + t2;
  TRY(ralloc(c, &c->amp, c->nvec));
  TRY(ralloc(c, &c->amp_new, c->nvec));
  TRY(ralloc(c, &c->err, c->nvec));
  TRY(ralloc(c, &c->R, c->nvec));
  c->diis_m = c->opt.diis_dim;
  if (c->diis_m >= 2) {
    TRY(ralloc(c, &c->dvec, (long)c->diis_m * c->nvec));
    TRY(ralloc(c, &c->derr, (long)c->diis_m * c->nvec));
    TRY(ralloc(c, &c->dcoef, c->diis_m));
    c->B.assign((size_t)c->diis_m * c->diis_m, 0.0);
  }
  // planner scratch: any permuted operand is amplitude-, o^3 v- or o^4-sized
  const long scap = t2 + o * o * o * v + o * o * o * o;
  for (int k = 0; k < 3; ++k) {
    TRY(ralloc(c, &c->sc.p[k], scap));
    c->sc.cap[k] = scap;
  }
  TRY(ralloc(c, &c->eo, o));
  TRY(ralloc(c, &c->ev, v));
  // reduction partials: the DIIS Gram row writes one block partial per stored
  // vector (reduce_blocks_for gives at most 1024 blocks; diis_dim <= 64)
  TRY(ralloc(c, &c->partial, std::max(1L << 16, 1024L * c->diis_m)));
  TRY(ralloc(c, &c->scal, 4096));
  HIPTRY(hipMemcpyAsync(c->eo, c->h_eo.data(), o * 8, hipMemcpyHostToDevice, c->stream));
  HIPTRY(hipMemcpyAsync(c->ev, c->h_ev.data(), v * 8, hipMemcpyHostToDevice, c->stream));
  for (auto& e : c->ev_) HIPTRY(hipEventCreate(&e));
  // MO integrals in HBM for the set-up only
  const size_t n4 = (size_t)n * n * n * n;
  double *mo = nullptr, *tmp = nullptr, *dC = nullptr;
  int rc = CCSD_OK;
  if (hipMalloc(&mo, n4 * 8) != hipSuccess) rc = CCSD_ERR_BUDGET;
  if (rc == CCSD_OK && p->C &&
      (hipMalloc(&tmp, n4 * 8) != hipSuccess || hipMalloc(&dC, (size_t)n * n * 8) != hipSuccess))
    rc = CCSD_ERR_BUDGET;
  if (rc == CCSD_OK && hipMemcpyAsync(mo, p->eri, n4 * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess)
    rc = CCSD_ERR_HIP;
  if (rc == CCSD_OK && p->C) {
    if (hipMemcpyAsync(dC, p->C, (size_t)n * n * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess) rc = CCSD_ERR_HIP;
    if (rc == CCSD_OK) rc = ao_to_mo_device(mo, tmp, dC, n, c->ws, c->stream);
  }
  if (rc == CCSD_OK) {
    GpuExec x{c};
    rc = rplan::rccsd_build_static(x, mo, n, nocc, nf, S);
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(mo);
  if (tmp) (void)hipFree(tmp);
  if (dC) (void)hipFree(dC);
  if (rc != CCSD_OK) return rc;
  HIPTRY(hipGetLastError());
  return rmp2_start(c);
}

}  // namespace
}  // namespace ccsd

using namespace ccsd;

extern "C" {

size_t ccsd_rhf_plan_bytes(int n_basis, int n_occ, int n_frozen) {
  if (n_basis < 2 || n_frozen < 0 || n_occ - n_frozen < 1 || n_basis - n_occ < 1) return 0;
  size_t tr = 0;
  return rplan_bytes(n_basis, n_occ - n_frozen, n_basis - n_occ, 8, true, &tr) + tr;
}

int ccsd_rhf_create(const ccsd_rhf_problem* p, const ccsd_options* o, ccsd_rhf_ctx** out) {
  if (!p || !out || !p->eps || !p->eri) return CCSD_ERR_ARG;
  *out = nullptr;
  const int n = p->n_basis, nocc = p->n_occ, nf = p->n_frozen;
  if (n < 2 || n > 1024 || nf < 0 || nocc - nf < 1 || n - nocc < 1) return CCSD_ERR_ARG;
  const size_t n4 = (size_t)n * n * n * n;
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(p->eps[i])) return CCSD_ERR_NONFINITE;
  if (p->C)
    for (size_t i = 0; i < (size_t)n * n; ++i)
      if (!std::isfinite(p->C[i])) return CCSD_ERR_NONFINITE;
  for (size_t i = 0; i < n4; ++i)
    if (!std::isfinite(p->eri[i])) return CCSD_ERR_NONFINITE;
  ccsd_options opt;
  ccsd_default_options(&opt);
  if (o) opt = *o;
  if (opt.precision != CCSD_PREC_FP64) return CCSD_ERR_ARG;  // the closed-shell path is FP64
  if (opt.diis_dim < 0 || opt.diis_dim > 64) return CCSD_ERR_ARG;
  const long oc = nocc - nf, vi = n - nocc;
  std::vector<double> eo(p->eps + nf, p->eps + nocc), ev(p->eps + nocc, p->eps + n);
  for (double a : eo)
    for (double b : ev)
      if (std::fabs(a - b) < 1e-10) return CCSD_ERR_DENOMINATOR;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CCSD_ERR_NO_DEVICE;
  if (opt.device < 0 || opt.device >= ndev) return CCSD_ERR_ARG;
  ccsd_rhf_ctx* h = new (std::nothrow) ccsd_rhf_ctx();
  if (!h) return CCSD_ERR_ARG;
  RCtx* c = &h->c;
  c->opt = opt;
  c->device = opt.device;
  c->o = oc;
  c->v = vi;
  c->h_eo = eo;
  c->h_ev = ev;
  int rc = hipSetDevice(c->device) == hipSuccess ? CCSD_OK : CCSD_ERR_HIP;
  if (rc == CCSD_OK && hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = CCSD_ERR_HIP;
  if (rc == CCSD_OK) rc = rsetup(c, p);
  if (rc != CCSD_OK) {
    if (!c->last_error.empty() && rc != CCSD_ERR_BUDGET) fprintf(stderr, "ccsd_rhf_create: %s\n", c->last_error.c_str());
    rrelease(c);
    delete h;
    return rc;
  }
  *out = h;
  return CCSD_OK;
}

void ccsd_rhf_destroy(ccsd_rhf_ctx* h) {
  if (!h) return;
  rrelease(&h->c);
  delete h;
}

int ccsd_rhf_solve(ccsd_rhf_ctx* h, double tol, int max_iter, ccsd_result* out) {
  if (!h || !(tol >= 0.0) || max_iter < 1) return CCSD_ERR_ARG;
  RCtx* c = &h->c;
  HIPTRY(hipSetDevice(c->device));
  ccsd_result r;
  std::memset(&r, 0, sizeof(r));
  double E;
  TRY(renergy(c, c->amp, &E));
  if (!std::isfinite(E)) return CCSD_ERR_NONFINITE;
  r.e_mp2 = c->e_mp2;
  rdiis_reset(c);
  bool converged = false;
  int rc = CCSD_OK, it = 0;
  double dE = 0.0, dt2 = 0.0;
  auto t0 = std::chrono::steady_clock::now();
  while (it < max_iter) {
    const double E_old = E;
    auto ti = std::chrono::steady_clock::now();
    bool fell = false;
    rc = rupdate_once(c, it + 1 >= c->opt.diis_start, &fell, &dt2);
    if (rc != CCSD_OK) break;
    if (fell) r.diis_fallbacks++;
    r.fp64_iterations++;
    rc = renergy(c, c->amp, &E);
    if (rc != CCSD_OK) break;
    ++it;
    dE = std::fabs(E - E_old);
    if (c->opt.verbose) {
      const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - ti).count();
      printf("Energi: %.12f Steg: %d Time/iter: %.6f\n", E, it + 1, sec);
      fflush(stdout);
    }
    if (!std::isfinite(E)) {
      rc = CCSD_ERR_NONFINITE;
      break;
    }
    if (dE < tol) {
      converged = true;
      break;
    }
  }
  auto t1 = std::chrono::steady_clock::now();
  r.e_corr = E;
  r.iterations = it;
  r.converged = converged ? 1 : 0;
  r.last_delta_e = dE;
  r.max_delta_t2 = dt2;
  r.seconds_total = std::chrono::duration<double>(t1 - t0).count();
  r.seconds_per_iter = it ? r.seconds_total / it : 0.0;
  if (out) *out = r;
  if (rc != CCSD_OK) return rc;
  return converged ? CCSD_OK : CCSD_ERR_NOT_CONVERGED;
}

int ccsd_rhf_iterate(ccsd_rhf_ctx* h, int n, double* energies) {
  if (!h || n < 0) return CCSD_ERR_ARG;
  RCtx* c = &h->c;
  HIPTRY(hipSetDevice(c->device));
  for (int k = 0; k < n; ++k) {
    bool fell = false;
    double dt2 = 0.0, E = 0.0;
    TRY(rupdate_once(c, true, &fell, &dt2));
    TRY(renergy(c, c->amp, &E));
    if (energies) energies[k] = E;
  }
  return CCSD_OK;
}

int ccsd_rhf_reset(ccsd_rhf_ctx* h) {
  if (!h) return CCSD_ERR_ARG;
  RCtx* c = &h->c;
  HIPTRY(hipSetDevice(c->device));
  rdiis_reset(c);
  return rmp2_start(c);
}

int ccsd_rhf_residual(ccsd_rhf_ctx* h, const double* t1, const double* t2, double* r1, double* r2) {
  if (!h || !t1 || !t2 || !r1 || !r2) return CCSD_ERR_ARG;
  RCtx* c = &h->c;
  HIPTRY(hipSetDevice(c->device));
  const long ov = c->o * c->v, t2n = c->nvec - ov;
  double* in = c->amp_new;  // scratch vector (the next update rewrites it)
  HIPTRY(hipMemcpyAsync(in, t1, ov * 8, hipMemcpyHostToDevice, c->stream));
  HIPTRY(hipMemcpyAsync(in + ov, t2, t2n * 8, hipMemcpyHostToDevice, c->stream));
  TRY(rrhs(c, in, false));
  HIPTRY(hipMemcpyAsync(r1, c->R, ov * 8, hipMemcpyDeviceToHost, c->stream));
  HIPTRY(hipMemcpyAsync(r2, c->R + ov, t2n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPTRY(hipStreamSynchronize(c->stream));
  return CCSD_OK;
}

int ccsd_rhf_energy(ccsd_rhf_ctx* h, const double* t1, const double* t2, double* e) {
  if (!h || !t1 || !t2 || !e) return CCSD_ERR_ARG;
  RCtx* c = &h->c;
  HIPTRY(hipSetDevice(c->device));
  const long ov = c->o * c->v, t2n = c->nvec - ov;
  double* in = c->amp_new;
  HIPTRY(hipMemcpyAsync(in, t1, ov * 8, hipMemcpyHostToDevice, c->stream));
  HIPTRY(hipMemcpyAsync(in + ov, t2, t2n * 8, hipMemcpyHostToDevice, c->stream));
  return renergy(c, in, e);
}

int ccsd_rhf_get_amplitudes(ccsd_rhf_ctx* h, double* t1, double* t2) {
  if (!h || !t1 || !t2) return CCSD_ERR_ARG;
  RCtx* c = &h->c;
  HIPTRY(hipSetDevice(c->device));
  const long ov = c->o * c->v;
  HIPTRY(hipMemcpyAsync(t1, c->amp, ov * 8, hipMemcpyDeviceToHost, c->stream));
  HIPTRY(hipMemcpyAsync(t2, c->amp + ov, (c->nvec - ov) * 8, hipMemcpyDeviceToHost, c->stream));
  HIPTRY(hipStreamSynchronize(c->stream));
  return CCSD_OK;
}

int ccsd_rhf_set_amplitudes(ccsd_rhf_ctx* h, const double* t1, const double* t2) {
  if (!h || !t1 || !t2) return CCSD_ERR_ARG;
  RCtx* c = &h->c;
  HIPTRY(hipSetDevice(c->device));
  const long ov = c->o * c->v;
  HIPTRY(hipMemcpyAsync(c->amp, t1, ov * 8, hipMemcpyHostToDevice, c->stream));
  HIPTRY(hipMemcpyAsync(c->amp + ov, t2, (c->nvec - ov) * 8, hipMemcpyHostToDevice, c->stream));
  HIPTRY(hipStreamSynchronize(c->stream));
  rdiis_reset(c);
  return CCSD_OK;
}

int ccsd_rhf_last_timings(ccsd_rhf_ctx* h, double* ms, int cap) {
  if (!h || !ms || cap < 1) return CCSD_ERR_ARG;
  RCtx* c = &h->c;
  ms[0] = c->timings_ms[0];
  return CCSD_OK;
}

const char* ccsd_rhf_last_error(ccsd_rhf_ctx* h) { return h ? h->c.last_error.c_str() : ""; }

}  // extern "C"
