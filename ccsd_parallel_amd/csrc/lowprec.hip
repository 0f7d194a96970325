// Mixed-precision GEMM path: FP32 MFMA (v_mfma_f32_16x16x4f32 through the
// generic engine) and BF16 hi/lo split (v_mfma_f32_16x16x32_bf16, three
// products hi*hi + hi*lo + lo*hi accumulated in FP32).
//
// x = hi + lo with hi = bf16(x), lo = bf16(x - hi) keeps ~16 significant bits
// of each operand; the dropped lo*lo term and FP32 accumulation put the
// product error near 2^-16 relative, which the final FP64 sweeps remove.
#include <algorithm>
#include <cstring>

#include "bf16.hpp"
#include "lowprec.hpp"

namespace ccsd {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kThreads = 256;
inline long pad64(long k) { return ((k + 63) / 64) * 64; }
inline int grid_for(long n) {
  long b = (n + kThreads - 1) / kThreads;
  return (int)std::max<long>(1, std::min<long>(b, 1L << 20));
}
#define GRID_LOOP(idx, n) \
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < (n); idx += (long)gridDim.x * blockDim.x)

__global__ void conv_f32_kernel(float* dst, const double* src, long rows, long K, long sld, long ldk, double scale) {
  const long n = rows * ldk;
  GRID_LOOP(idx, n) {
    long r = idx / ldk, k = idx % ldk;
    dst[idx] = (k < K) ? (float)(scale * src[r * sld + k]) : 0.0f;
  }
}

__global__ void conv_bf16x2_kernel(uint16_t* hi, uint16_t* lo, const double* src, long rows, long K, long sld,
                                   long ldk, double scale) {
  const long n = rows * ldk;
  GRID_LOOP(idx, n) {
    long r = idx / ldk, k = idx % ldk;
    uint16_t h = 0, l = 0;
    if (k < K) bf16_split(scale * src[r * sld + k], h, l);
    hi[idx] = h;
    lo[idx] = l;
  }
}

// y[m] = alpha * sum_k A[m][k] x[k] + beta * y[m], A given as FP32 (A32) or as a
// BF16 hi/lo pair (row stride ld, a multiple of 64); one wavefront per row,
// 8 consecutive k per lane (16-byte loads) and a scalar tail.
__global__ void gemv_lp_kernel(double* y, const float* A32, const uint16_t* Ah, const uint16_t* Al, long ld,
                               const double* __restrict__ x, long M, long K, double alpha, double beta) {
  const int lane = threadIdx.x & 63;
  const long row = blockIdx.x * (long)(blockDim.x / 64) + (threadIdx.x >> 6);
  if (row >= M) return;
  const long K8 = K & ~7L;
  double s = 0.0;
  if (A32) {
    const float* a = A32 + row * ld;
    for (long k0 = lane * 8; k0 < K8; k0 += 512) {
      const float4 u = *(const float4*)(a + k0), v = *(const float4*)(a + k0 + 4);
      s += (double)u.x * x[k0] + (double)u.y * x[k0 + 1] + (double)u.z * x[k0 + 2] + (double)u.w * x[k0 + 3];
      s += (double)v.x * x[k0 + 4] + (double)v.y * x[k0 + 5] + (double)v.z * x[k0 + 6] + (double)v.w * x[k0 + 7];
    }
    for (long k = K8 + lane; k < K; k += 64) s += (double)a[k] * x[k];
  } else {
    const uint16_t* h = Ah + row * ld;
    const uint16_t* l = Al + row * ld;
    for (long k0 = lane * 8; k0 < K8; k0 += 512) {
      const uint4 hv = *(const uint4*)(h + k0), lv = *(const uint4*)(l + k0);
      const uint32_t hw[4] = {hv.x, hv.y, hv.z, hv.w}, lw[4] = {lv.x, lv.y, lv.z, lv.w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double a0 = (double)bf16_to_f((uint16_t)(hw[q] & 0xffff)) + (double)bf16_to_f((uint16_t)(lw[q] & 0xffff));
        const double a1 = (double)bf16_to_f((uint16_t)(hw[q] >> 16)) + (double)bf16_to_f((uint16_t)(lw[q] >> 16));
        s += a0 * x[k0 + 2 * q] + a1 * x[k0 + 2 * q + 1];
      }
    }
    for (long k = K8 + lane; k < K; k += 64) s += ((double)bf16_to_f(h[k]) + (double)bf16_to_f(l[k])) * x[k];
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) y[row] = alpha * s + (beta != 0.0 ? beta * y[row] : 0.0);
}

// Output of a BF16 hi/lo product: FP32 C[row*ldc + col] = acc, or (C64 set)
// the FP64 C64[row*ldc + col] = alpha * acc + beta * C64, optionally also
// split into the hi/lo pair Oh/Ol[row*ldo + col] for use as a later operand.
struct Bf16Out {
  float* C = nullptr;
  double* C64 = nullptr;
  long ldc = 0;
  double alpha = 1.0, beta = 0.0;
  uint16_t *Oh = nullptr, *Ol = nullptr;
  long ldo = 0;
};

struct Bf16Args {
  const uint16_t *Ahi, *Alo, *Bhi, *Blo;
  long lda, ldb;
  int M, N, Kp;  // Kp: K rounded up to 32 (operands zero padded)
  int tiles_m, tiles_n, tiles_pad;
  Bf16Out out;
};

__device__ __forceinline__ void bf16_store(const Bf16Out& o, int row, int col, float acc) {
  if (o.C64) {
    double* c = o.C64 + (long)row * o.ldc + col;
    double v = o.alpha * (double)acc;
    if (o.beta != 0.0) v += o.beta * (*c);
    *c = v;
    if (o.Oh) {
      uint16_t h, l;
      bf16_split(v, h, l);
      o.Oh[(long)row * o.ldo + col] = h;
      o.Ol[(long)row * o.ldo + col] = l;
    }
  } else {
    o.C[(long)row * o.ldc + col] = acc;
  }
}

// C[m][n] = sum_k A[m][k] B[n][k] with A = Ahi + Alo, B = Bhi + Blo, dropping lo*lo.
template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void bf16x3_nt_kernel(Bf16Args p) {
  constexpr int NT = 64 * WM * WN;
  constexpr int BK = 32;
  constexpr int LDK = BK + 8;  // 80-byte row pitch: 16 lanes of one k-slice hit distinct banks
  constexpr int TM = BM / WM / 16, TN = BN / WN / 16;
  constexpr int CA = 2 * BM * 4 / NT;  // 16-byte chunks per thread (hi+lo)
  constexpr int CB = 2 * BN * 4 / NT;
  static_assert(CA * NT == 2 * BM * 4 && CB * NT == 2 * BN * 4, "tile/threads mismatch");
  __shared__ __attribute__((aligned(16))) uint16_t sA[2][BM * LDK];
  __shared__ __attribute__((aligned(16))) uint16_t sB[2][BN * LDK];

  const int gx = p.tiles_pad;  // XCD-aware order when a multiple of 8 (see bf16x3_nt)
  const int tile = (gx & 7) ? (int)blockIdx.x : (int)((blockIdx.x & 7) * (gx >> 3) + (blockIdx.x >> 3));
  if (tile >= p.tiles_m * p.tiles_n) return;
  const int tm = tile % p.tiles_m, tn = tile / p.tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / WN) * (BM / WM), wn0 = (wave % WN) * (BN / WN);

  uint4 ra[CA], rb[CB];
  auto load = [&](int k0) {
#pragma unroll
    for (int it = 0; it < CA; ++it) {
      int u = it * NT + tid;
      int hl = u / (BM * 4), rem = u % (BM * 4), m = rem >> 2, ch = rem & 3;
      const uint16_t* src = (hl ? p.Alo : p.Ahi) + (long)(m0 + m) * p.lda + k0 + ch * 8;
      ra[it] = (m0 + m < p.M) ? *(const uint4*)src : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int it = 0; it < CB; ++it) {
      int u = it * NT + tid;
      int hl = u / (BN * 4), rem = u % (BN * 4), n = rem >> 2, ch = rem & 3;
      const uint16_t* src = (hl ? p.Blo : p.Bhi) + (long)(n0 + n) * p.ldb + k0 + ch * 8;
      rb[it] = (n0 + n < p.N) ? *(const uint4*)src : make_uint4(0, 0, 0, 0);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int it = 0; it < CA; ++it) {
      int u = it * NT + tid;
      int hl = u / (BM * 4), rem = u % (BM * 4), m = rem >> 2, ch = rem & 3;
      *(uint4*)&sA[hl][m * LDK + ch * 8] = ra[it];
    }
#pragma unroll
    for (int it = 0; it < CB; ++it) {
      int u = it * NT + tid;
      int hl = u / (BN * 4), rem = u % (BN * 4), n = rem >> 2, ch = rem & 3;
      *(uint4*)&sB[hl][n * LDK + ch * 8] = rb[it];
    }
  };

  f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0, 0, 0, 0};

  const int nk = p.Kp / BK;
  if (nk > 0) load(0);
  for (int kt = 0; kt < nk; ++kt) {
    __syncthreads();  // previous compute finished reading LDS
    store();
    __syncthreads();
    if (kt + 1 < nk) load((kt + 1) * BK);
    const int kofs = 8 * (lane >> 4);
    bf16x8_t ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int r = (wm0 + i * 16 + (lane & 15)) * LDK + kofs;
      ah[i] = *(const bf16x8_t*)&sA[0][r];
      al[i] = *(const bf16x8_t*)&sA[1][r];
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int r = (wn0 + j * 16 + (lane & 15)) * LDK + kofs;
      bh[j] = *(const bf16x8_t*)&sB[0][r];
      bl[j] = *(const bf16x8_t*)&sB[1][r];
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
      }
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int row = m0 + wm0 + i * 16 + 4 * (lane >> 4) + r;
        int col = n0 + wn0 + j * 16 + (lane & 15);
        if (row < p.M && col < p.N) bf16_store(p.out, row, col, acc[i][j][r]);
      }
}

// Large-tile variant: 8 waves, each a 64 x 128 sub-tile (4 x 8 MFMA tiles, 128
// accumulator registers), so one 256 x 256 (or 128 x 512) block tile reaches
// ~190 flop per byte loaded -- the 128 x 128 tile above needs ~25 TB/s of L2
// bandwidth to feed 2.5 PF/s.  Per k-step the A fragments of the wave are kept
// in registers and the B fragments are streamed one 16-column tile at a time.
// DBUF: double-buffered LDS (one barrier per k-tile; 160 KB at 256 x 256).
template <int BM, int BN, int WM, int WN, bool DBUF>
__global__ __launch_bounds__(64 * WM * WN) void bf16x3_nt_big_kernel(Bf16Args p) {
  constexpr int NT = 64 * WM * WN;
  constexpr int BK = 32;
  constexpr int LDK = BK + 8;
  constexpr int TM = BM / WM / 16, TN = BN / WN / 16;
  constexpr int CA = 2 * BM * 4 / NT;
  constexpr int CB = 2 * BN * 4 / NT;
  static_assert(CA * NT == 2 * BM * 4 && CB * NT == 2 * BN * 4, "tile/threads mismatch");
  constexpr int NBUF = DBUF ? 2 : 1;
  __shared__ __attribute__((aligned(16))) uint16_t sA[NBUF][2][BM * LDK];
  __shared__ __attribute__((aligned(16))) uint16_t sB[NBUF][2][BN * LDK];

  const int gx = p.tiles_pad;  // XCD-aware order when a multiple of 8 (see bf16x3_nt)
  const int tile = (gx & 7) ? (int)blockIdx.x : (int)((blockIdx.x & 7) * (gx >> 3) + (blockIdx.x >> 3));
  if (tile >= p.tiles_m * p.tiles_n) return;
  const int tm = tile % p.tiles_m, tn = tile / p.tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / WN) * (BM / WM), wn0 = (wave % WN) * (BN / WN);

  uint4 ra[CA], rb[CB];
  auto load = [&](int k0) {
#pragma unroll
    for (int it = 0; it < CA; ++it) {
      int u = it * NT + tid;
      int hl = u / (BM * 4), rem = u % (BM * 4), m = rem >> 2, ch = rem & 3;
      const uint16_t* src = (hl ? p.Alo : p.Ahi) + (long)(m0 + m) * p.lda + k0 + ch * 8;
      ra[it] = (m0 + m < p.M) ? *(const uint4*)src : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int it = 0; it < CB; ++it) {
      int u = it * NT + tid;
      int hl = u / (BN * 4), rem = u % (BN * 4), n = rem >> 2, ch = rem & 3;
      const uint16_t* src = (hl ? p.Blo : p.Bhi) + (long)(n0 + n) * p.ldb + k0 + ch * 8;
      rb[it] = (n0 + n < p.N) ? *(const uint4*)src : make_uint4(0, 0, 0, 0);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int it = 0; it < CA; ++it) {
      int u = it * NT + tid;
      int hl = u / (BM * 4), rem = u % (BM * 4), m = rem >> 2, ch = rem & 3;
      *(uint4*)&sA[buf][hl][m * LDK + ch * 8] = ra[it];
    }
#pragma unroll
    for (int it = 0; it < CB; ++it) {
      int u = it * NT + tid;
      int hl = u / (BN * 4), rem = u % (BN * 4), n = rem >> 2, ch = rem & 3;
      *(uint4*)&sB[buf][hl][n * LDK + ch * 8] = rb[it];
    }
  };

  f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0, 0, 0, 0};

  auto compute = [&](int buf) {
    const int kofs = 8 * (lane >> 4);
    bf16x8_t ah[TM], al[TM];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int r = (wm0 + i * 16 + (lane & 15)) * LDK + kofs;
      ah[i] = *(const bf16x8_t*)&sA[buf][0][r];
      al[i] = *(const bf16x8_t*)&sA[buf][1][r];
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int r = (wn0 + j * 16 + (lane & 15)) * LDK + kofs;
      const bf16x8_t bh = *(const bf16x8_t*)&sB[buf][0][r];
      const bf16x8_t bl = *(const bf16x8_t*)&sB[buf][1][r];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh, acc[i][j], 0, 0, 0);
      }
    }
  };

  const int nk = p.Kp / BK;
  if (nk > 0) {
    load(0);
    store(0);
    __syncthreads();
  }
  for (int kt = 0; kt < nk; ++kt) {
    const bool more = kt + 1 < nk;
    if (more) load((kt + 1) * BK);
    if (DBUF) {
      compute(kt & 1);
      if (more) store((kt + 1) & 1);
      __syncthreads();
    } else {
      compute(0);
      __syncthreads();
      if (more) {
        store(0);
        __syncthreads();
      }
    }
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int row = m0 + wm0 + i * 16 + 4 * (lane >> 4) + r;
        int col = n0 + wn0 + j * 16 + (lane & 15);
        if (row < p.M && col < p.N) bf16_store(p.out, row, col, acc[i][j][r]);
      }
}

template <int BM, int BN, int WM, int WN, bool DBUF>
hipError_t launch_big(Bf16Args p, hipStream_t s) {
  p.tiles_m = (p.M + BM - 1) / BM;
  p.tiles_n = (p.N + BN - 1) / BN;
  long tiles = (long)p.tiles_m * p.tiles_n;
  p.tiles_pad = tiles >= 32 ? (int)(((tiles + 7) / 8) * 8) : (int)tiles;
  hipLaunchKernelGGL((bf16x3_nt_big_kernel<BM, BN, WM, WN, DBUF>), dim3(p.tiles_pad), dim3(64 * WM * WN), 0, s, p);
  return hipGetLastError();
}

// Tile choice by modelled time: padding x wave quantisation over 256 CUs,
// divided by the tile's relative throughput (measured on MI355X).
hipError_t bf16x3_nt(const uint16_t* Ahi, const uint16_t* Alo, long lda, const uint16_t* Bhi, const uint16_t* Blo,
                     long ldb, const Bf16Out& out, int M, int N, int K, hipStream_t s, int force = -1) {
  Bf16Args p;
  p.Ahi = Ahi; p.Alo = Alo; p.Bhi = Bhi; p.Blo = Blo; p.out = out;
  p.lda = lda; p.ldb = ldb;
  p.M = M; p.N = N; p.Kp = ((K + 31) / 32) * 32;
  struct Opt { int BM, BN; double rate; };
  const Opt opts[3] = {{128, 128, 1.0}, {256, 256, 1.65}, {128, 512, 1.5}};
  int best = 0;
  double bt = 1e300;
  for (int c = 0; c < 3; ++c) {
    long tm = (M + opts[c].BM - 1) / opts[c].BM, tn = (N + opts[c].BN - 1) / opts[c].BN;
    double pad = double(tm * opts[c].BM) * double(tn * opts[c].BN) / (double(M) * N);
    double tiles = double(tm * tn);
    double rounds = std::ceil(tiles / 256.0);
    double t = pad * (rounds * 256.0 / tiles) / opts[c].rate;
    if (t < bt) { bt = t; best = c; }
  }
  if (best == 1) best = 3;  // 256 x 256: the double-buffered variant measured ~1% faster
  if (force >= 0 && force < 4) best = force;
  switch (best) {
    case 1: return launch_big<256, 256, 4, 2, false>(p, s);
    case 2: return launch_big<128, 512, 2, 4, false>(p, s);
    case 3: return launch_big<256, 256, 4, 2, true>(p, s);
    default: break;
  }
  constexpr int BM = 128, BN = 128;
  p.tiles_m = (M + BM - 1) / BM;
  p.tiles_n = (N + BN - 1) / BN;
  long tiles = (long)p.tiles_m * p.tiles_n;
  p.tiles_pad = tiles >= 32 ? (int)(((tiles + 7) / 8) * 8) : (int)tiles;
  hipLaunchKernelGGL((bf16x3_nt_kernel<BM, BN, 2, 2>), dim3(p.tiles_pad), dim3(256), 0, s, p);
  return hipGetLastError();
}

}  // namespace

size_t lowprec_plan_bytes(const Dims& d, int precision) {
  if (precision == CCSD_PREC_FP64) return 0;
  const long lpv = pad64(d.pv), lov = pad64(d.ov), O = d.O, V = d.V;
  // dominant-contraction operands: vvvv, ovvv, Doovv, tau, Tring, WT (fp32: 4 B; bf16 hi+lo: 2 x 2 B)
  long elems = d.pv * lpv + d.ov * lpv + d.ov * lov + d.po * lpv + 2 * d.ov * lov;
  size_t bytes = (size_t)elems * 4;
  if (precision == CCSD_PREC_BF16X3) bytes += (size_t)(d.ov * lpv + d.po * lpv) * 4;  // FP32 ovvv, tau as well
  // secondary FP32 copies: oovv, oovv_vu, oovv_ou, G, ovvv_u; t1, T2u, tauu, Tv, W1, W3, W3u, W2, Fae, Fmi
  long sec = d.po * d.pv + d.po * V * V + O * O * d.pv + d.ov * d.ov + d.ov * V * V;
  sec += d.ov + 2 * O * O * d.pv + d.po * V * V + d.po * d.po + 2 * O * d.po * V + O * O * O * V + d.po * d.ov +
         V * V + O * O;
  return bytes + (size_t)sec * 4;
}

template <typename T>
static int lp_alloc(Ctx* c, T** p, long n) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, (size_t)std::max<long>(n, 1) * sizeof(T));
  if (e != hipSuccess) {
    c->last_error = hipGetErrorString(e);
    return CCSD_ERR_HIP;
  }
  c->lp->allocs.push_back(q);
  c->bytes += (size_t)std::max<long>(n, 1) * sizeof(T);
  *p = (T*)q;
  return CCSD_OK;
}

#define TRY(x)                      \
  do {                              \
    int _rc = (x);                  \
    if (_rc != CCSD_OK) return _rc; \
  } while (0)
#define HIPTRY(x)                              \
  do {                                         \
    hipError_t _e = (x);                       \
    if (_e != hipSuccess) {                    \
      c->last_error = hipGetErrorString(_e);   \
      return CCSD_ERR_HIP;                     \
    }                                          \
  } while (0)

static void to_f32(Ctx* c, float* dst, const double* src, long n) {
  hipLaunchKernelGGL(conv_f32_kernel, dim3(grid_for(n)), dim3(kThreads), 0, c->stream, dst, src, 1L, n, n, n, 1.0);
}

int lowprec_allocate(Ctx* c, int precision) {
  if (precision == CCSD_PREC_FP64) return CCSD_OK;
  if (precision != CCSD_PREC_FP32 && precision != CCSD_PREC_BF16X3) return CCSD_ERR_ARG;
  if (!c->lp) c->lp = new LowPrec();
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  const long O = d.O, V = d.V;
  L->ld_pv = pad64(d.pv);
  L->ld_ov = pad64(d.ov);
  hipStream_t s = c->stream;
  if (!L->have_sec) {
    // FP32 copies shared by both modes: ovvv and tau (k-contiguous, padded), the
    // secondary integrals and the per-iteration operands
    TRY(lp_alloc(c, &L->ovvv32, d.ov * L->ld_pv));
    TRY(lp_alloc(c, &L->tau32, d.po * L->ld_pv));
    TRY(lp_alloc(c, &L->oovv_f, d.po * d.pv));
    TRY(lp_alloc(c, &L->oovv_vu_f, d.po * V * V));
    TRY(lp_alloc(c, &L->oovv_ou_f, O * O * d.pv));
    TRY(lp_alloc(c, &L->G_f, d.ov * d.ov));
    TRY(lp_alloc(c, &L->ovvv_u_f, d.ov * V * V));
    TRY(lp_alloc(c, &L->t1_f, d.ov));
    TRY(lp_alloc(c, &L->T2u_f, O * O * d.pv));
    TRY(lp_alloc(c, &L->tauu_f, O * O * d.pv));
    TRY(lp_alloc(c, &L->Tv_f, d.po * V * V));
    TRY(lp_alloc(c, &L->W1_f, d.po * d.po));
    TRY(lp_alloc(c, &L->W3_f, O * d.po * V));
    TRY(lp_alloc(c, &L->W3u_f, O * O * O * V));
    TRY(lp_alloc(c, &L->W2_f, d.po * d.ov));
    TRY(lp_alloc(c, &L->Fae_f, V * V));
    TRY(lp_alloc(c, &L->Fmi_f, O * O));
    hipLaunchKernelGGL(conv_f32_kernel, dim3(grid_for(d.ov * L->ld_pv)), dim3(kThreads), 0, s, L->ovvv32, c->ovvv,
                       d.ov, d.pv, d.pv, L->ld_pv, 1.0);
    to_f32(c, L->oovv_f, c->oovv, d.po * d.pv);
    to_f32(c, L->oovv_vu_f, c->oovv_vu, d.po * V * V);
    to_f32(c, L->oovv_ou_f, c->oovv_ou, O * O * d.pv);
    to_f32(c, L->G_f, c->G, d.ov * d.ov);
    to_f32(c, L->ovvv_u_f, c->ovvv_u, d.ov * V * V);
    L->have_sec = 1;
  }
  if (precision == CCSD_PREC_FP32 && !L->have_fp32) {
    TRY(lp_alloc(c, &L->vvvv32, d.pv * L->ld_pv));
    TRY(lp_alloc(c, &L->Doovv32, d.ov * L->ld_ov));
    TRY(lp_alloc(c, &L->Tring32, d.ov * L->ld_ov));
    TRY(lp_alloc(c, &L->WT32, d.ov * L->ld_ov));
    // the unpack writes Tring32 for columns < ov only: zero the padding once
    HIPTRY(hipMemsetAsync(L->Tring32, 0, (size_t)d.ov * L->ld_ov * 4, s));
    hipLaunchKernelGGL(conv_f32_kernel, dim3(grid_for(d.pv * L->ld_pv)), dim3(kThreads), 0, s, L->vvvv32, c->vvvv,
                       d.pv, d.pv, d.pv, L->ld_pv, 1.0);
    hipLaunchKernelGGL(conv_f32_kernel, dim3(grid_for(d.ov * L->ld_ov)), dim3(kThreads), 0, s, L->Doovv32, c->Doovv,
                       d.ov, d.ov, d.ov, L->ld_ov, 1.0);
    L->have_fp32 = 1;
  }
  if (precision == CCSD_PREC_BF16X3 && !L->have_bf16) {
    TRY(lp_alloc(c, &L->vvvv_h, d.pv * L->ld_pv));
    TRY(lp_alloc(c, &L->vvvv_l, d.pv * L->ld_pv));
    TRY(lp_alloc(c, &L->ovvv_h, d.ov * L->ld_pv));
    TRY(lp_alloc(c, &L->ovvv_l, d.ov * L->ld_pv));
    TRY(lp_alloc(c, &L->Doovv_h, d.ov * L->ld_ov));
    TRY(lp_alloc(c, &L->Doovv_l, d.ov * L->ld_ov));
    TRY(lp_alloc(c, &L->tau_h, d.po * L->ld_pv));
    TRY(lp_alloc(c, &L->tau_l, d.po * L->ld_pv));
    TRY(lp_alloc(c, &L->Tring_h, d.ov * L->ld_ov));
    TRY(lp_alloc(c, &L->Tring_l, d.ov * L->ld_ov));
    TRY(lp_alloc(c, &L->WT_h, d.ov * L->ld_ov));
    TRY(lp_alloc(c, &L->WT_l, d.ov * L->ld_ov));
    // WT_h/WT_l (W_mbej epilogue) and Tring_h/Tring_l (unpack) are written for
    // columns < ov only: zero the padding once
    for (uint16_t* q : {L->WT_h, L->WT_l, L->Tring_h, L->Tring_l})
      HIPTRY(hipMemsetAsync(q, 0, (size_t)d.ov * L->ld_ov * 2, s));
    hipLaunchKernelGGL(conv_bf16x2_kernel, dim3(grid_for(d.pv * L->ld_pv)), dim3(kThreads), 0, s, L->vvvv_h,
                       L->vvvv_l, c->vvvv, d.pv, d.pv, d.pv, L->ld_pv, 1.0);
    hipLaunchKernelGGL(conv_bf16x2_kernel, dim3(grid_for(d.ov * L->ld_pv)), dim3(kThreads), 0, s, L->ovvv_h,
                       L->ovvv_l, c->ovvv, d.ov, d.pv, d.pv, L->ld_pv, 1.0);
    hipLaunchKernelGGL(conv_bf16x2_kernel, dim3(grid_for(d.ov * L->ld_ov)), dim3(kThreads), 0, s, L->Doovv_h,
                       L->Doovv_l, c->Doovv, d.ov, d.ov, d.ov, L->ld_ov, 1.0);
    L->have_bf16 = 1;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    c->last_error = hipGetErrorString(e);
    return CCSD_ERR_HIP;
  }
  return CCSD_OK;
}

void lowprec_free(Ctx* c) {
  if (!c->lp) return;
  for (void* p : c->lp->allocs) (void)hipFree(p);
  delete c->lp;
  c->lp = nullptr;
}

static void conv(Ctx* c, int prec, const double* src, long rows, long K, long ld, float* d32, uint16_t* dh,
                 uint16_t* dl) {
  long n = rows * ld;
  if (prec == CCSD_PREC_FP32)
    hipLaunchKernelGGL(conv_f32_kernel, dim3(grid_for(n)), dim3(kThreads), 0, c->stream, d32, src, rows, K, K, ld,
                       1.0);
  else
    hipLaunchKernelGGL(conv_bf16x2_kernel, dim3(grid_for(n)), dim3(kThreads), 0, c->stream, dh, dl, src, rows, K,
                       K, ld, 1.0);
}

int lowprec_amplitudes(Ctx* c, int prec) {
  TRY(lowprec_allocate(c, prec));
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  hipLaunchKernelGGL(conv_f32_kernel, dim3(grid_for(d.po * L->ld_pv)), dim3(kThreads), 0, c->stream, L->tau32, c->tau,
                     d.po, d.pv, d.pv, L->ld_pv, 1.0);
  if (prec == CCSD_PREC_BF16X3) conv(c, prec, c->tau, d.po, d.pv, L->ld_pv, nullptr, L->tau_h, L->tau_l);
  return CCSD_OK;
}

RingOut lowprec_ring_out(Ctx* c, int prec) {
  LowPrec* L = c->lp;
  RingOut r;
  r.ld = L->ld_ov;
  if (prec == CCSD_PREC_FP32) {
    r.f = L->Tring32;
  } else {
    r.h = L->Tring_h;
    r.l = L->Tring_l;
  }
  return r;
}

static int gemv_lp(Ctx* c, int prec, const float* A32, const uint16_t* Ah, const uint16_t* Al, const double* x,
                   double* y, double alpha, double beta) {
  const Dims& d = c->d;
  const int rows_per_block = 4;
  const bool f32 = prec == CCSD_PREC_FP32;
  hipLaunchKernelGGL(gemv_lp_kernel, dim3((unsigned)((d.ov + rows_per_block - 1) / rows_per_block)),
                     dim3(64 * rows_per_block), 0, c->stream, y, f32 ? A32 : nullptr, f32 ? nullptr : Ah,
                     f32 ? nullptr : Al, c->lp->ld_ov, x, d.ov, d.ov, alpha, beta);
  return CCSD_OK;
}

int lowprec_fme(Ctx* c, int prec, const double* t1) {
  LowPrec* L = c->lp;
  return gemv_lp(c, prec, L->Doovv32, L->Doovv_h, L->Doovv_l, t1, c->Fme, 1.0, 1.0);
}

int lowprec_r1_ring(Ctx* c, int prec) {
  LowPrec* L = c->lp;
  return gemv_lp(c, prec, L->Tring32, L->Tring_h, L->Tring_l, c->Fme, c->R1, 1.0, 1.0);
}

// C64 (M x N, ldc) = alpha * A B^T + beta * C64, operands as stored low-precision
// pairs; in BF16 mode the product can also be emitted as a hi/lo pair (Oh/Ol, ldo).
static int nt_product(Ctx* c, int prec, const float* A32, const uint16_t* Ah, const uint16_t* Al, long lda,
                      const float* B32, const uint16_t* Bh, const uint16_t* Bl, long ldb, int M, int N, int K,
                      double* C, long ldc, double alpha, double beta, uint16_t* Oh = nullptr,
                      uint16_t* Ol = nullptr, long ldo = 0) {
  hipError_t e;
  if (prec == CCSD_PREC_FP32) {
    GemmDesc<float> g;
    g.A = A32; g.B = B32; g.Cd = C;
    g.M = M; g.N = N; g.K = K;
    g.sAm = lda; g.sAk = 1; g.sBk = 1; g.sBn = ldb; g.sCm = ldc; g.sCn = 1;
    g.alpha = (float)alpha; g.beta = (float)beta;  // 1, 0.5, 0: exact in FP32
    e = gemm<float>(g, c->ws, c->stream);
  } else {
    Bf16Out o;
    o.C64 = C; o.ldc = ldc; o.alpha = alpha; o.beta = beta;
    o.Oh = Oh; o.Ol = Ol; o.ldo = ldo;
    e = bf16x3_nt(Ah, Al, lda, Bh, Bl, ldb, o, M, N, K, c->stream);
  }
  if (e != hipSuccess) {
    c->last_error = hipGetErrorString(e);
    return CCSD_ERR_HIP;
  }
  return CCSD_OK;
}

int lowprec_w2(Ctx* c, int prec) {
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  return nt_product(c, prec, L->tau32, L->tau_h, L->tau_l, L->ld_pv, L->ovvv32, L->ovvv_h, L->ovvv_l, L->ld_pv,
                    (int)d.po, (int)d.ov, (int)d.pv, c->W2, d.ov, 1.0, 1.0);
}

int lowprec_wmbej(Ctx* c, int prec) {
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  // Doovv is symmetric, so its rows serve as the k-contiguous B operand.  In BF16
  // mode the finished W_mbej is also written as the ring GEMM's hi/lo operand.
  return nt_product(c, prec, L->Tring32, L->Tring_h, L->Tring_l, L->ld_ov, L->Doovv32, L->Doovv_h, L->Doovv_l,
                    L->ld_ov, (int)d.ov, (int)d.ov, (int)d.ov, c->WT, d.ov, 0.5, 1.0, L->WT_h, L->WT_l, L->ld_ov);
}

int lowprec_ladder(Ctx* c, int prec) {
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  // <ef||ab> = <ab||ef>: the packed vvvv block is symmetric, rows are k-contiguous B.
  return nt_product(c, prec, L->tau32, L->tau_h, L->tau_l, L->ld_pv, L->vvvv32, L->vvvv_h, L->vvvv_l, L->ld_pv,
                    (int)d.po, (int)d.pv, (int)d.pv, c->Rp, d.pv, 1.0, 1.0);
}

int lowprec_ring(Ctx* c, int prec) {
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  // BF16: WT_h/WT_l come from the W_mbej epilogue (WT is final there)
  if (prec == CCSD_PREC_FP32) conv(c, prec, c->WT, d.ov, d.ov, L->ld_ov, L->WT32, nullptr, nullptr);
  return nt_product(c, prec, L->Tring32, L->Tring_h, L->Tring_l, L->ld_ov, L->WT32, L->WT_h, L->WT_l, L->ld_ov,
                    (int)d.ov, (int)d.ov, (int)d.ov, c->X, d.ov, 1.0, 0.0);
}

GemmDesc<float> f32_desc(const GemmDesc<double>& g, const float* A32, const float* B32) {
  GemmDesc<float> f;
  f.A = A32; f.B = B32; f.C = nullptr; f.Cd = g.C;
  f.M = g.M; f.N = g.N; f.K = g.K;
  f.sAm = g.sAm; f.sAk = g.sAk; f.sBk = g.sBk; f.sBn = g.sBn; f.sCm = g.sCm; f.sCn = g.sCn;
  f.nb1 = g.nb1; f.nb2 = g.nb2;
  f.sA1 = g.sA1; f.sA2 = g.sA2; f.sB1 = g.sB1; f.sB2 = g.sB2; f.sC1 = g.sC1; f.sC2 = g.sC2;
  f.alpha = (float)g.alpha; f.beta = (float)g.beta;
  f.reduce_batch = g.reduce_batch;
  return f;
}

int lowprec_gemm_nt(const double* A, const double* B, double* C, int M, int N, int K, double alpha, double beta,
                    int prec, hipStream_t s) {
  const int tile = (prec >> 8) & 0xff;  // 0: automatic, else forced tile configuration + 1
  prec &= 0xff;
  const long ld = pad64(K);
  float *A32 = nullptr, *B32 = nullptr;
  uint16_t *Ah = nullptr, *Al = nullptr, *Bh = nullptr, *Bl = nullptr;
  std::vector<void*> owned;
  auto al = [&](void** p, size_t b) {
    hipError_t e = hipMalloc(p, std::max<size_t>(b, 16));
    if (e == hipSuccess) owned.push_back(*p);
    return e;
  };
  hipError_t e = hipSuccess;
  if (prec == CCSD_PREC_FP32) {
    e = al((void**)&A32, (size_t)M * ld * 4);
    if (e == hipSuccess) e = al((void**)&B32, (size_t)N * ld * 4);
  } else {
    e = al((void**)&Ah, (size_t)M * ld * 2);
    if (e == hipSuccess) e = al((void**)&Al, (size_t)M * ld * 2);
    if (e == hipSuccess) e = al((void**)&Bh, (size_t)N * ld * 2);
    if (e == hipSuccess) e = al((void**)&Bl, (size_t)N * ld * 2);
  }
  void* wsp = nullptr;
  if (e == hipSuccess) e = al(&wsp, 16u << 20);
  if (e == hipSuccess) {
    if (prec == CCSD_PREC_FP32) {
      hipLaunchKernelGGL(conv_f32_kernel, dim3(grid_for(M * ld)), dim3(kThreads), 0, s, A32, A, (long)M, (long)K,
                         (long)K, ld, 1.0);
      hipLaunchKernelGGL(conv_f32_kernel, dim3(grid_for(N * ld)), dim3(kThreads), 0, s, B32, B, (long)N, (long)K,
                         (long)K, ld, 1.0);
      GemmDesc<float> g;
      g.A = A32; g.B = B32; g.Cd = C;
      g.M = M; g.N = N; g.K = K;
      g.sAm = ld; g.sAk = 1; g.sBk = 1; g.sBn = ld; g.sCm = N; g.sCn = 1;
      g.alpha = (float)alpha; g.beta = (float)beta;
      GemmWorkspace ws{wsp, 16u << 20};
      e = gemm<float>(g, ws, s, nullptr, tile - 1, -1);
    } else {
      hipLaunchKernelGGL(conv_bf16x2_kernel, dim3(grid_for(M * ld)), dim3(kThreads), 0, s, Ah, Al, A, (long)M,
                         (long)K, (long)K, ld, 1.0);
      hipLaunchKernelGGL(conv_bf16x2_kernel, dim3(grid_for(N * ld)), dim3(kThreads), 0, s, Bh, Bl, B, (long)N,
                         (long)K, (long)K, ld, 1.0);
      Bf16Out o;
      o.C64 = C; o.ldc = N; o.alpha = alpha; o.beta = beta;
      e = bf16x3_nt(Ah, Al, ld, Bh, Bl, ld, o, M, N, K, s, tile - 1);
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  for (void* p : owned) (void)hipFree(p);
  return e == hipSuccess ? CCSD_OK : CCSD_ERR_HIP;
}

}  // namespace ccsd
