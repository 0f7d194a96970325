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
#include "strassen.hpp"

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

// BF16 hi/lo operands are stored interleaved per 8-element chunk: row r of a
// matrix with leading dimension ld (multiple of 64) holds, for chunk c = k/8,
// the 8 hi values followed by the 8 lo values at r*2*ld + 16*c.  One k-tile of
// 32 elements of a row is then one contiguous 128-byte line.
__device__ __forceinline__ long il_index(long row, long k, long ld) { return row * 2 * ld + (k >> 3) * 16 + (k & 7); }

// dst (interleaved, rows x ld) = split(scale * src[r*sld + k]) for k < K, 0 beyond;
// one thread per 8-element chunk (two 16-byte stores).
// Whole chunks of a 16-byte aligned source with a row stride that keeps the
// alignment are read with 16-byte loads.
template <typename S>
__global__ void conv_bf16il_kernel(uint16_t* dst, const S* src, long rows, long K, long sld, long ld,
                                   double scale) {
  const long chunks = ld / 8, n = rows * chunks;
  constexpr int kPer16 = 16 / (int)sizeof(S);
  const bool vec = (reinterpret_cast<uintptr_t>(src) % 16 == 0) && (sld % kPer16 == 0);
  GRID_LOOP(idx, n) {
    const long r = idx / chunks, c = idx % chunks;
    uint16_t h[8], l[8];
    if (vec && c * 8 + 8 <= K) {
      S v[8];
      const uint4* p = reinterpret_cast<const uint4*>(src + r * sld + c * 8);
#pragma unroll
      for (int q = 0; q < 8 / kPer16; ++q) *reinterpret_cast<uint4*>(v + q * kPer16) = p[q];
#pragma unroll
      for (int e = 0; e < 8; ++e) bf16_split(scale * (double)v[e], h[e], l[e]);
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const long k = c * 8 + e;
        h[e] = 0;
        l[e] = 0;
        if (k < K) bf16_split(scale * (double)src[r * sld + k], h[e], l[e]);
      }
    }
    uint4 hv, lv;
    hv.x = h[0] | ((uint32_t)h[1] << 16); hv.y = h[2] | ((uint32_t)h[3] << 16);
    hv.z = h[4] | ((uint32_t)h[5] << 16); hv.w = h[6] | ((uint32_t)h[7] << 16);
    lv.x = l[0] | ((uint32_t)l[1] << 16); lv.y = l[2] | ((uint32_t)l[3] << 16);
    lv.z = l[4] | ((uint32_t)l[5] << 16); lv.w = l[6] | ((uint32_t)l[7] << 16);
    uint4* o = (uint4*)(dst + r * 2 * ld + c * 16);
    o[0] = hv;
    o[1] = lv;
  }
}

// dst (interleaved, rows x ld) = split(src[k*sld + r]): the transpose of a
// row-major source; threads run along r so the source reads coalesce.
__global__ void conv_bf16il_t_kernel(uint16_t* dst, const double* src, long rows, long K, long sld, long ld) {
  const long chunks = ld / 8, n = rows * chunks;
  GRID_LOOP(idx, n) {
    const long c = idx / rows, r = idx % rows;
    uint16_t h[8], l[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const long k = c * 8 + e;
      h[e] = 0;
      l[e] = 0;
      if (k < K) bf16_split(src[k * sld + r], h[e], l[e]);
    }
    uint4 hv, lv;
    hv.x = h[0] | ((uint32_t)h[1] << 16); hv.y = h[2] | ((uint32_t)h[3] << 16);
    hv.z = h[4] | ((uint32_t)h[5] << 16); hv.w = h[6] | ((uint32_t)h[7] << 16);
    lv.x = l[0] | ((uint32_t)l[1] << 16); lv.y = l[2] | ((uint32_t)l[3] << 16);
    lv.z = l[4] | ((uint32_t)l[5] << 16); lv.w = l[6] | ((uint32_t)l[7] << 16);
    uint4* o = (uint4*)(dst + r * 2 * ld + c * 16);
    o[0] = hv;
    o[1] = lv;
  }
}

// y[m] = alpha * sum_k A[m][k] x[k] + beta * y[m], A given as FP32 (A32, row
// stride ld) or as interleaved BF16 hi/lo rows (Ab); one wavefront per row,
// 8 consecutive k per lane (16-byte loads) and a scalar tail.
__global__ void gemv_lp_kernel(double* y, const float* A32, const uint16_t* Ab, long ld,
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
    const uint16_t* a = Ab + row * 2 * ld;
    for (long k0 = lane * 8; k0 < K8; k0 += 512) {
      const uint4 hv = *(const uint4*)(a + 2 * k0), lv = *(const uint4*)(a + 2 * k0 + 8);
      const uint32_t hw[4] = {hv.x, hv.y, hv.z, hv.w}, lw[4] = {lv.x, lv.y, lv.z, lv.w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double a0 = (double)bf16_to_f((uint16_t)(hw[q] & 0xffff)) + (double)bf16_to_f((uint16_t)(lw[q] & 0xffff));
        const double a1 = (double)bf16_to_f((uint16_t)(hw[q] >> 16)) + (double)bf16_to_f((uint16_t)(lw[q] >> 16));
        s += a0 * x[k0 + 2 * q] + a1 * x[k0 + 2 * q + 1];
      }
    }
    for (long k = K8 + lane; k < K; k += 64) {
      const long i = il_index(0, k, ld);
      s += ((double)bf16_to_f(a[i]) + (double)bf16_to_f(a[i + 8])) * x[k];
    }
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) y[row] = alpha * s + (beta != 0.0 ? beta * y[row] : 0.0);
}

// Output of a BF16 hi/lo product: FP32 C[row*ldc + col] = acc, or (C64 set)
// the FP64 C64[row*ldc + col] = alpha * acc + beta * C64, optionally also
// written as interleaved hi/lo rows Ob (leading dimension ldo) for use as a
// later operand.
// With C32 set instead, the FP32 C32 is the accumulator: v = alpha * acc +
// beta * C32, stored back to C32 unless c32_keep is false (when only the hi/lo
// operand Ob is wanted).
struct Bf16Out {
  float* C = nullptr;
  double* C64 = nullptr;
  float* C32 = nullptr;
  bool c32_keep = true;
  long ldc = 0;
  double alpha = 1.0, beta = 0.0;
  uint16_t* Ob = nullptr;
  long ldo = 0;
};

struct Bf16Args {
  const uint16_t *A, *B;  // interleaved hi/lo rows, leading dimensions lda / ldb (elements)
  long lda, ldb;
  int M, N, Kp;           // Kp: K rounded up to 32 (operands zero padded)
  int tiles_m, tiles_n, tiles_pad, group;
  Bf16Out out;
  // Split tail (tail_split > 0): workgroup x runs on XCD x % 8; each XCD first
  // takes xcd_full whole tiles of the range [x * xcd_full, (x + 1) * xcd_full),
  // then units of the tail tiles [8 * xcd_full, tiles), each cut into
  // tail_split chunks of tail_kt k-tiles, dealt round-robin over the XCDs.  A
  // tail unit writes its BM x BN FP32 partial to tail_partial[unit]; a second
  // kernel sums them in a fixed order.
  int tail_split = 0, xcd_full = 0, tail_units = 0, tail_kt = 0;
  float* tail_partial = nullptr;
};

// Tile index -> (tile row, tile column): column-major, or walked in groups of
// `group` tile rows.  A bijection over all tiles; the split-tail reduction
// uses the same map.
__device__ __forceinline__ void bf16_tile_coords(const Bf16Args& p, int tile, int& tm, int& tn) {
  tm = tile % p.tiles_m;
  tn = tile / p.tiles_m;
  if (p.group > 0) {
    const int per = p.group * p.tiles_n, g = tile / per, first = g * p.group;
    const int gs = min(p.tiles_m - first, p.group), r = tile % per;
    tm = first + r % gs;
    tn = r / gs;
  }
}

__device__ __forceinline__ void bf16_store(const Bf16Out& o, int row, int col, float acc) {
  if (o.C64 || o.C32) {
    const long off = (long)row * o.ldc + col;
    double v = o.alpha * (double)acc;
    if (o.C64) {
      if (o.beta != 0.0) v += o.beta * o.C64[off];
      o.C64[off] = v;
    } else {
      if (o.beta != 0.0) v += o.beta * (double)o.C32[off];
      if (o.c32_keep) o.C32[off] = (float)v;
    }
    if (o.Ob) {
      uint16_t h, l;
      bf16_split(v, h, l);
      const long i = il_index(row, col, o.ldo);
      o.Ob[i] = h;
      o.Ob[i + 8] = l;
    }
  } else {
    o.C[(long)row * o.ldc + col] = acc;
  }
}

// A wave-uniform pointer held in scalar registers: loads then take the
// scalar-base form (64-bit base + 32-bit per-lane offset), and the compiler
// cannot fold the per-lane offset into hoisted 64-bit address registers.
__device__ __forceinline__ const char* uniform_ptr(const char* p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (const char*)(((uint64_t)hi << 32) | lo);
}

// C[m][n] = sum_k A[m][k] B[n][k] with A = Ahi + Alo, B = Bhi + Blo, dropping
// lo*lo: three v_mfma_f32_16x16x32_bf16 per 16x16x32 step, FP32 accumulation.
//
// Block tile BM x BN over WM x WN waves (wave tile (BM/WM) x (BN/WN)), k-tiles
// of 32, LDS double-buffered in two arrays per operand (256 x 256: 128 KB, one
// block per CU; 128 x 128 on 4 waves: 64 KB, two blocks per CU).
//
// Operands go from global memory straight into LDS (global_load_lds_dwordx4):
// lane l of a wave writes the 16 bytes at the wave's LDS base + 16 l, so
// piece u = it*NT + tid lands at byte 16 u of the stage, i.e. tile row u/8,
// slot u%8, and the lane loads the global chunk that belongs in that slot.  A
// row keeps its 128 bytes (hi0 lo0 hi1 lo1 .. lo3 of one 32-deep k-tile) with
// chunk q at slot q ^ (row & 7) ^ ((row >> 3) & 1): 16 lanes store two whole
// rows and the 16 rows of one MFMA operand fetch read without bank conflicts
// (within each half of the 16 rows the slots are distinct, and rows of equal
// parity in the two halves take slots of opposite parity).  Since every pass
// moves NT/8 rows (a multiple of 16), a thread's chunk index is the same in
// every pass.  Each load is a scalar base (tile row start + k offset) plus a
// 32-bit per-lane offset that holds the row, clamped to the last row (rows
// past M / N only feed accumulator rows that are never stored); no registers
// stage the data and no LDS store instructions run.
//
// Per k-tile: wait for this wave's loads of the tile, barrier (every wave's
// loads have landed and every wave is done with the previous tile, whose
// buffer is refilled next), issue the next tile's loads, then the MFMAs.  The
// three products of a 16 x 16 output block are issued as three passes over the
// wave's row blocks, so dependent MFMAs are TM instructions apart.  Measured
// at 12000^3 on MI355X against staging the loads through registers (the
// previous form of this kernel): 7.97 -> 7.26 ms.
//
// Tile order: XCD-aware (linear workgroup x runs on XCD x % 8, so XCD j takes a
// contiguous range of tiles); with group > 0 the tiles of that range are
// walked in groups of `group` tile rows so the 32 concurrent tiles of an XCD
// share operand panels in its L2.
template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN, (WM * WN == 4) ? 2 : 1) void bf16x3_nt_kernel(Bf16Args p) {
  constexpr int NT = 64 * WM * WN;
  constexpr int BK = 32;
  constexpr int RP = 2 * BK;  // row pitch in bf16 elements: 128 bytes, hi and lo interleaved per 8
  constexpr int TM = BM / WM / 16, TN = BN / WN / 16;
  // 16-byte pieces per thread per k-tile; a partial last pass drops whole waves
  constexpr int CA = (BM * 8 + NT - 1) / NT, CB = (BN * 8 + NT - 1) / NT;
  static_assert((NT / 8) % 16 == 0, "a pass must cover whole 16-row swizzle periods");
  static_assert((BM * 8) % 64 == 0 && (BN * 8) % 64 == 0, "passes end on wave boundaries");
  __shared__ __attribute__((aligned(1024))) uint16_t sA0[BM * RP];
  __shared__ __attribute__((aligned(1024))) uint16_t sA1[BM * RP];
  __shared__ __attribute__((aligned(1024))) uint16_t sB0[BN * RP];
  __shared__ __attribute__((aligned(1024))) uint16_t sB1[BN * RP];

  int tile, kt0 = 0, kt1 = p.Kp / 32;  // k-tile range of this unit
  float* tail_out = nullptr;           // this unit's FP32 partial (split tail tiles)
  if (p.tail_split > 0) {
    const int x = (int)(blockIdx.x & 7), r = (int)(blockIdx.x >> 3);
    if (r < p.xcd_full) {
      tile = x * p.xcd_full + r;
    } else {
      const int v = (r - p.xcd_full) * 8 + x;
      if (v >= p.tail_units) return;
      tile = 8 * p.xcd_full + v / p.tail_split;
      kt0 = (v % p.tail_split) * p.tail_kt;
      kt1 = min(kt1, kt0 + p.tail_kt);
      tail_out = p.tail_partial + (long)v * (BM * BN);
    }
  } else {
    const int gx = p.tiles_pad;
    tile = (gx & 7) ? (int)blockIdx.x : (int)((blockIdx.x & 7) * (gx >> 3) + (blockIdx.x >> 3));
  }
  if (tile >= p.tiles_m * p.tiles_n) return;
  int tm, tn;
  bf16_tile_coords(p, tile, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / WN) * (BM / WM), wn0 = (wave % WN) * (BN / WN);
  const int wv = __builtin_amdgcn_readfirstlane(wave);

  // per-lane byte offsets from the tile's first row: row (it*NT/8 + tid/8), clamped, chunk q
  const int mr = tid >> 3, sl = tid & 7;
  const int q = sl ^ (mr & 7) ^ ((mr >> 3) & 1);
  uint32_t offa[CA], offb[CB];
#pragma unroll
  for (int it = 0; it < CA; ++it)
    offa[it] = (uint32_t)(((long)(min(m0 + it * (NT / 8) + mr, p.M - 1) - m0) * 2 * p.lda + q * 8) * 2);
#pragma unroll
  for (int it = 0; it < CB; ++it)
    offb[it] = (uint32_t)(((long)(min(n0 + it * (NT / 8) + mr, p.N - 1) - n0) * 2 * p.ldb + q * 8) * 2);
  const uint16_t* rowa = p.A + (long)m0 * 2 * p.lda;
  const uint16_t* rowb = p.B + (long)n0 * 2 * p.ldb;

  f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0, 0, 0, 0};

  const int nk = kt1 - kt0;
  auto issue = [&](int t, uint16_t* la, uint16_t* lb) {
    const char* ba = uniform_ptr((const char*)(rowa + 2L * (kt0 + t) * BK));
    const char* bb = uniform_ptr((const char*)(rowb + 2L * (kt0 + t) * BK));
#pragma unroll
    for (int it = 0; it < CA; ++it)
      if ((it + 1) * NT <= BM * 8 || it * NT + wv * 64 < BM * 8)
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(ba + offa[it]),
                                         (__attribute__((address_space(3))) void*)(la + (it * NT + wv * 64) * 8), 16,
                                         0, 0);
#pragma unroll
    for (int it = 0; it < CB; ++it)
      if ((it + 1) * NT <= BN * 8 || it * NT + wv * 64 < BN * 8)
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(bb + offb[it]),
                                         (__attribute__((address_space(3))) void*)(lb + (it * NT + wv * 64) * 8), 16,
                                         0, 0);
  };
  // operand fetch: lane l reads tile row 16i + (l & 15) (a multiple of 16 plus
  // l & 15, so its swizzle is (l & 7) ^ ((l >> 3) & 1)), k-slice l >> 4
  const int swz = (lane & 7) ^ ((lane >> 3) & 1), ks = lane >> 4;
  const int oh = ((2 * ks) ^ swz) << 3, ol = ((2 * ks + 1) ^ swz) << 3;
  struct AFrag { bf16x8_t h[TM], l[TM]; };
  struct BFrag { bf16x8_t h, l; };
  auto read_a = [&](const uint16_t* a, AFrag& f) {
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int r = (wm0 + i * 16 + (lane & 15)) * RP;
      f.l[i] = *(const bf16x8_t*)&a[r + ol];
      f.h[i] = *(const bf16x8_t*)&a[r + oh];
    }
  };
  auto read_b = [&](const uint16_t* b, int j, BFrag& f) {
    const int r = (wn0 + j * 16 + (lane & 15)) * RP;
    f.h = *(const bf16x8_t*)&b[r + oh];
    f.l = *(const bf16x8_t*)&b[r + ol];
  };
  // column block j: the three products of each 16 x 16 block as three passes
  // over the wave's row blocks (dependent MFMAs TM instructions apart)
  auto mma_j = [&](const AFrag& fa, const BFrag& fb, int j) {
#pragma unroll
    for (int i = 0; i < TM; ++i) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa.l[i], fb.h, acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < TM; ++i) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa.h[i], fb.l, acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < TM; ++i) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa.h[i], fb.h, acc[i][j], 0, 0, 0);
  };
  // vmcnt(0): every load this wave issued (the next tile's) has written LDS
  constexpr int kVmcnt0 = 0x3f70;
  AFrag fa0, fa1;
  BFrag fb[2];
  // One k-tile t held in LDS (a, b), its A fragments in fa and its first B
  // fragment in fb[P]: column block j reads block j+1's B fragment before its
  // MFMAs.  Before the last block: wait for this wave's loads of tile t+1,
  // barrier (every wave's loads of t+1 have landed and every wave has read its
  // last fragment of t), issue the loads of tile t+2 into (a, b), and read
  // tile t+1's A fragments into fn and its first B fragment while the last
  // block's MFMAs run.  P is the parity of fb at block 0 (TN may be odd).
  auto ktile = [&](int t, uint16_t* a, uint16_t* b, const uint16_t* na, const uint16_t* nb, const AFrag& fa,
                  AFrag& fn, auto P) {
    constexpr int p0 = decltype(P)::value;
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      if (j + 1 < TN) {
        read_b(b, j + 1, fb[(p0 + j + 1) & 1]);
      } else {
        // also after the last tile (uniform; the reads of the idle buffer are
        // in bounds and unused): no branch around the reads, so the MFMAs
        // below wait only for this block's B fragment
        __builtin_amdgcn_s_waitcnt(kVmcnt0);
        __syncthreads();
        if (t + 2 < nk) issue(t + 2, a, b);
        read_a(na, fn);
        read_b(nb, 0, fb[(p0 + TN) & 1]);
      }
      __builtin_amdgcn_sched_barrier(0);  // the fragment reads (and loads) go out before this block's MFMAs
      mma_j(fa, fb[(p0 + j) & 1], j);
    }
  };
  issue(0, sA0, sB0);
  if (nk > 1) issue(1, sA1, sB1);
  __builtin_amdgcn_s_waitcnt(kVmcnt0);
  __syncthreads();
  read_a(sA0, fa0);
  read_b(sB0, 0, fb[0]);
  using P0 = std::integral_constant<int, 0>;
  using P1 = std::integral_constant<int, TN & 1>;
  for (int kt = 0; kt < nk; kt += 2) {
    ktile(kt, sA0, sB0, sA1, sB1, fa0, fa1, P0{});
    if (kt + 1 < nk) ktile(kt + 1, sA1, sB1, sA0, sB0, fa1, fa0, P1{});
  }
  if (tail_out) {
    // whole BM x BN tile in tile coordinates; the reduction drops rows and
    // columns past M and N
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          tail_out[(wm0 + i * 16 + 4 * (lane >> 4) + r) * BN + wn0 + j * 16 + (lane & 15)] = acc[i][j][r];
    return;
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm0 + i * 16 + 4 * (lane >> 4) + r;
        const int col = n0 + wn0 + j * 16 + (lane & 15);
        if (row < p.M && col < p.N) bf16_store(p.out, row, col, acc[i][j][r]);
      }
}

// Split tail tiles: out(tile) = sum_s partial[tile][s], the chunks summed in
// order (FP32, as the kernel accumulates), through the same output path.
template <int BM, int BN>
__global__ void bf16_tail_reduce_kernel(Bf16Args p, int ntail) {
  constexpr long per = (long)BM * BN;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < ntail * per;
       idx += (long)gridDim.x * blockDim.x) {
    const long tt = idx / per, e = idx % per;
    int tm, tn;
    bf16_tile_coords(p, 8 * p.xcd_full + (int)tt, tm, tn);
    const int row = tm * BM + (int)(e / BN), col = tn * BN + (int)(e % BN);
    if (row >= p.M || col >= p.N) continue;
    float v = 0.0f;
    for (int sp = 0; sp < p.tail_split; ++sp) v += p.tail_partial[(tt * p.tail_split + sp) * per + e];
    bf16_store(p.out, row, col, v);
  }
}

// Launch with the split tail when a workspace is given and it saves at least
// 1% of the modelled rounds: the tiles past the last full round of every XCD
// run as k-chunks (at least 8 k-tiles each), reduced afterwards.
template <int BM, int BN, int WM, int WN>
hipError_t launch_bf16(Bf16Args p, hipStream_t s, const GemmWorkspace* ws) {
  p.tiles_m = (p.M + BM - 1) / BM;
  p.tiles_n = (p.N + BN - 1) / BN;
  const long tiles = (long)p.tiles_m * p.tiles_n;
  p.tiles_pad = tiles >= 32 ? (int)(((tiles + 7) / 8) * 8) : (int)tiles;
  long tail = 0;
  if (ws && ws->ptr) {
    constexpr long kCusPerXcd = 32;  // MI355X: 256 CUs in 8 XCDs
    constexpr long slots_xcd = kCusPerXcd * ((WM * WN == 4) ? 2 : 1);  // resident blocks per XCD
    const long xcd_full = (tiles / 8) / slots_xcd * slots_xcd;
    const long rest = tiles - 8 * xcd_full;
    const int nk = p.Kp / 32;
    if (rest > 0) {
      const double r0 = std::ceil(std::ceil(double(tiles) / 8) / slots_xcd);
      double best_r = 0.99 * r0;
      for (int sp = 2; sp <= 8; ++sp) {
        const int kt = (nk + sp - 1) / sp, se = (nk + kt - 1) / kt;
        if (se < 2 || kt < 8) continue;
        if (double(rest) * se * BM * BN * sizeof(float) > double(ws->bytes)) continue;
        const double r = double(xcd_full) / slots_xcd +
                         std::ceil(std::ceil(double(rest * se) / 8) / slots_xcd) / se + 0.01 * se;
        if (r < best_r) {
          best_r = r;
          p.tail_split = se;
          p.tail_kt = kt;
        }
      }
      if (p.tail_split > 0) {
        tail = rest;
        p.xcd_full = (int)xcd_full;
        p.tail_units = (int)(rest * p.tail_split);
        p.tail_partial = (float*)ws->ptr;
      }
    }
  }
  const unsigned grid = p.tail_split ? (unsigned)(8 * (p.xcd_full + (p.tail_units + 7) / 8)) : (unsigned)p.tiles_pad;
  hipLaunchKernelGGL((bf16x3_nt_kernel<BM, BN, WM, WN>), dim3(grid), dim3(64 * WM * WN), 0, s, p);
  if (tail > 0) {
    const long n = tail * BM * BN;
    hipLaunchKernelGGL((bf16_tail_reduce_kernel<BM, BN>), dim3((unsigned)std::min<long>((n + 255) / 256, 4096)),
                       dim3(256), 0, s, p, (int)tail);
  }
  return hipGetLastError();
}

// C = A B^T on interleaved BF16 hi/lo operands.  Tile shape by modelled time
// (padding x wave quantisation over 256 CUs / relative throughput measured on
// MI355X): 256 x 256 on 8 waves for large products, 160 x 320 on 8 waves when
// M = P2(O) = 780 (5 x 160 = 800 rows instead of 4 x 256 = 1024), 128 x 128 on
// 4 waves for small ones.  The grouped tile order pays on square products with
// many tile rows (measured ~3% on 12000^3) and not on the P2(O)-row ones.
// 256 x 48 on 4 waves (wave tile 64 x 48) serves N <= 64: the t1 product of
// W_mbej has N = O = 40 (83% of a 48-wide tile against 31% of a 128-wide one).
// force (tests): 0 = 128 x 128, 1 = 256 x 256 automatic order, 2 = 256 x 256
// plain order, 3 = 256 x 256 grouped order, 4 = 160 x 320, 5 = 256 x 48.
hipError_t bf16x3_nt(const uint16_t* A, long lda, const uint16_t* B, long ldb, const Bf16Out& out, int M, int N,
                     int K, hipStream_t s, int force = -1, const GemmWorkspace* ws = nullptr) {
  if (M <= 0 || N <= 0) return hipSuccess;
  // the kernel clamps its k-tile index to [0, K/32 - 1]: it needs one tile
  if (K <= 0) return hipErrorInvalidValue;
  Bf16Args p;
  p.A = A; p.B = B; p.out = out;
  p.lda = lda; p.ldb = ldb;
  p.M = M; p.N = N; p.Kp = ((K + 31) / 32) * 32;
  p.group = 0;
  struct Opt { int BM, BN; double rate; };
  const Opt opts[4] = {{128, 128, 1.0}, {256, 256, 1.65}, {160, 320, 1.55}, {256, 48, 1.45}};
  int best = 0;
  double bt = 1e300;
  for (int c = 0; c < 4; ++c) {
    if (c == 3 && N > 64) continue;  // 256 x 48: for N of one occupied-orbital count
    long tm = (M + opts[c].BM - 1) / opts[c].BM, tn = (N + opts[c].BN - 1) / opts[c].BN;
    double pad = double(tm * opts[c].BM) * double(tn * opts[c].BN) / (double(M) * N);
    double tiles = double(tm * tn);
    double rounds = std::ceil(tiles / 256.0);
    double t = pad * (rounds * 256.0 / tiles) / opts[c].rate;
    if (t < bt) { bt = t; best = c; }
  }
  int group = (best == 1 && (M + 255) / 256 >= 16) ? 8 : 0;
  if (force == 0) best = 0;
  if (force >= 1 && force <= 3) {
    best = 1;
    if (force == 2) group = 0;
    if (force == 3) group = 8;
  }
  if (force == 4) best = 2;
  if (force == 5) best = 3;
  if (best == 0) return launch_bf16<128, 128, 2, 2>(p, s, ws);
  if (best == 2) return launch_bf16<160, 320, 2, 4>(p, s, ws);
  if (best == 3) return launch_bf16<256, 48, 4, 1>(p, s, ws);
  p.group = group;
  return launch_bf16<256, 256, 4, 2>(p, s, ws);
}

}  // namespace

size_t lowprec_plan_bytes(const Dims& d, int precision) {
  if (precision == CCSD_PREC_FP64) return 0;
  const long lpv = pad64(d.pv), lov = pad64(d.ov), O = d.O, V = d.V;
  // dominant-contraction operands: vvvv, ovvv, Doovv, tau, Tring, WT (fp32: 4 B; bf16 hi+lo: 2 x 2 B)
  long elems = d.pv * lpv + d.ov * lpv + d.ov * lov + d.po * lpv + 2 * d.ov * lov;
  size_t bytes = (size_t)elems * 4;
  if (precision == CCSD_PREC_BF16X3) {
    bytes += (size_t)(d.ov * lpv + d.po * lpv) * 4;  // FP32 ovvv, tau as well
    // BF16 operands of the W_mnij, hh-ladder, Y and W_mbej t1 products: oovv,
    // W1^T, tau^T, Tv, Fae, the unpacked <mb||ef> and t1
    const long lpo = pad64(d.po), lv = pad64(V);
    bytes += (size_t)(d.po * lpv + d.po * lpo + d.pv * lpo + d.po * V * lv + V * lv + d.ov * V * lv + O * lv) * 4;
  }
  // secondary FP32 copies: oovv, oovv_vu, oovv_ou, G, ovvv_u; t1, T2u, Tv, W1, W3, W3u, W2, Fae, Fmi
  long sec = d.po * d.pv + d.po * V * V + O * O * d.pv + d.ov * d.ov + d.ov * V * V;
  sec += d.ov + O * O * d.pv + d.po * V * V + d.po * d.po + 2 * O * d.po * V + O * O * O * V + d.po * d.ov +
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

static void conv_il(Ctx* c, uint16_t* dst, const double* src, long rows, long K, long ld) {
  hipLaunchKernelGGL(conv_bf16il_kernel<double>, dim3(grid_for(rows * ld / 8)), dim3(kThreads), 0, c->stream, dst, src, rows,
                     K, K, ld, 1.0);
}

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
  L->ld_po = pad64(d.po);
  L->ld_v = pad64(V);
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
    // the unpack writes Tring32, the W_mbej build WT32, for columns < ov only:
    // zero the padding once
    for (float* q : {L->Tring32, L->WT32}) HIPTRY(hipMemsetAsync(q, 0, (size_t)d.ov * L->ld_ov * 4, s));
    hipLaunchKernelGGL(conv_f32_kernel, dim3(grid_for(d.pv * L->ld_pv)), dim3(kThreads), 0, s, L->vvvv32, c->vvvv,
                       d.pv, d.pv, d.pv, L->ld_pv, 1.0);
    hipLaunchKernelGGL(conv_f32_kernel, dim3(grid_for(d.ov * L->ld_ov)), dim3(kThreads), 0, s, L->Doovv32, c->Doovv,
                       d.ov, d.ov, d.ov, L->ld_ov, 1.0);
    L->have_fp32 = 1;
  }
  if (precision == CCSD_PREC_BF16X3 && !L->have_bf16) {
    TRY(lp_alloc(c, &L->vvvv_b, 2 * d.pv * L->ld_pv));
    TRY(lp_alloc(c, &L->ovvv_b, 2 * d.ov * L->ld_pv));
    TRY(lp_alloc(c, &L->Doovv_b, 2 * d.ov * L->ld_ov));
    TRY(lp_alloc(c, &L->tau_b, 2 * d.po * L->ld_pv));
    TRY(lp_alloc(c, &L->Tring_b, 2 * d.ov * L->ld_ov));
    TRY(lp_alloc(c, &L->WT_b, 2 * d.ov * L->ld_ov));
    TRY(lp_alloc(c, &L->oovv_b, 2 * d.po * L->ld_pv));
    TRY(lp_alloc(c, &L->W1T_b, 2 * d.po * L->ld_po));
    TRY(lp_alloc(c, &L->tauT_b, 2 * d.pv * L->ld_po));
    TRY(lp_alloc(c, &L->Tv_b, 2 * d.po * V * L->ld_v));
    TRY(lp_alloc(c, &L->Fae_b, 2 * V * L->ld_v));
    TRY(lp_alloc(c, &L->ovvv_u_b, 2 * d.ov * V * L->ld_v));
    TRY(lp_alloc(c, &L->t1_b, 2 * O * L->ld_v));
    // WT_b (W_mbej epilogue) and Tring_b (unpack) are written for columns < ov
    // only: zero the padding once
    for (uint16_t* q : {L->WT_b, L->Tring_b}) HIPTRY(hipMemsetAsync(q, 0, (size_t)d.ov * L->ld_ov * 4, s));
    conv_il(c, L->vvvv_b, c->vvvv, d.pv, d.pv, L->ld_pv);
    conv_il(c, L->ovvv_b, c->ovvv, d.ov, d.pv, L->ld_pv);
    conv_il(c, L->Doovv_b, c->Doovv, d.ov, d.ov, L->ld_ov);
    conv_il(c, L->oovv_b, c->oovv, d.po, d.pv, L->ld_pv);
    conv_il(c, L->ovvv_u_b, c->ovvv_u, d.ov * V, V, L->ld_v);
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


int lowprec_amplitudes(Ctx* c, int prec) {
  TRY(lowprec_allocate(c, prec));
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  hipLaunchKernelGGL(conv_f32_kernel, dim3(grid_for(d.po * L->ld_pv)), dim3(kThreads), 0, c->stream, L->tau32, c->tau,
                     d.po, d.pv, d.pv, L->ld_pv, 1.0);
  if (prec == CCSD_PREC_BF16X3) conv_il(c, L->tau_b, c->tau, d.po, d.pv, L->ld_pv);
  return CCSD_OK;
}

RingOut lowprec_ring_out(Ctx* c, int prec) {
  LowPrec* L = c->lp;
  RingOut r;
  r.ld = L->ld_ov;
  if (prec == CCSD_PREC_FP32) {
    r.f = L->Tring32;
  } else {
    r.b = L->Tring_b;
  }
  return r;
}

static int gemv_lp(Ctx* c, int prec, const float* A32, const uint16_t* Ab, const double* x, double* y,
                   double alpha, double beta) {
  const Dims& d = c->d;
  const int rows_per_block = 4;
  const bool f32 = prec == CCSD_PREC_FP32;
  hipLaunchKernelGGL(gemv_lp_kernel, dim3((unsigned)((d.ov + rows_per_block - 1) / rows_per_block)),
                     dim3(64 * rows_per_block), 0, c->stream, y, f32 ? A32 : nullptr, f32 ? nullptr : Ab,
                     c->lp->ld_ov, x, d.ov, d.ov, alpha, beta);
  return CCSD_OK;
}

int lowprec_fme(Ctx* c, int prec, const double* t1) {
  LowPrec* L = c->lp;
  return gemv_lp(c, prec, L->Doovv32, L->Doovv_b, t1, c->Fme, 1.0, 1.0);
}

int lowprec_r1_ring(Ctx* c, int prec) {
  LowPrec* L = c->lp;
  return gemv_lp(c, prec, L->Tring32, L->Tring_b, c->Fme, c->R1, 1.0, 1.0);
}

// C64 (M x N, ldc) = alpha * A B^T + beta * C64 on the stored low-precision
// operands; in BF16 mode the product can also be emitted as an interleaved
// hi/lo operand (Ob, leading dimension ldo).  With C null the FP32 C32 is the
// destination instead (in BF16 mode stored back only if keep32).
// FP32 square (OV)^3 product: one level of Strassen (strassen.hip) when the
// context holds its temporaries, else the engine.  With beta 1 (FP32 result)
// the Strassen product goes to `scratch` (M x N FP32, contiguous) and is added
// after.
static hipError_t f32_square(Ctx* c, const GemmDesc<float>& g, float* scratch) {
  if (c->strassen_tmp && g.M == c->d.ov) {
    GemmDesc<float> h = g;
    const bool acc = g.beta != 0.0f;
    if (acc) {
      h.C = scratch;
      h.Cd = nullptr;
      h.sCm = g.N;
      h.beta = 0.0f;
    }
    if ((!acc || (scratch && g.C && !g.Cd && g.beta == 1.0f)) && strassen_applies(h, 2)) {
      hipError_t e = gemm_strassen(h, c->strassen_tmp, c->ws, c->stream);
      if (e == hipSuccess && acc) e = strassen_accumulate(g.C, g.sCm, scratch, g.N, g.M, g.N, c->stream);
      return e;
    }
  }
  return gemm<float>(g, c->ws, c->stream);
}

// square: an (OV)^3 product (FP32: through f32_square, with `scratch` for beta 1)
static int nt_product(Ctx* c, int prec, const float* A32, const uint16_t* Ab, long lda, const float* B32,
                      const uint16_t* Bb, long ldb, int M, int N, int K, double* C, long ldc, double alpha,
                      double beta, uint16_t* Ob = nullptr, long ldo = 0, float* C32 = nullptr, bool keep32 = true,
                      bool square = false, float* scratch = nullptr) {
  hipError_t e;
  if (prec == CCSD_PREC_FP32) {
    GemmDesc<float> g;
    g.A = A32; g.B = B32; g.Cd = C;
    if (!C) g.C = C32;
    g.M = M; g.N = N; g.K = K;
    g.sAm = lda; g.sAk = 1; g.sBk = 1; g.sBn = ldb; g.sCm = ldc; g.sCn = 1;
    g.alpha = (float)alpha; g.beta = (float)beta;  // 1, 0.5, 0: exact in FP32
    e = square ? f32_square(c, g, scratch) : gemm<float>(g, c->ws, c->stream);
  } else {
    Bf16Out o;
    o.C64 = C; o.ldc = ldc; o.alpha = alpha; o.beta = beta;
    if (!C) { o.C32 = C32; o.c32_keep = keep32; }
    o.Ob = Ob; o.ldo = ldo;
    e = bf16x3_nt(Ab, lda, Bb, ldb, o, M, N, K, c->stream, -1, &c->ws);
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
  return nt_product(c, prec, L->tau32, L->tau_b, L->ld_pv, L->ovvv32, L->ovvv_b, L->ld_pv, (int)d.po, (int)d.ov,
                    (int)d.pv, c->W2, d.ov, 1.0, 0.0);
}

float* lowprec_wmbej_acc(Ctx* c, int prec, long* ld) {
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  if (prec == CCSD_PREC_FP32) {
    *ld = L->ld_ov;
    return L->WT32;
  }
  *ld = d.ov;
  return reinterpret_cast<float*>(c->X) + d.ov * d.ov;
}

int lowprec_wmbej(Ctx* c, int prec) {
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  long ld;
  float* W = lowprec_wmbej_acc(c, prec, &ld);
  // Doovv is symmetric, so its rows serve as the k-contiguous B operand.  FP32:
  // W is the ring GEMM's operand itself, and X (free until the ring GEMM)
  // holds the Strassen product before it is added.  BF16: the finished W_mbej
  // is written only as the ring GEMM's hi/lo operand.
  return nt_product(c, prec, L->Tring32, L->Tring_b, L->ld_ov, L->Doovv32, L->Doovv_b, L->ld_ov, (int)d.ov,
                    (int)d.ov, (int)d.ov, nullptr, ld, 0.5, 1.0, L->WT_b, L->ld_ov, W, false, true,
                    reinterpret_cast<float*>(c->X));
}

int lowprec_ladder(Ctx* c, int prec) {
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  // Rp[ij][ab] += sum_{e<f} tau[ij][ef] <ab||ef>: row ab of the packed vvvv
  // block is the k-contiguous B row of output column ab.
  return nt_product(c, prec, L->tau32, L->tau_b, L->ld_pv, L->vvvv32, L->vvvv_b, L->ld_pv, (int)d.po, (int)d.pv,
                    (int)d.pv, c->Rp, d.pv, 1.0, 1.0);
}

int lowprec_ring(Ctx* c, int prec) {
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  // the W_mbej operand (WT32 / WT_b) is final after lowprec_wmbej
  return nt_product(c, prec, L->Tring32, L->Tring_b, L->ld_ov, L->WT32, L->WT_b, L->ld_ov, (int)d.ov, (int)d.ov,
                    (int)d.ov, c->X, d.ov, 1.0, 0.0, nullptr, 0, nullptr, true, true);
}

// ---------------------------------------------------------------- adjoint (Lambda) pieces
static int bf16_nt(Ctx* c, const uint16_t* A, long lda, const uint16_t* B, long ldb, long M, long N, long K,
                   double* C, long ldc, double alpha, double beta) {
  Bf16Out o;
  o.C64 = C; o.ldc = ldc; o.alpha = alpha; o.beta = beta;
  hipError_t e = bf16x3_nt(A, lda, B, ldb, o, (int)M, (int)N, (int)K, c->stream, -1, &c->ws);
  if (e != hipSuccess) {
    c->last_error = hipGetErrorString(e);
    return CCSD_ERR_HIP;
  }
  return CCSD_OK;
}

// W1 += <mn||ef> tau_ij^ef (C[mn][ij]; W_1 of Fill_W1_and_W3 :659-1029): both
// operands k-contiguous over the virtual pairs (tau_b is the ladder's operand).
int lowprec_w1_bf16(Ctx* c) {
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  return bf16_nt(c, L->oovv_b, L->ld_pv, L->tau_b, L->ld_pv, d.po, d.po, d.pv, c->W1, d.po, 1.0, 1.0);
}

// Rp[ij][ab] += sum_mn W1[mn][ij] tau[mn][ab] (the hh ladder of Fill_t2_new
// :2001-2458): as C = A B^T with A = W1^T and B = tau^T, both formed here.
int lowprec_hh_bf16(Ctx* c) {
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(conv_bf16il_t_kernel, dim3(grid_for(d.po * L->ld_po / 8)), dim3(kThreads), 0, s, L->W1T_b,
                     c->W1, d.po, d.po, d.po, L->ld_po);
  hipLaunchKernelGGL(conv_bf16il_t_kernel, dim3(grid_for(d.pv * L->ld_po / 8)), dim3(kThreads), 0, s, L->tauT_b,
                     c->tau, d.pv, d.po, d.pv, L->ld_po);
  HIPTRY(hipGetLastError());
  return bf16_nt(c, L->W1T_b, L->ld_po, L->tauT_b, L->ld_po, d.po, d.pv, d.po, c->Rp, d.pv, 1.0, 1.0);
}

// Y[ij][a][b] = sum_e t_ij^ae F_be (Fill_t2_new's P(ab) F_ae term): rows (ij, a)
// of Tv against rows b of F_ae, both k-contiguous over e.
int lowprec_y_bf16(Ctx* c) {
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  const long V = d.V;
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(conv_bf16il_kernel<float>, dim3(grid_for(d.po * V * L->ld_v / 8)), dim3(kThreads), 0, s,
                     L->Tv_b, L->Tv_f, d.po * V, V, V, L->ld_v, 1.0);
  hipLaunchKernelGGL(conv_bf16il_kernel<double>, dim3(grid_for(V * L->ld_v / 8)), dim3(kThreads), 0, s, L->Fae_b,
                     c->Fae, V, V, V, L->ld_v, 1.0);
  HIPTRY(hipGetLastError());
  return bf16_nt(c, L->Tv_b, L->ld_v, L->Fae_b, L->ld_v, d.po * V, V, V, c->Y, V, 1.0, 0.0);
}

// S[(m,b,e)][j] = sum_f <mb||ef> t_j^f into FP32 S32 (rows of O), the t1 term
// of W_mbej (W_4, Fill_W4_MPI :4781-5089): rows of the unpacked block against
// the rows of t1, both k-contiguous over f, on the 256 x 48 tile.
int lowprec_wmbej_s_bf16(Ctx* c, const double* t1, float* S32) {
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  const long V = d.V;
  conv_il(c, L->t1_b, t1, d.O, V, L->ld_v);
  HIPTRY(hipGetLastError());
  Bf16Out o;
  o.C = S32;
  o.ldc = d.O;
  o.alpha = 1.0;
  o.beta = 0.0;
  hipError_t e = bf16x3_nt(L->ovvv_u_b, L->ld_v, L->t1_b, L->ld_v, o, (int)(d.ov * V), (int)d.O, (int)V, c->stream,
                           -1, &c->ws);
  if (e != hipSuccess) {
    c->last_error = hipGetErrorString(e);
    return CCSD_ERR_HIP;
  }
  return CCSD_OK;
}

int lowprec_adj_setup(Ctx* c) {
  TRY(lowprec_allocate(c, CCSD_PREC_BF16X3));
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  const long O = d.O, V = d.V;
  hipStream_t s = c->stream;
  if (!L->have_adj) {
    TRY(lp_alloc(c, &L->WTt_b, 2 * d.ov * L->ld_ov));
    TRY(lp_alloc(c, &L->WTbar_b, 2 * d.ov * L->ld_ov));
    TRY(lp_alloc(c, &L->W2bar_b, 2 * d.po * L->ld_ov));
    TRY(lp_alloc(c, &L->lam1_f, d.ov));
    TRY(lp_alloc(c, &L->lam2_f, d.po * L->ld_pv));
    TRY(lp_alloc(c, &L->Yb_f, d.po * V * V));
    TRY(lp_alloc(c, &L->Zb_f, O * O * d.pv));
    TRY(lp_alloc(c, &L->Faeb_f, V * V));
    TRY(lp_alloc(c, &L->W1b_f, d.po * d.po));
    TRY(lp_alloc(c, &L->Sb_f, O * V * V * O));
    TRY(lp_alloc(c, &L->Ub_f, O * O * d.ov));
    TRY(lp_alloc(c, &L->W3b_f, O * d.po * V));
    TRY(lp_alloc(c, &L->Fmib_f, O * O));
    L->have_adj = 1;
  }
  // FP32 copies of the forward intermediates at the converged amplitudes (the
  // FP64 right-hand side just computed), for the secondary reverse products
  const double* t1 = c->amp;
  k_to_f32(L->t1_f, t1, d.ov, s);
  k_to_f32(L->T2u_f, c->T2u, O * O * d.pv, s);
  k_to_f32(L->Tv_f, c->Tv, d.po * V * V, s);
  k_to_f32(L->W1_f, c->W1, d.po * d.po, s);
  k_to_f32(L->W3_f, c->W3, O * d.po * V, s);
  k_to_f32(L->W2_f, c->W2, d.po * d.ov, s);
  k_to_f32(L->Fmi_f, c->Fmi, O * O, s);
  k_to_f32(L->Fae_f, c->Fae, V * V, s);
  hipLaunchKernelGGL(conv_f32_kernel, dim3(grid_for(d.po * L->ld_pv)), dim3(kThreads), 0, s, L->tau32, c->tau, d.po,
                     d.pv, d.pv, L->ld_pv, 1.0);
  if (!L->have_ovvvT) {  // <ia||bc>^T: rows (b<c), k = (i,a); static
    TRY(lp_alloc(c, &L->ovvvT_b, 2 * d.pv * L->ld_ov));
    hipLaunchKernelGGL(conv_bf16il_t_kernel, dim3(grid_for(d.pv * L->ld_ov / 8)), dim3(kThreads), 0, s, L->ovvvT_b,
                       c->ovvv, d.pv, d.ov, d.pv, L->ld_ov);
    L->have_ovvvT = 1;
  }
  // fixed at the converged amplitudes: W_mbej^T and Tring (FP64 forward intermediates)
  hipLaunchKernelGGL(conv_bf16il_t_kernel, dim3(grid_for(d.ov * L->ld_ov / 8)), dim3(kThreads), 0, s, L->WTt_b,
                     c->WT, d.ov, d.ov, d.ov, L->ld_ov);
  conv_il(c, L->Tring_b, c->Tring, d.ov, d.ov, L->ld_ov);
  HIPTRY(hipGetLastError());
  return CCSD_OK;
}

int lowprec_adj_lambda(Ctx* c, const double* lam) {
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  hipStream_t s = c->stream;
  const double* lam2 = lam + d.ov;
  k_to_f32(L->lam1_f, lam, d.ov, s);
  hipLaunchKernelGGL(conv_f32_kernel, dim3(grid_for(d.po * L->ld_pv)), dim3(kThreads), 0, s, L->lam2_f, lam2, d.po,
                     d.pv, d.pv, L->ld_pv, 1.0);
  k_unpack_vir_ring_lp(L->Yb_f, RingOut{}, lam2, d, s);
  k_unpack_occ_f32(L->Zb_f, L->Zb_f, lam2, lam2, d, s);
  HIPTRY(hipGetLastError());
  return CCSD_OK;
}

int lowprec_adj_ring(Ctx* c, double* Tringb, double* WTb, const double* lam2) {
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  // Xb (symmetric) as a BF16 operand, in the forward's WT_b buffer (rebuilt by every forward step)
  RingOut r;
  r.b = L->WT_b;
  r.ld = L->ld_ov;
  k_unpack_vir_ring_lp(nullptr, r, lam2, d, c->stream);
  TRY(bf16_nt(c, L->WT_b, L->ld_ov, L->WTt_b, L->ld_ov, d.ov, d.ov, d.ov, Tringb, d.ov, 1.0, 1.0));
  return bf16_nt(c, L->WT_b, L->ld_ov, L->Tring_b, L->ld_ov, d.ov, d.ov, d.ov, WTb, d.ov, 1.0, 1.0);
}

int lowprec_adj_ladder(Ctx* c, double* taub, const double* lam2) {
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  conv_il(c, L->tau_b, lam2, d.po, d.pv, L->ld_pv);  // forward's tau_b as scratch
  return bf16_nt(c, L->tau_b, L->ld_pv, L->vvvv_b, L->ld_pv, d.po, d.pv, d.pv, taub, d.pv, 1.0, 1.0);
}

int lowprec_adj_wmbej(Ctx* c, double* Tringb, const double* WTb) {
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  conv_il(c, L->WTbar_b, WTb, d.ov, d.ov, L->ld_ov);
  return bf16_nt(c, L->WTbar_b, L->ld_ov, L->Doovv_b, L->ld_ov, d.ov, d.ov, d.ov, Tringb, d.ov, 0.5, 1.0);
}

int lowprec_adj_w2(Ctx* c, double* taub, const double* W2b) {
  LowPrec* L = c->lp;
  const Dims& d = c->d;
  conv_il(c, L->W2bar_b, W2b, d.po, d.ov, L->ld_ov);
  return bf16_nt(c, L->W2bar_b, L->ld_ov, L->ovvvT_b, L->ld_ov, d.po, d.pv, d.ov, taub, d.pv, 1.0, 1.0);
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
  const bool strassen = (prec >> 17) & 1;  // FP32: one level of Strassen (M = N = K even, beta 0)
  prec &= 0xff;
  const long ld = pad64(K);
  float *A32 = nullptr, *B32 = nullptr;
  uint16_t *Ab = nullptr, *Bb = nullptr;
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
    e = al((void**)&Ab, (size_t)M * ld * 4);
    if (e == hipSuccess) e = al((void**)&Bb, (size_t)N * ld * 4);
  }
  void* wsp = nullptr;
  constexpr size_t kWs = 64u << 20;
  if (e == hipSuccess) e = al(&wsp, kWs);
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
      GemmWorkspace ws{wsp, kWs};
      if (strassen) {
        if (!strassen_applies(g, 2)) {
          for (void* q : owned) (void)hipFree(q);
          return CCSD_ERR_ARG;
        }
        double* tmp = nullptr;
        e = al((void**)&tmp, strassen_temp_elems(M) * sizeof(double));
        if (e == hipSuccess) e = gemm_strassen(g, tmp, ws, s);
      } else {
        e = gemm<float>(g, ws, s, nullptr, tile - 1, -1);
      }
    } else if (strassen) {
      for (void* q : owned) (void)hipFree(q);
      return CCSD_ERR_ARG;
    } else {
      hipLaunchKernelGGL(conv_bf16il_kernel<double>, dim3(grid_for(M * ld / 8)), dim3(kThreads), 0, s, Ab, A, (long)M,
                         (long)K, (long)K, ld, 1.0);
      hipLaunchKernelGGL(conv_bf16il_kernel<double>, dim3(grid_for(N * ld / 8)), dim3(kThreads), 0, s, Bb, B, (long)N,
                         (long)K, (long)K, ld, 1.0);
      Bf16Out o;
      o.C64 = C; o.ldc = N; o.alpha = alpha; o.beta = beta;
      GemmWorkspace ws{wsp, kWs};
      e = bf16x3_nt(Ab, ld, Bb, ld, o, M, N, K, s, tile - 1, &ws);
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  for (void* p : owned) (void)hipFree(p);
  return e == hipSuccess ? CCSD_OK : CCSD_ERR_HIP;
}

}  // namespace ccsd
