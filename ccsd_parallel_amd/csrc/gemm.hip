// MFMA GEMM engine (gfx950).  See gemm.hpp for the contract.
//
// Design (CDNA4-first, not a CUDA warp tiling):
//  * 64-wide wavefronts, each computing a (BM/WM) x (BN/WN) sub-tile as a grid
//    of 16x16 MFMA tiles: v_mfma_f64_16x16x4f64 (FP64) or v_mfma_f32_16x16x4f32.
//  * A and B tiles are staged through LDS in k-major order ([k][m], [k][n]) so
//    that a wave's operand fetch (lane l reads row l%16 of k-slice l/16) is a
//    contiguous 16-element run per k; double-buffered, one barrier per k-tile.
//  * Global loads are coalesced along whichever operand stride is unit
//    (template switches AK / BNM select the loop order); any strides work.
//  * Workgroup ids are remapped so that the tiles sharing one B column block
//    run on the same XCD (the dispatcher deals workgroups round-robin over the
//    8 XCDs, each with its own L2).
//  * Split-K and batch reductions write per-slice partials and are summed in a
//    fixed order by a second kernel: results are deterministic.
#include "gemm.hpp"
#include <algorithm>
#include <cstdio>

namespace ccsd {

namespace {

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  using acc_t = double __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // FP64 16x16x4 output: lane l, element r -> row l/16 + 4r, col l%16
  static __device__ __forceinline__ int out_row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
  using acc_t = float __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // FP32 16x16x4 output: lane l, element r -> row 4*(l/16) + r, col l%16
  static __device__ __forceinline__ int out_row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// C[off] = alpha * s + beta * C[off], into the FP64 destination Cd when given.
template <typename T>
__device__ __forceinline__ void write_out(const GemmDesc<T>& d, long off, T s) {
  if (d.Cd) {
    double v = (double)d.alpha * (double)s;
    if (d.beta != T(0)) v += (double)d.beta * d.Cd[off];
    d.Cd[off] = v;
  } else {
    T v = d.alpha * s;
    if (d.beta != T(0)) v += d.beta * d.C[off];
    d.C[off] = v;
  }
}

template <typename T>
struct KArgs {
  GemmDesc<T> d;
  T* partial;      // [splits][nbatch][M][N] when non-null
  // tiles_pad: tiles rounded up to a multiple of 8 when the XCD-aware order is
  // used; equal to tiles (not a multiple of 8) for the plain order
  int tiles_m, tiles_n, tiles_pad;
  int splits, kchunk;
};

template <typename T, int BM, int BN, int BK, int WM, int WN, bool AK, bool BNM>
__global__ __launch_bounds__(64 * WM * WN) void gemm_kernel(KArgs<T> p) {
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = BM / WM / 16;
  constexpr int TN = BN / WN / 16;
  static_assert(TM >= 1 && TN >= 1, "wave tile too small");
  // LDS layouts.  A k-contiguous operand stays k-contiguous ([m][LDK]): the
  // 16-byte stores of a row are contiguous and, with LDK = BK + 2 doubles, the
  // 16 rows of one MFMA operand fetch fall in distinct banks.  An m/n-contiguous
  // operand is stored [k][LDM]; the pitch BM + 16 puts consecutive k-rows of a
  // fetch in opposite bank halves.
  constexpr int LDK = BK + (sizeof(T) == 8 ? 2 : 4);
  constexpr int LDMA = BM + (sizeof(T) == 8 ? 16 : 32);
  constexpr int LDMB = BN + (sizeof(T) == 8 ? 16 : 32);
  constexpr int ASZ = AK ? BM * LDK : BK * LDMA;
  constexpr int BSZ = BNM ? BK * LDMB : BN * LDK;
  __shared__ __attribute__((aligned(16))) T As[2][ASZ];
  __shared__ __attribute__((aligned(16))) T Bs[2][BSZ];
  auto aidx = [](int m, int k) { return AK ? m * LDK + k : k * LDMA + m; };
  auto bidx = [](int k, int n) { return BNM ? k * LDMB + n : n * LDK + k; };

  const GemmDesc<T>& d = p.d;
  // XCD-aware tile id: hardware places linear workgroup x on XCD x % 8, so
  // XCD j takes the contiguous tile range [j*gx/8, (j+1)*gx/8).  With few tiles
  // (gx not a multiple of 8) the plain order spreads the batches and K-splits
  // over all XCDs instead of stacking every real tile on the low ones.
  const int gx = p.tiles_pad;
  const int tile = (gx & 7) ? (int)blockIdx.x : (int)((blockIdx.x & 7) * (gx >> 3) + (blockIdx.x >> 3));
  if (tile >= p.tiles_m * p.tiles_n) return;
  const int tm = tile % p.tiles_m;
  const int tn = tile / p.tiles_m;
  const int batch = blockIdx.y;
  const int split = blockIdx.z;
  const int b1 = batch / d.nb2, b2 = batch % d.nb2;

  const T* __restrict__ A = d.A + b1 * d.sA1 + b2 * d.sA2;
  const T* __restrict__ B = d.B + b1 * d.sB1 + b2 * d.sB2;
  const int m0 = tm * BM, n0 = tn * BN;
  const int kbeg = split * p.kchunk;
  const int kend = min(d.K, kbeg + p.kchunk);

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wy = wave / WN, wx = wave % WN;
  const int wm0 = wy * (BM / WM), wn0 = wx * (BN / WN);

  constexpr int AU = BM * BK / 2;
  constexpr int BU = BK * BN / 2;
  constexpr int AIT = (AU + NT - 1) / NT;
  constexpr int BIT = (BU + NT - 1) / NT;
  T ra[AIT][2], rb[BIT][2];

  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int it = 0; it < AIT; ++it) {
      int u = it * NT + tid;
      int m, kk;
      if (AK) { m = u / (BK / 2); kk = (u % (BK / 2)) * 2; }
      else    { kk = u / (BM / 2); m = (u % (BM / 2)) * 2; }
      int gm = m0 + m, gk = k0 + kk;
      T v0 = 0, v1 = 0;
      if (u < AU) {
        if (AK) {
          if (gm < d.M) {
            if (gk < kend) v0 = A[gm * d.sAm + gk * d.sAk];
            if (gk + 1 < kend) v1 = A[gm * d.sAm + (gk + 1) * d.sAk];
          }
        } else {
          if (gk < kend) {
            if (gm < d.M) v0 = A[gm * d.sAm + gk * d.sAk];
            if (gm + 1 < d.M) v1 = A[(gm + 1) * d.sAm + gk * d.sAk];
          }
        }
      }
      ra[it][0] = v0; ra[it][1] = v1;
    }
#pragma unroll
    for (int it = 0; it < BIT; ++it) {
      int u = it * NT + tid;
      int n, kk;
      if (BNM) { kk = u / (BN / 2); n = (u % (BN / 2)) * 2; }
      else     { n = u / (BK / 2); kk = (u % (BK / 2)) * 2; }
      int gn = n0 + n, gk = k0 + kk;
      T v0 = 0, v1 = 0;
      if (u < BU) {
        if (BNM) {
          if (gk < kend) {
            if (gn < d.N) v0 = B[gk * d.sBk + gn * d.sBn];
            if (gn + 1 < d.N) v1 = B[gk * d.sBk + (gn + 1) * d.sBn];
          }
        } else {
          if (gn < d.N) {
            if (gk < kend) v0 = B[gk * d.sBk + gn * d.sBn];
            if (gk + 1 < kend) v1 = B[(gk + 1) * d.sBk + gn * d.sBn];
          }
        }
      }
      rb[it][0] = v0; rb[it][1] = v1;
    }
  };
  auto store_tiles = [&](int buf) {
#pragma unroll
    for (int it = 0; it < AIT; ++it) {
      int u = it * NT + tid;
      if (u >= AU) break;
      if (AK) { int m = u / (BK / 2), kk = (u % (BK / 2)) * 2;
                As[buf][aidx(m, kk)] = ra[it][0]; As[buf][aidx(m, kk + 1)] = ra[it][1]; }
      else    { int kk = u / (BM / 2), m = (u % (BM / 2)) * 2;
                As[buf][aidx(m, kk)] = ra[it][0]; As[buf][aidx(m + 1, kk)] = ra[it][1]; }
    }
#pragma unroll
    for (int it = 0; it < BIT; ++it) {
      int u = it * NT + tid;
      if (u >= BU) break;
      if (BNM) { int kk = u / (BN / 2), n = (u % (BN / 2)) * 2;
                 Bs[buf][bidx(kk, n)] = rb[it][0]; Bs[buf][bidx(kk, n + 1)] = rb[it][1]; }
      else     { int n = u / (BK / 2), kk = (u % (BK / 2)) * 2;
                 Bs[buf][bidx(kk, n)] = rb[it][0]; Bs[buf][bidx(kk + 1, n)] = rb[it][1]; }
    }
  };

  using acc_t = typename Mfma<T>::acc_t;
  acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  const int ntiles = (kend > kbeg) ? (kend - kbeg + BK - 1) / BK : 0;
  if (ntiles > 0) {
    load_tiles(kbeg);
    store_tiles(0);
    __syncthreads();
  }
  for (int kt = 0; kt < ntiles; ++kt) {
    const int cur = kt & 1;
    const bool more = (kt + 1) < ntiles;
    if (more) load_tiles(kbeg + (kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kr = ks + (lane >> 4);
      T a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) a[i] = As[cur][aidx(wm0 + i * 16 + (lane & 15), kr)];
#pragma unroll
      for (int j = 0; j < TN; ++j) b[j] = Bs[cur][bidx(kr, wn0 + j * 16 + (lane & 15))];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = Mfma<T>::mma(a[i], b[j], acc[i][j]);
    }
    if (more) store_tiles(cur ^ 1);
    __syncthreads();
  }

  // Epilogue (output row of each accumulator element from the MFMA trait).
  if (p.partial) {
    const long MN = (long)d.M * d.N;
    const int nbatch = d.nb1 * d.nb2;
    T* W = p.partial + ((long)split * nbatch + batch) * MN;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int row = m0 + wm0 + i * 16 + Mfma<T>::out_row(lane, r);
          int col = n0 + wn0 + j * 16 + (lane & 15);
          if (row < d.M && col < d.N) W[(long)row * d.N + col] = acc[i][j][r];
        }
  } else {
    const long cbase = b1 * d.sC1 + b2 * d.sC2;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int row = m0 + wm0 + i * 16 + Mfma<T>::out_row(lane, r);
          int col = n0 + wn0 + j * 16 + (lane & 15);
          if (row < d.M && col < d.N) write_out(d, cbase + row * d.sCm + col * d.sCn, (T)acc[i][j][r]);
        }
  }
}

// Many partials per output (long split-K on a small C): one workgroup per output
// element, threads stride over the slices, fixed-order tree sum.
template <typename T>
__global__ void gemm_reduce_wide_kernel(GemmDesc<T> d, const T* __restrict__ W, int splits) {
  const long MN = (long)d.M * d.N;
  const int nbatch = d.nb1 * d.nb2;
  const long e = blockIdx.x;  // element within one output matrix
  const int bo = blockIdx.y;  // output batch (0 when reduce_batch)
  const int nslices = d.reduce_batch ? splits * nbatch : splits;
  T s = 0;
  for (int q = threadIdx.x; q < nslices; q += blockDim.x) {
    long slice = d.reduce_batch ? q : (long)q * nbatch + bo;
    s += W[slice * MN + e];
  }
  __shared__ T sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int row = (int)(e / d.N), col = (int)(e % d.N);
    long off = row * d.sCm + col * d.sCn;
    if (!d.reduce_batch) off += (bo / d.nb2) * d.sC1 + (bo % d.nb2) * d.sC2;
    write_out(d, off, sh[0]);
  }
}

// Sum partials [splits][nbatch][M][N] in a fixed order and apply alpha/beta.
template <typename T>
__global__ void gemm_reduce_kernel(GemmDesc<T> d, const T* __restrict__ W, int splits) {
  const long MN = (long)d.M * d.N;
  const int nbatch = d.nb1 * d.nb2;
  const int nout = d.reduce_batch ? 1 : nbatch;
  const long total = MN * nout;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int bo = (int)(idx / MN);
    const long e = idx % MN;
    const int row = (int)(e / d.N), col = (int)(e % d.N);
    T s = 0;
    for (int sp = 0; sp < splits; ++sp) {
      if (d.reduce_batch) {
        for (int b = 0; b < nbatch; ++b) s += W[((long)sp * nbatch + b) * MN + e];
      } else {
        s += W[((long)sp * nbatch + bo) * MN + e];
      }
    }
    long off = row * d.sCm + col * d.sCn;
    if (!d.reduce_batch) off += (bo / d.nb2) * d.sC1 + (bo % d.nb2) * d.sC2;
    write_out(d, off, s);
  }
}

// y(m) = alpha * sum_k A(m,k) x(k) + beta * y(m): one wavefront per row,
// lanes stride over k (coalesced when sAk == 1).  Used for N == 1.
template <typename T>
__global__ void gemv_kernel(GemmDesc<T> d) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (row >= d.M) return;
  const T* a = d.A + (long)row * d.sAm;
  T s = 0;
  for (int k = lane; k < d.K; k += 64) s += a[k * d.sAk] * d.B[k * d.sBk];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) write_out(d, (long)row * d.sCm, s);
}

// rate: FP64 TF/s of the config on a 12000^3 GEMM on MI355X (tile-sweep probe);
// slots: resident workgroups per CU (LDS / register limited).
struct Cfg { int BM, BN, BK, WM, WN; float rate; int slots; };
constexpr Cfg kCfgs[] = {
    {128, 256, 16, 2, 4, 54.4f, 1},
    {256, 128, 16, 4, 2, 54.0f, 1},
    {128, 128, 16, 2, 2, 39.1f, 2},
    {64, 128, 16, 2, 2, 45.8f, 2},
    {128, 64, 16, 2, 2, 45.6f, 2},
    {64, 64, 16, 2, 2, 44.7f, 4},
    {32, 64, 16, 1, 2, 25.0f, 4},
    {64, 32, 16, 2, 1, 25.0f, 4},
    // 48-wide tiles for the O = 40 edges (M or N of 40 fills 83% instead of 62%)
    {128, 48, 16, 4, 1, 42.0f, 2},
    {48, 128, 16, 1, 4, 42.0f, 2},
    {48, 64, 16, 1, 2, 38.0f, 4},
    {64, 48, 16, 2, 1, 38.0f, 4},
};
constexpr int kNumCus = 256;
constexpr int kNumCfgs = sizeof(kCfgs) / sizeof(kCfgs[0]);

template <typename T, int BM, int BN, int BK, int WM, int WN>
hipError_t launch_cfg(const KArgs<T>& ka, bool AK, bool BNM, dim3 grid, hipStream_t s) {
  constexpr int NT = 64 * WM * WN;
  if (AK && BNM)        hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WM, WN, true, true>), grid, dim3(NT), 0, s, ka);
  else if (AK && !BNM)  hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WM, WN, true, false>), grid, dim3(NT), 0, s, ka);
  else if (!AK && BNM)  hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WM, WN, false, true>), grid, dim3(NT), 0, s, ka);
  else                  hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WM, WN, false, false>), grid, dim3(NT), 0, s, ka);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_any(int c, const KArgs<T>& ka, bool AK, bool BNM, dim3 grid, hipStream_t s) {
  switch (c) {
    case 0: return launch_cfg<T, 128, 256, 16, 2, 4>(ka, AK, BNM, grid, s);
    case 1: return launch_cfg<T, 256, 128, 16, 4, 2>(ka, AK, BNM, grid, s);
    case 2: return launch_cfg<T, 128, 128, 16, 2, 2>(ka, AK, BNM, grid, s);
    case 3: return launch_cfg<T, 64, 128, 16, 2, 2>(ka, AK, BNM, grid, s);
    case 4: return launch_cfg<T, 128, 64, 16, 2, 2>(ka, AK, BNM, grid, s);
    case 5: return launch_cfg<T, 64, 64, 16, 2, 2>(ka, AK, BNM, grid, s);
    case 6: return launch_cfg<T, 32, 64, 16, 1, 2>(ka, AK, BNM, grid, s);
    case 7: return launch_cfg<T, 64, 32, 16, 2, 1>(ka, AK, BNM, grid, s);
    case 8: return launch_cfg<T, 128, 48, 16, 4, 1>(ka, AK, BNM, grid, s);
    case 9: return launch_cfg<T, 48, 128, 16, 1, 4>(ka, AK, BNM, grid, s);
    case 10: return launch_cfg<T, 48, 64, 16, 1, 2>(ka, AK, BNM, grid, s);
    default: return launch_cfg<T, 64, 48, 16, 2, 1>(ka, AK, BNM, grid, s);
  }
}

}  // namespace

template <typename T>
hipError_t gemm(const GemmDesc<T>& d, const GemmWorkspace& ws, hipStream_t stream,
                GemmLaunchInfo* info, int force_config, int force_split) {
  if (d.M <= 0 || d.N <= 0) return hipSuccess;
  const int nbatch = d.nb1 * d.nb2;
  if (nbatch <= 0 || nbatch > 65535) return hipErrorInvalidValue;
  if (d.K <= 0) {
    // C = beta*C (alpha*0): reuse the reduce kernel with zero splits.
    if (d.reduce_batch) {
      GemmDesc<T> e = d; e.nb1 = 1; e.nb2 = 1; e.reduce_batch = false;
      hipLaunchKernelGGL(gemm_reduce_kernel<T>, dim3(256), dim3(256), 0, stream, e, (const T*)nullptr, 0);
    } else {
      hipLaunchKernelGGL(gemm_reduce_kernel<T>, dim3(256), dim3(256), 0, stream, d, (const T*)nullptr, 0);
    }
    return hipGetLastError();
  }
  if (d.N == 1 && nbatch == 1 && force_config < 0) {
    const int rows_per_block = 4;
    hipLaunchKernelGGL(gemv_kernel<T>, dim3((d.M + rows_per_block - 1) / rows_per_block), dim3(64 * rows_per_block),
                       0, stream, d);
    if (info) *info = GemmLaunchInfo{d.M, 1, 1, 1, -1};
    return hipGetLastError();
  }
  // A is "k-contiguous" unless its m stride is the unit one (and likewise B).
  const bool AK = !(d.sAm == 1 && d.sAk != 1);
  const bool BNM = !(d.sBk == 1 && d.sBn != 1);

  // Pick the tile shape with the least modelled time: padding waste times
  // wave quantisation over all CUs, divided by the shape's measured rate.
  int best = 0;
  double best_t = 1e300;
  for (int c = 0; c < kNumCfgs; ++c) {
    const Cfg& g = kCfgs[c];
    long tm = (d.M + g.BM - 1) / g.BM, tn = (d.N + g.BN - 1) / g.BN;
    double pad = (double(tm * g.BM) * double(tn * g.BN)) / (double(d.M) * d.N);
    double tiles = double(tm * tn) * nbatch;
    double slots = double(kNumCus) * g.slots;
    double rounds = std::ceil(tiles / slots);
    double quant = rounds * slots / tiles;
    // few tiles and a long K are fixed by split-K below; do not penalise them twice
    if (tiles < slots && d.K >= 16 * g.BK * 8) quant = 1.0;
    double t = pad * quant / g.rate;
    if (t < best_t) { best_t = t; best = c; }
  }
  if (force_config >= 0 && force_config < kNumCfgs) best = force_config;
  const Cfg& g = kCfgs[best];
  const int tiles_m = (d.M + g.BM - 1) / g.BM;
  const int tiles_n = (d.N + g.BN - 1) / g.BN;
  const long tiles = (long)tiles_m * tiles_n;

  // split-K when the output tiles cannot fill the chip and K is long
  int splits = 1;
  const bool need_partial_for_batch = d.reduce_batch && nbatch > 1;
  if (force_split > 0) {
    splits = force_split;
  } else {
    long work = tiles * nbatch;
    const long target = 2L * kNumCus * g.slots;
    if (work < target) {
      long want = (target + work - 1) / work;
      long kmax = std::max(1, d.K / (g.BK * 8));
      splits = (int)std::max(1L, std::min(want, kmax));
    }
  }
  const size_t part_bytes_per = (size_t)d.M * d.N * nbatch * sizeof(T);
  bool use_partial = splits > 1 || need_partial_for_batch;
  if (use_partial) {
    size_t maxs = ws.bytes / std::max<size_t>(1, part_bytes_per);
    if (maxs < 1) {
      if (need_partial_for_batch) return hipErrorOutOfMemory;
      splits = 1;
      use_partial = false;
    } else {
      splits = (int)std::min<size_t>(splits, maxs);
      splits = std::min(splits, 65535);
    }
  }
  int kchunk = (d.K + splits - 1) / splits;
  kchunk = ((kchunk + g.BK - 1) / g.BK) * g.BK;
  splits = (d.K + kchunk - 1) / kchunk;
  use_partial = splits > 1 || need_partial_for_batch;

  KArgs<T> ka;
  ka.d = d;
  ka.partial = use_partial ? (T*)ws.ptr : nullptr;
  ka.tiles_m = tiles_m;
  ka.tiles_n = tiles_n;
  ka.tiles_pad = tiles >= 32 ? (int)(((tiles + 7) / 8) * 8) : (int)tiles;
  ka.splits = splits;
  ka.kchunk = kchunk;
  if (ka.tiles_pad <= 0 || splits > 65535) return hipErrorInvalidValue;
  dim3 grid(ka.tiles_pad, nbatch, splits);
  hipError_t e = launch_any<T>(best, ka, AK, BNM, grid, stream);
  if (e != hipSuccess) return e;
  if (use_partial) {
    const long MN = (long)d.M * d.N;
    const int nout_b = d.reduce_batch ? 1 : nbatch;
    const long slices = d.reduce_batch ? (long)splits * nbatch : splits;
    if (slices >= 64 && MN <= (1L << 20) && nout_b <= 65535) {
      hipLaunchKernelGGL(gemm_reduce_wide_kernel<T>, dim3((unsigned)MN, nout_b), dim3(256), 0, stream, d,
                         (const T*)ws.ptr, splits);
    } else {
      long nout = MN * nout_b;
      int blocks = (int)std::min<long>((nout + 255) / 256, 4096);
      hipLaunchKernelGGL(gemm_reduce_kernel<T>, dim3(blocks), dim3(256), 0, stream, d, (const T*)ws.ptr, splits);
    }
    e = hipGetLastError();
  }
  if (info) *info = GemmLaunchInfo{tiles_m, tiles_n, nbatch, splits, best};
  return e;
}

template hipError_t gemm<double>(const GemmDesc<double>&, const GemmWorkspace&, hipStream_t, GemmLaunchInfo*, int, int);
template hipError_t gemm<float>(const GemmDesc<float>&, const GemmWorkspace&, hipStream_t, GemmLaunchInfo*, int, int);

}  // namespace ccsd
