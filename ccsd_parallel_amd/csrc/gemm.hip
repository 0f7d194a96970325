// MFMA GEMM engine (gfx950).  See gemm.hpp for the contract.
//
// Design (CDNA4-first, not a CUDA warp tiling):
//  * 64-wide wavefronts, each computing a (BM/WM) x (BN/WN) sub-tile as a grid
//    of 16x16 MFMA tiles: v_mfma_f64_16x16x4f64 (FP64) or v_mfma_f32_16x16x4f32.
//  * A and B tiles are staged through LDS ([m][k] with a padded pitch for a
//    k-contiguous operand, [k][m] otherwise) so that a wave's operand fetch
//    (lane l reads row l%16 of k-slice l/16) puts its 16 rows in distinct
//    banks; LDS is double-buffered with one barrier per k-tile, and the MFMA
//    operands are double-buffered in registers (k-step s+1 read during k-step
//    s).
//  * Global loads are coalesced along whichever operand stride is unit
//    (template switches AK / BNM select the loop order); any strides work.
//    Where the strides allow, each thread loads 16-byte vectors (two doubles
//    or four floats; FP32 falls back to 8-byte pairs); rows/columns past M/N
//    are clamped, not masked, and only the ragged last k-tile is
//    bounds-checked.
//  * Tiles are compiled for two or more resident workgroups per CU (4-wave
//    tiles at 256 registers, 8-wave FP64 tiles at 128), so one workgroup's
//    barrier wait is covered by another's MFMAs.
//  * The tiles past the last full round of the GPU run as k-chunks whose
//    partials are reduced afterwards (split tail).
//  * Workgroup ids are remapped so that each XCD (the dispatcher deals
//    workgroups round-robin over the 8 XCDs, each with its own L2) runs a
//    contiguous, balanced range of tiles that share operand panels.
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
  // Split tail (tail_split > 0; one batch, no split-K): workgroup x runs on
  // XCD x % 8; each XCD first takes xcd_full whole tiles of the contiguous
  // range [x * xcd_full, (x + 1) * xcd_full), then units of the tail tiles
  // [8 * xcd_full, tiles), each cut into tail_split k-chunks of tail_kchunk,
  // dealt round-robin over the XCDs.  A tail unit writes its BM x BN partial
  // to tail_partial[unit]; a second kernel sums them in a fixed order.
  int tail_split, xcd_full, tail_units, tail_kchunk;
  T* tail_partial;
  // Tile walk: 0 for column-major over the tiles_m x tiles_n grid; g > 0 walks
  // bands of g tile rows, column-major inside a band.
  int group;
};

// Tile id -> (tile row, tile column).  In bands of `group` tile rows a run of
// consecutive ids (one XCD's slots at a time) covers a window about
// group x (run / group) tiles instead of one tall column strip, so each
// A and B panel it reads is shared by several resident workgroups
// (12000^3 in 128 x 128 tiles: 3% faster with bands of 8 rows, measured).
__device__ __forceinline__ void tile_coords(int tile, int tiles_m, int tiles_n, int group, int& tm, int& tn) {
  if (group <= 0) {
    tm = tile % tiles_m;
    tn = tile / tiles_m;
    return;
  }
  const int band = group * tiles_n, first = (tile / band) * group;
  const int rows = min(tiles_m - first, group), r = tile % band;
  tm = first + r % rows;
  tn = r / rows;
}

// VWT: elements per vector load (2: 16-byte double pairs or 8-byte float
// pairs, 4: 16-byte float quads), 0 for single-element loads (in units of 2).
template <typename T, int BM, int BN, int BK, int WM, int WN, bool AK, bool BNM, int WPE, int VWT>
__global__ __launch_bounds__(64 * WM * WN, WPE) void gemm_kernel(KArgs<T> p) {
  constexpr bool VEC = VWT > 0;
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = BM / WM / 16;
  constexpr int TN = BN / WN / 16;
  static_assert(TM >= 1 && TN >= 1 && TM * WM * 16 == BM && TN * WN * 16 == BN, "wave tile");
  // LDS layouts.  A k-contiguous operand stays k-contiguous ([m][LDK]): with
  // LDK = BK + 2 doubles the 16 rows of one MFMA operand fetch fall in distinct
  // banks.  An m/n-contiguous operand is stored [k][LDM] with LDM = 16 mod 32
  // elements (BM + 16 or BM + 32): the k-rows of one fetch then start 16 or 32
  // banks apart, which gfx950 serves without conflicts for both 8-byte and
  // 4-byte reads (measured: pitches 80 and 144 conflict-free; 64, 96, 128 and
  // 160 two-way conflicted).  Every global load brings a vector of VW
  // elements that are adjacent in LDS too (FP32 pitches are multiples of 4,
  // keeping 16-byte stores aligned; FP32 LDK = BK + 4).
  constexpr int LDK = BK + (sizeof(T) == 8 ? 2 : 4);
  constexpr int LDMA = BM + ((BM & 31) ? 32 : 16);
  constexpr int LDMB = BN + ((BN & 31) ? 32 : 16);
  static_assert(LDMA % 32 == 16 && LDMB % 32 == 16, "m/n pitch");
  constexpr int ASZ = AK ? BM * LDK : BK * LDMA;
  constexpr int BSZ = BNM ? BK * LDMB : BN * LDK;
  __shared__ __attribute__((aligned(16))) T As[2][ASZ];
  __shared__ __attribute__((aligned(16))) T Bs[2][BSZ];
  // load unit: VW consecutive elements, one vector load (VEC) and one LDS store
  constexpr int VW = VEC ? VWT : 2;
  using pair_t = T __attribute__((ext_vector_type(VW)));

  const GemmDesc<T>& d = p.d;
  // XCD-aware tile id: hardware places linear workgroup x on XCD x % 8, so
  // XCD j takes the contiguous tile range [j*T/8, (j+1)*T/8) of the T tiles
  // (6 or 7 tiles each for T = 49: no XCD is left with only padding).  With
  // few tiles (tiles_pad not a multiple of 8) the plain order spreads the
  // batches and K-splits over all XCDs instead of stacking every real tile on
  // the low ones.
  const int ntiles = p.tiles_m * p.tiles_n;
  const int batch = blockIdx.y;
  const int split = blockIdx.z;
  int tile = (int)blockIdx.x;
  int kbeg = split * p.kchunk;
  int kend = min(d.K, kbeg + p.kchunk);
  T* tail_out = nullptr;  // this unit's compact partial (split tail tiles)
  if (p.tail_split > 0) {
    const int x = (int)(blockIdx.x & 7), r = (int)(blockIdx.x >> 3);
    if (r < p.xcd_full) {
      tile = x * p.xcd_full + r;
    } else {
      const int v = (r - p.xcd_full) * 8 + x;
      if (v >= p.tail_units) return;
      tile = 8 * p.xcd_full + v / p.tail_split;
      kbeg = (v % p.tail_split) * p.tail_kchunk;
      kend = min(d.K, kbeg + p.tail_kchunk);
      tail_out = p.tail_partial + (long)v * (BM * BN);
    }
  } else if ((p.tiles_pad & 7) == 0) {
    const int x = (int)(blockIdx.x & 7);
    tile = (int)(((long)x * ntiles) >> 3) + (int)(blockIdx.x >> 3);
    if (tile >= (int)(((long)(x + 1) * ntiles) >> 3)) return;
  }
  if (tile >= ntiles) return;
  int tm, tn;
  tile_coords(tile, p.tiles_m, p.tiles_n, p.group, tm, tn);
  const int b1 = batch / d.nb2, b2 = batch % d.nb2;

  const T* __restrict__ A = d.A + b1 * d.sA1 + b2 * d.sA2;
  const T* __restrict__ B = d.B + b1 * d.sB1 + b2 * d.sB2;
  const int m0 = tm * BM, n0 = tn * BN;
  const int M = d.M, N = d.N;
  const long sAm = d.sAm, sAk = d.sAk, sBk = d.sBk, sBn = d.sBn;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wy = wave / WN, wx = wave % WN;
  const int wm0 = wy * (BM / WM), wn0 = wx * (BN / WN);

  static_assert(BK % VW == 0 && BM % VW == 0 && BN % VW == 0, "vector width");
  constexpr int AU = BM * BK / VW;
  constexpr int BU = BK * BN / VW;
  constexpr int AIT = (AU + NT - 1) / NT;
  constexpr int BIT = (BU + NT - 1) / NT;
  pair_t ra[AIT], rb[BIT];

  // Unit u of a tile -> (row, k) of its VW-element vector.  A vector runs
  // along the operand's contiguous index: k for AK / !BNM, m for !AK, n for BNM.
  auto a_unit = [](int u, int& m, int& kk) {
    if (AK) { m = u / (BK / VW); kk = (u % (BK / VW)) * VW; }
    else    { kk = u / (BM / VW); m = (u % (BM / VW)) * VW; }
  };
  auto b_unit = [](int u, int& n, int& kk) {
    if (BNM) { kk = u / (BN / VW); n = (u % (BN / VW)) * VW; }
    else     { n = u / (BK / VW); kk = (u % (BK / VW)) * VW; }
  };
  // Rows past M and columns past N only feed outputs that the epilogue drops,
  // so their loads are clamped into range rather than masked (a mask would
  // have to wait for the load).  Only k is checked, and only in the ragged
  // last tile: full tiles (`checked` false) load without any test.
  // VEC: one load per VW-element vector (the host guarantees unit stride
  // along the vector, pitches that are multiples of VW and a base aligned to
  // the vector; M or N a multiple of VW when the vector runs along m or n).
  // Paired loads of an operand whose k stride is not the unit one: each
  // thread's pair addresses at k0 = 0 are computed once, and a tile adds the
  // uniform k0 * pitch (scalar arithmetic) instead of a per-lane 64-bit
  // multiply.  k-contiguous operands address rows directly (row + k0 + kk).
  const T* a_base[AK ? 1 : AIT];
  const T* b_base[BNM ? BIT : 1];
  if constexpr (VEC && !AK) {
#pragma unroll
    for (int it = 0; it < AIT; ++it) {
      int m, kk;
      a_unit(min(it * NT + tid, AU - 1), m, kk);
      a_base[it] = A + (long)kk * sAk + min(m0 + m, M - VW);
    }
  }
  if constexpr (VEC && BNM) {
#pragma unroll
    for (int it = 0; it < BIT; ++it) {
      int n, kk;
      b_unit(min(it * NT + tid, BU - 1), n, kk);
      b_base[it] = B + (long)kk * sBk + min(n0 + n, N - VW);
    }
  }
  auto load_tiles = [&](int k0, bool checked) {
    if (VEC && !checked) {
#pragma unroll
      for (int it = 0; it < AIT; ++it) {
        const int u = it * NT + tid;
        if (AU % NT != 0 && u >= AU) continue;
        int m, kk;
        a_unit(u, m, kk);
        const T* src;
        if constexpr (AK) src = A + (long)min(m0 + m, M - 1) * sAm + (k0 + kk);
        else src = a_base[it] + (long)k0 * sAk;
        ra[it] = *reinterpret_cast<const pair_t*>(src);
      }
#pragma unroll
      for (int it = 0; it < BIT; ++it) {
        const int u = it * NT + tid;
        if (BU % NT != 0 && u >= BU) continue;
        int n, kk;
        b_unit(u, n, kk);
        const T* src;
        if constexpr (BNM) src = b_base[it] + (long)k0 * sBk;
        else src = B + (long)min(n0 + n, N - 1) * sBn + (k0 + kk);
        rb[it] = *reinterpret_cast<const pair_t*>(src);
      }
      return;
    }
#pragma unroll
    for (int it = 0; it < AIT; ++it) {
      const int u = it * NT + tid;
      int m, kk;
      a_unit(u, m, kk);
      const int gm = m0 + m, gk = k0 + kk;
      pair_t v;
#pragma unroll
      for (int e = 0; e < VW; ++e) v[e] = T(0);
      if (AU % NT == 0 || u < AU) {
        if (AK) {
          const T* row = A + (long)min(gm, M - 1) * sAm;
#pragma unroll
          for (int e = 0; e < VW; ++e)
            if (!checked || gk + e < kend) v[e] = row[(long)(gk + e) * sAk];
        } else if (!checked || gk < kend) {
          const T* kr = A + (long)gk * sAk;
#pragma unroll
          for (int e = 0; e < VW; ++e) v[e] = kr[min(gm + e, M - 1) * sAm];
        }
      }
      ra[it] = v;
    }
#pragma unroll
    for (int it = 0; it < BIT; ++it) {
      const int u = it * NT + tid;
      int n, kk;
      b_unit(u, n, kk);
      const int gn = n0 + n, gk = k0 + kk;
      pair_t v;
#pragma unroll
      for (int e = 0; e < VW; ++e) v[e] = T(0);
      if (BU % NT == 0 || u < BU) {
        if (!BNM) {
          const T* col = B + (long)min(gn, N - 1) * sBn;
#pragma unroll
          for (int e = 0; e < VW; ++e)
            if (!checked || gk + e < kend) v[e] = col[(long)(gk + e) * sBk];
        } else if (!checked || gk < kend) {
          const T* kr = B + (long)gk * sBk;
#pragma unroll
          for (int e = 0; e < VW; ++e) v[e] = kr[min(gn + e, N - 1) * sBn];
        }
      }
      rb[it] = v;
    }
  };
  auto store_tiles = [&](int buf) {
#pragma unroll
    for (int it = 0; it < AIT; ++it) {
      const int u = it * NT + tid;
      if (AU % NT != 0 && u >= AU) break;
      int m, kk;
      a_unit(u, m, kk);
      *reinterpret_cast<pair_t*>(&As[buf][AK ? m * LDK + kk : kk * LDMA + m]) = ra[it];
    }
#pragma unroll
    for (int it = 0; it < BIT; ++it) {
      const int u = it * NT + tid;
      if (BU % NT != 0 && u >= BU) break;
      int n, kk;
      b_unit(u, n, kk);
      *reinterpret_cast<pair_t*>(&Bs[buf][BNM ? kk * LDMB + n : n * LDK + kk]) = rb[it];
    }
  };

  using acc_t = typename Mfma<T>::acc_t;
  acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  // One k-step (4 deep) of MFMA operands from LDS buffer `cur`, and its MFMAs.
  struct Frag { T a[TM], b[TN]; };
  auto frag = [&](int cur, int ks, Frag& f) {
    const int kr = ks + (lane >> 4);
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int m = wm0 + i * 16 + (lane & 15);
      f.a[i] = As[cur][AK ? m * LDK + kr : kr * LDMA + m];
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int n = wn0 + j * 16 + (lane & 15);
      f.b[j] = Bs[cur][BNM ? kr * LDMB + n : n * LDK + kr];
    }
  };
  auto mma = [&](const Frag& f) {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = Mfma<T>::mma(f.a[i], f.b[j], acc[i][j]);
  };
  auto compute = [&](int cur) {
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      Frag f;
      frag(cur, ks, f);
      mma(f);
    }
  };

  // Full k-tiles run in the double-buffered loop (the next tiles' loads in
  // flight during this tile's MFMAs, one barrier per tile); the ragged tail,
  // if any, follows as one checked tile.  Keeping the checked loads out of the
  // loop keeps each loaded pair in its own registers until its LDS store.
  // Paired loads: each iteration first stores tile kt+1 (loaded during the
  // previous iteration) into the idle buffer, released by the previous
  // barrier, then loads tile kt+2 into the same registers and computes tile
  // kt, so the LDS stores overlap the MFMAs.  Loads past the last full tile
  // reload it (harmless: that buffer is not read again before the next
  // barrier).  Element-wise loads hold twice the address registers and keep
  // the load / compute / store order, which fits without spilling.
  // The MFMA operands are double-buffered in registers as well: the LDS reads
  // of k-step s+1 are issued before the MFMAs of k-step s.  The tile's one
  // barrier sits before its last k-step, after every read of buffer `cur`
  // and every store of tile kt+1: behind it the first operands of tile kt+1
  // are read while the last MFMAs of tile kt run.
  const int nfull = (kend > kbeg) ? (kend - kbeg) / BK : 0;
  if constexpr (VEC) {
    constexpr int S = BK / 4;
    static_assert(S % 2 == 0, "operand sets alternate per k-step");
    Frag f[2];
    if (nfull > 0) {
      load_tiles(kbeg, false);
      store_tiles(0);
      load_tiles(kbeg + min(1, nfull - 1) * BK, false);
      __syncthreads();
      frag(0, 0, f[0]);
    }
    for (int kt = 0; kt < nfull; ++kt) {
      const int cur = kt & 1;
      store_tiles(cur ^ 1);
      // keep the loads below the stores (one register set)
      __builtin_amdgcn_sched_barrier(0);
      load_tiles(kbeg + min(kt + 2, nfull - 1) * BK, false);
#pragma unroll
      for (int st = 0; st < S; ++st) {
        if (st + 1 < S) {
          frag(cur, (st + 1) * 4, f[(st + 1) & 1]);
        } else {
          __syncthreads();
          // after the last tile this reads the idle buffer's reloaded copy of
          // it: in bounds, and never used
          frag(cur ^ 1, 0, f[0]);
        }
        mma(f[st & 1]);
      }
    }
    // the tail below stores into buffer 0, which the reads just above may
    // still be using
    __syncthreads();
  } else {
    if (nfull > 0) {
      load_tiles(kbeg, false);
      store_tiles(0);
      __syncthreads();
    }
    for (int kt = 0; kt < nfull; ++kt) {
      const int cur = kt & 1;
      const bool more = (kt + 1) < nfull;
      if (more) load_tiles(kbeg + (kt + 1) * BK, false);
      compute(cur);
      if (more) store_tiles(cur ^ 1);
      __syncthreads();
    }
  }
  if (kbeg + nfull * BK < kend) {
    load_tiles(kbeg + nfull * BK, true);
    store_tiles(0);
    __syncthreads();
    compute(0);
  }

  // Epilogue (output row of each accumulator element from the MFMA trait).
  if (tail_out) {
    // whole BM x BN tile in tile coordinates; the reduction drops rows and
    // columns past M and N
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = wm0 + i * 16 + Mfma<T>::out_row(lane, r);
          const int col = wn0 + j * 16 + (lane & 15);
          tail_out[row * BN + col] = acc[i][j][r];
        }
  } else if (p.partial) {
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

// Many partials per output (long split-K on a small C): a workgroup owns 64
// consecutive output elements; each of its 4 waves sums every 4th slice with
// one contiguous 512-byte load per slice (4 independent chains, so several
// loads are in flight), then the 4 waves' sums are added in a fixed order.
template <typename T>
__global__ __launch_bounds__(256) void gemm_reduce_wide_kernel(GemmDesc<T> d, const T* __restrict__ W, int splits) {
  const long MN = (long)d.M * d.N;
  const int nbatch = d.nb1 * d.nb2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long e = (long)blockIdx.x * 64 + lane;  // element within one output matrix
  const int bo = blockIdx.y;                    // output batch (0 when reduce_batch)
  const int nslices = d.reduce_batch ? splits * nbatch : splits;
  auto at = [&](int q) { return W[(d.reduce_batch ? (long)q : (long)q * nbatch + bo) * MN + e]; };
  T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  if (e < MN) {
    int q = wave;
    for (; q + 12 < nslices; q += 16) {
      s0 += at(q);
      s1 += at(q + 4);
      s2 += at(q + 8);
      s3 += at(q + 12);
    }
    for (; q < nslices; q += 4) s0 += at(q);
  }
  __shared__ T sh[4][64];
  sh[wave][lane] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (wave == 0 && e < MN) {
    const T v = (sh[0][lane] + sh[1][lane]) + (sh[2][lane] + sh[3][lane]);
    const int row = (int)(e / d.N), col = (int)(e % d.N);
    long off = row * d.sCm + col * d.sCn;
    if (!d.reduce_batch) off += (bo / d.nb2) * d.sC1 + (bo % d.nb2) * d.sC2;
    write_out(d, off, v);
  }
}

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

// Split tail tiles: C(tile) = alpha * sum_s W[tile][s] + beta * C, the splits
// summed in order; W holds whole bm x bn tiles, rows and columns past M and N
// are dropped here.
template <typename T>
__global__ void gemm_tail_reduce_kernel(GemmDesc<T> d, const T* __restrict__ W, int first_tile, int tiles_m,
                                        int tiles_n, int group, int ntail, int nsplit, int bm, int bn) {
  const long per = (long)bm * bn;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < ntail * per;
       idx += (long)gridDim.x * blockDim.x) {
    const long tt = idx / per, e = idx % per;
    int tm, tn;
    tile_coords(first_tile + (int)tt, tiles_m, tiles_n, group, tm, tn);
    const int row = tm * bm + (int)(e / bn), col = tn * bn + (int)(e % bn);
    if (row >= d.M || col >= d.N) continue;
    T s = 0;
    for (int sp = 0; sp < nsplit; ++sp) s += W[(tt * nsplit + sp) * per + e];
    write_out(d, (long)row * d.sCm + (long)col * d.sCn, s);
  }
}

// gemv over an m-contiguous A (sAm == 1): workgroup (x, y) owns 64
// consecutive outputs and the k-range y of gridDim.y equal ranges; its 4 waves
// take every 4th k of the range (one coalesced 64-element read of A per k, two
// independent sums), and the 4 wave sums are added in a fixed order into
// partial[y][m] (or, with one range, straight into C).
template <typename T>
__global__ __launch_bounds__(256) void gemv_t_kernel(GemmDesc<T> d, T* __restrict__ partial) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long m = (long)blockIdx.x * 64 + lane;
  const long chunk = (d.K + gridDim.y - 1) / gridDim.y;
  const long k0 = blockIdx.y * chunk, k1 = min((long)d.K, k0 + chunk);
  T s0 = 0, s1 = 0;
  if (m < d.M) {
    long k = k0 + wave;
    for (; k + 4 < k1; k += 8) {
      s0 += d.A[m + k * d.sAk] * d.B[k * d.sBk];
      s1 += d.A[m + (k + 4) * d.sAk] * d.B[(k + 4) * d.sBk];
    }
    for (; k < k1; k += 4) s0 += d.A[m + k * d.sAk] * d.B[k * d.sBk];
  }
  __shared__ T sh[4][64];
  sh[wave][lane] = s0 + s1;
  __syncthreads();
  if (wave == 0 && m < d.M) {
    const T v = (sh[0][lane] + sh[1][lane]) + (sh[2][lane] + sh[3][lane]);
    if (gridDim.y == 1) write_out(d, m * d.sCm, v);
    else partial[blockIdx.y * (long)d.M + m] = v;
  }
}

// C[m] from gemv_t_kernel's k-range partials, summed in range order.
template <typename T>
__global__ void gemv_t_reduce_kernel(GemmDesc<T> d, const T* __restrict__ partial, int nsplit) {
  for (long m = blockIdx.x * (long)blockDim.x + threadIdx.x; m < d.M; m += (long)gridDim.x * blockDim.x) {
    T s = 0;
    for (int y = 0; y < nsplit; ++y) s += partial[y * (long)d.M + m];
    write_out(d, m * d.sCm, s);
  }
}

// gemv over unit-stride rows with 16-byte alignment (the host checks): lane l
// loads pairs k0 + 128 j + 2 l, j = 0..3, into four independent sums, so a
// wavefront moves 2 KB of the row per iteration; the ragged end is read
// element-wise.  The sums are combined in a fixed order.
template <typename T>
__global__ void gemv_pairs_kernel(GemmDesc<T> d) {
  using pair_t = T __attribute__((ext_vector_type(2)));
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (row >= d.M) return;
  const T* a = d.A + (long)row * d.sAm;
  const T* x = d.B;
  T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  int k0 = 0;
  for (; k0 + 512 <= d.K; k0 += 512) {
    const int k = k0 + 2 * lane;
    const pair_t a0 = *(const pair_t*)(a + k), a1 = *(const pair_t*)(a + k + 128);
    const pair_t a2 = *(const pair_t*)(a + k + 256), a3 = *(const pair_t*)(a + k + 384);
    const pair_t x0 = *(const pair_t*)(x + k), x1 = *(const pair_t*)(x + k + 128);
    const pair_t x2 = *(const pair_t*)(x + k + 256), x3 = *(const pair_t*)(x + k + 384);
    s0 += a0[0] * x0[0] + a0[1] * x0[1];
    s1 += a1[0] * x1[0] + a1[1] * x1[1];
    s2 += a2[0] * x2[0] + a2[1] * x2[1];
    s3 += a3[0] * x3[0] + a3[1] * x3[1];
  }
  for (int k = k0 + lane; k < d.K; k += 64) s0 += a[k] * x[k];
  T s = (s0 + s1) + (s2 + s3);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) write_out(d, (long)row * d.sCm, s);
}

// One tile configuration: BM x BN x BK on WM x WN waves.  slots: workgroups
// resident per CU (LDS and register limited; the kernel is compiled for at
// least `slots` waves per SIMD when it has 4 waves, so that two or more
// workgroups overlap each other's barriers).  rate: FP64 TF/s with every slot
// busy and paired loads (12000^3 GEMM on MI355X, tile-sweep probe, divided
// by that shape's round fill as modelled below).
// pairs_only / k_contig_b_only: variants that exceed their register budget
// (spill) with element loads / with an n-contiguous B are left out of the
// automatic choice (forced configurations still run them, correctly).
// f64_only: measured slower than the 4-wave tiles in FP32 (the FP32 MFMA takes
// half the cycles, so the extra LDS reads per MFMA of the smaller wave tiles
// show), left out of the FP32 choice.
struct Cfg {
  int BM, BN, BK, WM, WN, slots;
  float rate;
  bool pairs_only = false, k_contig_b_only = false, f64_only = false;
};
constexpr Cfg kCfgs[] = {
    {128, 256, 16, 2, 4, 1, 64.9f},
    {256, 128, 16, 4, 2, 1, 65.0f},
    {128, 128, 16, 2, 2, 2, 68.9f},
    {64, 128, 16, 2, 2, 2, 63.2f},
    {128, 64, 16, 2, 2, 2, 62.4f},
    {64, 64, 16, 2, 2, 4, 60.7f},
    {32, 64, 16, 1, 2, 4, 44.7f},
    {64, 32, 16, 2, 1, 4, 40.2f},
    // 48-wide tiles for the O = 40 edges (M or N of 40 fills 83% instead of 62%)
    {128, 48, 16, 4, 1, 2, 55.4f},
    {48, 128, 16, 1, 4, 2, 62.0f},
    {48, 64, 16, 1, 2, 4, 50.0f},
    {64, 48, 16, 2, 1, 4, 47.0f},
    // 112 = 7 x 16 rows: P2(40) = 780 fills 99.5% (128-row tiles: 87%)
    {112, 128, 16, 1, 4, 2, 65.2f},
    {128, 112, 16, 4, 1, 2, 64.8f},
    // 8 waves, two workgroups per CU: 4 waves per SIMD (128 registers each)
    // cover each other's LDS and barrier waits (FP64 12000^3: 69-70 TF/s
    // against 67-68 for the 4-wave 128 x 128; FP64 iteration 183.5 -> 180.2
    // ms, same-box A/B)
    {128, 128, 16, 4, 2, 2, 72.6f, true, false, true},
    {128, 128, 16, 2, 4, 2, 72.4f, true, false, true},
    // paired loads of an n-contiguous B need more than its 128 registers
    // (spills), so it serves k-contiguous B only (W_mbij: 61.1 TF/s)
    {112, 128, 16, 1, 8, 2, 67.0f, true, true, true},
};
constexpr int kNumCus = 256;
constexpr int kTileBand = 8;  // tile rows per band of the tile walk (tile_coords)
constexpr int kNumCfgs = sizeof(kCfgs) / sizeof(kCfgs[0]);

template <typename T, int BM, int BN, int BK, int WM, int WN, int WPE, int VWT>
hipError_t launch_layout(const KArgs<T>& ka, bool AK, bool BNM, dim3 grid, hipStream_t s) {
  constexpr int NT = 64 * WM * WN;
  if (AK && BNM)
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WM, WN, true, true, WPE, VWT>), grid, dim3(NT), 0, s, ka);
  else if (AK && !BNM)
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WM, WN, true, false, WPE, VWT>), grid, dim3(NT), 0, s, ka);
  else if (!AK && BNM)
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WM, WN, false, true, WPE, VWT>), grid, dim3(NT), 0, s, ka);
  else
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WM, WN, false, false, WPE, VWT>), grid, dim3(NT), 0, s, ka);
  return hipGetLastError();
}

template <typename T, int C>
hipError_t launch_cfg(const KArgs<T>& ka, bool AK, bool BNM, int vw, dim3 grid, hipStream_t s) {
  constexpr Cfg g = kCfgs[C];
  // waves-per-SIMD floor handed to the compiler: a CU's 4 SIMDs hold slots x
  // (WM x WN) waves, which must fit 512 registers per SIMD lane between them
  constexpr int wpe = (g.WM * g.WN == 4)   ? g.slots
                      : (g.WM * g.WN == 8) ? 2 * g.slots
                                           : 1;
  if constexpr (sizeof(T) == 4) {
    if (vw == 4) return launch_layout<T, g.BM, g.BN, g.BK, g.WM, g.WN, wpe, 4>(ka, AK, BNM, grid, s);
  }
  if (vw == 2) return launch_layout<T, g.BM, g.BN, g.BK, g.WM, g.WN, wpe, 2>(ka, AK, BNM, grid, s);
  return launch_layout<T, g.BM, g.BN, g.BK, g.WM, g.WN, wpe, 0>(ka, AK, BNM, grid, s);
}

template <typename T>
hipError_t launch_any(int c, const KArgs<T>& ka, bool AK, bool BNM, int vec, dim3 grid, hipStream_t s) {
  switch (c) {
    case 0: return launch_cfg<T, 0>(ka, AK, BNM, vec, grid, s);
    case 1: return launch_cfg<T, 1>(ka, AK, BNM, vec, grid, s);
    case 2: return launch_cfg<T, 2>(ka, AK, BNM, vec, grid, s);
    case 3: return launch_cfg<T, 3>(ka, AK, BNM, vec, grid, s);
    case 4: return launch_cfg<T, 4>(ka, AK, BNM, vec, grid, s);
    case 5: return launch_cfg<T, 5>(ka, AK, BNM, vec, grid, s);
    case 6: return launch_cfg<T, 6>(ka, AK, BNM, vec, grid, s);
    case 7: return launch_cfg<T, 7>(ka, AK, BNM, vec, grid, s);
    case 8: return launch_cfg<T, 8>(ka, AK, BNM, vec, grid, s);
    case 9: return launch_cfg<T, 9>(ka, AK, BNM, vec, grid, s);
    case 10: return launch_cfg<T, 10>(ka, AK, BNM, vec, grid, s);
    case 11: return launch_cfg<T, 11>(ka, AK, BNM, vec, grid, s);
    case 12: return launch_cfg<T, 12>(ka, AK, BNM, vec, grid, s);
    case 13: return launch_cfg<T, 13>(ka, AK, BNM, vec, grid, s);
    case 14: return launch_cfg<T, 14>(ka, AK, BNM, vec, grid, s);
    case 15: return launch_cfg<T, 15>(ka, AK, BNM, vec, grid, s);
    default: return launch_cfg<T, 16>(ka, AK, BNM, vec, grid, s);
  }
}
static_assert(kNumCfgs == 17, "launch_any covers every configuration");

// W-element vector loads are legal when each operand has unit stride along
// its vector index, every other stride and batch stride is a multiple of W,
// the base is aligned to a vector, and a vector along m or n cannot straddle
// M or N.
template <typename T>
bool vectors_ok(const GemmDesc<T>& d, bool AK, bool BNM, long W) {
  auto mult = [W](long x) { return x % W == 0; };
  auto aligned = [W](const void* p) { return (reinterpret_cast<uintptr_t>(p) % (W * sizeof(T))) == 0; };
  const bool a = AK ? (d.sAk == 1 && mult(d.sAm)) : (d.sAm == 1 && mult(d.sAk) && mult(d.M));
  const bool b = BNM ? (d.sBn == 1 && mult(d.sBk) && mult(d.N)) : (d.sBk == 1 && mult(d.sBn));
  return a && b && mult(d.sA1) && mult(d.sA2) && mult(d.sB1) && mult(d.sB2) && aligned(d.A) && aligned(d.B);
}

}  // namespace

template <typename T>
hipError_t gemm(const GemmDesc<T>& d, const GemmWorkspace& ws, hipStream_t stream,
                GemmLaunchInfo* info, int force_config, int force_split, int force_pairs) {
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
    if (d.sAm == 1 && d.sAk != 1 && d.M >= 64) {
      // k-ranges of at least 256 so that about 1024 workgroups are in flight
      const long mb = (d.M + 63) / 64;
      long ny = std::min<long>(std::max<long>(1, 1024 / mb), std::max<long>(1, d.K / 256));
      ny = std::min<long>(ny, (long)(ws.bytes / sizeof(T) / std::max(1, d.M)));
      if (ny < 1 || !ws.ptr) ny = 1;
      hipLaunchKernelGGL(gemv_t_kernel<T>, dim3((unsigned)mb, (unsigned)ny), dim3(256), 0, stream, d, (T*)ws.ptr);
      if (ny > 1)
        hipLaunchKernelGGL(gemv_t_reduce_kernel<T>, dim3((unsigned)std::min<long>((d.M + 255) / 256, 4096)),
                           dim3(256), 0, stream, d, (const T*)ws.ptr, (int)ny);
      if (info) *info = GemmLaunchInfo{d.M, 1, 1, (int)ny, -1, 0};
      return hipGetLastError();
    }
    const int rows_per_block = 4;
    const bool pairs = d.sAk == 1 && d.sBk == 1 && d.sAm % 2 == 0 && d.K <= (1 << 30) &&
                       reinterpret_cast<uintptr_t>(d.A) % 16 == 0 && reinterpret_cast<uintptr_t>(d.B) % 16 == 0;
    if (pairs && force_pairs != 0)
      hipLaunchKernelGGL(gemv_pairs_kernel<T>, dim3((d.M + rows_per_block - 1) / rows_per_block),
                         dim3(64 * rows_per_block), 0, stream, d);
    else
      hipLaunchKernelGGL(gemv_kernel<T>, dim3((d.M + rows_per_block - 1) / rows_per_block),
                         dim3(64 * rows_per_block), 0, stream, d);
    if (info) *info = GemmLaunchInfo{d.M, 1, 1, 1, -1, 0};
    return hipGetLastError();
  }
  // A is "k-contiguous" unless its m stride is the unit one (and likewise B).
  const bool AK = !(d.sAm == 1 && d.sAk != 1);
  const bool BNM = !(d.sBk == 1 && d.sBn != 1);
  // widest legal vector: float quads (16 bytes), else pairs, else elements
  const int vw = force_pairs == 0 ? 0
                 : (sizeof(T) == 4 && vectors_ok(d, AK, BNM, 4)) ? 4
                 : vectors_ok(d, AK, BNM, 2) ? 2 : 0;
  const bool vec = vw > 0;

  // Tile shape and split-K count with the least modelled time: each of the
  // tiles x splits work units costs its padded flops at the configuration's
  // rate, units run in rounds of (CUs x slots), and a split adds one pass over
  // the partial slices.  The time is at least the HBM time of reading A once
  // per column tile and B once per row tile (two at most: neighbouring tiles
  // share panels in L2) plus the C traffic, which is what bounds skinny
  // products such as M = N = O with K = O V^2.  Splits need at least 8
  // k-tiles each and room for their partials in the workspace.
  const bool need_partial_for_batch = d.reduce_batch && nbatch > 1;
  const size_t part_bytes_per = (size_t)d.M * d.N * nbatch * sizeof(T);
  const long ws_slices = (long)(ws.bytes / std::max<size_t>(1, part_bytes_per));
  if (need_partial_for_batch && ws_slices < 1) return hipErrorOutOfMemory;
  const double flops = 2.0 * d.M * (double)d.N * d.K * nbatch;
  const double a_bytes = double(d.M) * d.K * nbatch * sizeof(T), b_bytes = double(d.K) * d.N * nbatch * sizeof(T);
  const double c_bytes = double(part_bytes_per) * (d.beta != T(0) ? 2 : 1);
  constexpr double kHbm = 6e12;  // sustained HBM bytes/s
  int best = 0, kchunk = d.K;
  double best_t = 1e300, best_tc = 1e300;
  for (int c = 0; c < kNumCfgs; ++c) {
    if (force_config >= 0 && force_config < kNumCfgs && c != force_config) continue;
    const Cfg& g = kCfgs[c];
    if (force_config < 0 && ((g.pairs_only && !vec) || (g.k_contig_b_only && BNM) || (g.f64_only && sizeof(T) != 8)))
      continue;
    const long tm = (d.M + g.BM - 1) / g.BM, tn = (d.N + g.BN - 1) / g.BN;
    const double pad = (double(tm * g.BM) * double(tn * g.BN)) / (double(d.M) * d.N);
    const long tiles = tm * tn * nbatch;
    const double slots = double(kNumCus) * g.slots;
    const bool xcd_order = tm * tn >= 32;  // the launch below pads to 8 and deals tiles per XCD
    const long ktiles = (d.K + g.BK - 1) / g.BK;
    const double t_hbm = (a_bytes * std::min(tn, 2L) + b_bytes * std::min(tm, 2L) + c_bytes) / kHbm;
    long s_lo = 1, s_hi = 1;
    if (force_split > 0) {
      s_lo = s_hi = force_split;
    } else if (ws_slices >= 2) {
      s_hi = std::max(1L, std::min({ktiles / 8, ws_slices, 1024L}));
    }
    long s_prev = -1;
    // split counts 1..8, then in steps of ~1/8 (at most ~40 candidates)
    for (long s = s_lo; s <= s_hi; s = (s < 8) ? s + 1 : s + s / 8) {
      int kc = (int)((d.K + s - 1) / s);
      kc = ((kc + g.BK - 1) / g.BK) * g.BK;
      const long sr = (d.K + kc - 1) / kc;
      if (sr == s_prev) continue;  // same k-chunking as the previous count
      s_prev = sr;
      if (sr > 1 && sr > ws_slices) continue;
      const double units = double(tiles) * sr;
      // rounds on the busiest XCD (each XCD holds 1/8 of the CUs)
      const double per_xcd = xcd_order ? std::ceil(double(tm * tn) / 8) * double(nbatch) * sr : std::ceil(units / 8);
      double rounds = std::ceil(per_xcd / (slots / 8));
      // whole-K tiles of a product with at least half a round of tiles: the
      // split tail below cuts the last partial round into k-chunks; model the
      // rounds it leaves (same rule as the launch)
      if (sr == 1 && nbatch == 1 && !d.reduce_batch && force_split <= 0 && tm * tn >= kNumCus / 2 * g.slots) {
        const long sx = (long)(kNumCus / 8) * g.slots, xf = (tm * tn / 8) / sx * sx, tl = tm * tn - 8 * xf;
        if (xf > 0 && tl > 0) {
          double br = 0.99 * rounds;
          for (int sp = 2; sp <= 8; ++sp) {
            int kq = (d.K + sp - 1) / sp;
            kq = ((kq + g.BK - 1) / g.BK) * g.BK;
            const int se = (d.K + kq - 1) / kq;
            if (se < 2 || kq / g.BK < 8) continue;
            if (double(tl) * se * g.BM * g.BN * sizeof(T) > double(ws.bytes)) continue;
            br = std::min(br, double(xf) / sx + std::ceil(std::ceil(double(tl * se) / 8) / sx) / se + 0.01 * se);
          }
          rounds = std::min(rounds, br);
        }
      }
      const double tc = rounds * slots / units * pad * flops / (g.rate * 1e12);
      double t = std::max(tc, t_hbm);
      if (sr > 1 || need_partial_for_batch) t += (sr + 2) * double(part_bytes_per) / 4e12 + 5e-6;
      // at the HBM floor, the least compute (least padding) breaks the tie
      if (t < best_t * (1 - 1e-9) || (t <= best_t * (1 + 1e-9) && tc < best_tc)) {
        best_t = t; best_tc = tc; best = c; kchunk = kc;
      }
    }
  }
  const Cfg& g = kCfgs[best];
  const int tiles_m = (d.M + g.BM - 1) / g.BM;
  const int tiles_n = (d.N + g.BN - 1) / g.BN;
  const long tiles = (long)tiles_m * tiles_n;
  const int splits = (d.K + kchunk - 1) / kchunk;
  const bool use_partial = splits > 1 || need_partial_for_batch;
  if (use_partial && (long)splits > ws_slices) return hipErrorOutOfMemory;

  // Split tail.  With whole-K tiles the last round on each XCD is partly
  // empty (12000^3 in 128 x 128 tiles: 1104.5 tiles per XCD on 64 slots, 18
  // rounds for 17.3 of work; measured 5% slower per flop than a shape of 17
  // full rounds).  The tiles past the last full round are cut into k-chunks
  // so that the tail costs a fraction of a round, then reduced in order.
  int tail_split = 0, tail_kchunk = 0;
  long xcd_full = 0, tail = 0;
  {
    const long slots_xcd = (long)(kNumCus / 8) * g.slots;
    xcd_full = (tiles / 8) / slots_xcd * slots_xcd;  // whole tiles per XCD
    tail = tiles - 8 * xcd_full;
    if (splits == 1 && nbatch == 1 && !d.reduce_batch && force_split <= 0 && tiles >= 32 && xcd_full > 0 &&
        tail > 0) {
      const double r0 = std::ceil(std::ceil(double(tiles) / 8) / slots_xcd);
      double best_r = 0.99 * r0;  // at least 1% fewer rounds
      for (int sp = 2; sp <= 8; ++sp) {
        int kc = (d.K + sp - 1) / sp;
        kc = ((kc + g.BK - 1) / g.BK) * g.BK;
        const int se = (d.K + kc - 1) / kc;
        if (se < 2 || kc / g.BK < 8) continue;
        if (double(tail) * se * g.BM * g.BN * sizeof(T) > double(ws.bytes)) continue;
        const double r = double(xcd_full) / slots_xcd +
                         std::ceil(std::ceil(double(tail * se) / 8) / slots_xcd) / se + 0.01 * se;
        if (r < best_r) { best_r = r; tail_split = se; tail_kchunk = kc; }
      }
    }
    if (tail_split == 0) xcd_full = tail = 0;
  }

  KArgs<T> ka;
  ka.d = d;
  ka.partial = use_partial ? (T*)ws.ptr : nullptr;
  ka.tiles_m = tiles_m;
  ka.tiles_n = tiles_n;
  ka.tiles_pad = tiles >= 32 ? (int)(((tiles + 7) / 8) * 8) : (int)tiles;
  ka.splits = splits;
  ka.kchunk = kchunk;
  ka.tail_split = tail_split;
  ka.xcd_full = (int)xcd_full;
  ka.tail_units = (int)(tail * tail_split);
  ka.tail_kchunk = tail_kchunk;
  ka.tail_partial = tail_split ? (T*)ws.ptr : nullptr;
  ka.group = tiles_m >= 2 * kTileBand && tiles_n >= kTileBand ? kTileBand : 0;
  if (ka.tiles_pad <= 0 || splits > 65535) return hipErrorInvalidValue;
  dim3 grid(ka.tiles_pad, nbatch, splits);
  if (tail_split) grid = dim3((unsigned)(8 * (xcd_full + (ka.tail_units + 7) / 8)), 1, 1);
  hipError_t e = launch_any<T>(best, ka, AK, BNM, vw, grid, stream);
  if (e != hipSuccess) return e;
  if (tail_split) {
    const long n = tail * g.BM * g.BN;
    const int blocks = (int)std::min<long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(gemm_tail_reduce_kernel<T>, dim3(blocks), dim3(256), 0, stream, d, (const T*)ws.ptr,
                       (int)(8 * xcd_full), tiles_m, tiles_n, ka.group, (int)tail, tail_split, g.BM, g.BN);
    e = hipGetLastError();
  }
  if (use_partial) {
    const long MN = (long)d.M * d.N;
    const int nout_b = d.reduce_batch ? 1 : nbatch;
    const long slices = d.reduce_batch ? (long)splits * nbatch : splits;
    if (slices >= 64 && MN <= (1L << 20) && nout_b <= 65535) {
      hipLaunchKernelGGL(gemm_reduce_wide_kernel<T>, dim3((unsigned)((MN + 63) / 64), nout_b), dim3(256), 0, stream,
                         d, (const T*)ws.ptr, splits);
    } else {
      long nout = MN * nout_b;
      int blocks = (int)std::min<long>((nout + 255) / 256, 4096);
      hipLaunchKernelGGL(gemm_reduce_kernel<T>, dim3(blocks), dim3(256), 0, stream, d, (const T*)ws.ptr, splits);
    }
    e = hipGetLastError();
  }
  if (info) *info = GemmLaunchInfo{tiles_m, tiles_n, nbatch, splits, best, vw, tail_split};
  return e;
}

template hipError_t gemm<double>(const GemmDesc<double>&, const GemmWorkspace&, hipStream_t, GemmLaunchInfo*, int, int,
                                 int);
template hipError_t gemm<float>(const GemmDesc<float>&, const GemmWorkspace&, hipStream_t, GemmLaunchInfo*, int, int, int);

}  // namespace ccsd
