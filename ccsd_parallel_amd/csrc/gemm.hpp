// Strided, batched MFMA GEMM engine for CDNA4 (gfx950).
//
//   C[b](m,n) = alpha * sum_k A[b](m,k) * B[b](k,n) + beta * C[b](m,n)
//
// Every CCSD contraction on the hot path (pp/hh ladders, ring, W/F builds) is
// expressed as one call of this engine over the packed antisymmetric layouts;
// the element addresses are
//   A: A + b1*sA1 + b2*sA2 + m*sAm + k*sAk   (and likewise for B, C)
// with the batch index z = b1*nb2 + b2.  When ``reduce_batch`` is set, all
// batches sum into one C (the batch strides of C are ignored).  Split-K and
// batch reductions go through a workspace and a second deterministic pass, so
// results are bit-reproducible run to run.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "gemm_desc.hpp"

namespace ccsd {



// Scratch the engine may use for split-K / batch-reduce partials.
struct GemmWorkspace {
  void* ptr = nullptr;
  size_t bytes = 0;
};

// Launch statistics of the last call (for tests and the profiler notes).
struct GemmLaunchInfo {
  int tiles_m = 0, tiles_n = 0, batches = 0, splits = 0, config = 0;
  int pairs = 0;       // elements per vector load (2 or 4), 0 for element loads
  int tail_split = 0;  // k-chunks of the tiles past the last full round (0: none)
};

// Workspace the engine's callers allocate for split-K / batch-reduce partials.
constexpr size_t kGemmWorkspaceBytes = 256ull << 20;

// force_config: tile configuration index (-1 = modelled choice); force_split:
// K splits (0 = modelled); force_pairs: 0 = element-wise loads even where
// paired loads are legal (-1 = automatic).
template <typename T>
hipError_t gemm(const GemmDesc<T>& d, const GemmWorkspace& ws, hipStream_t stream,
                GemmLaunchInfo* info = nullptr, int force_config = -1, int force_split = 0, int force_pairs = -1);

// Plain FP64 GEMV-like helper is just gemm with N=1.

}  // namespace ccsd
