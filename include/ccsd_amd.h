/* C ABI of the MI355X CCSD amplitude-update engine.
 *
 * The reference has no FFI: its boundary is the C++ class
 * CCSD_Memory_optimized (ccsd_memory_optimized.h:15-149) whose entry point
 * `double CCSD(double toler, bool print_stuff)` (ccsd_memory_optimized.cpp:16-567)
 * runs the whole amplitude iteration.  This header is the thin C layer that the
 * drop-in class (ccsd_memory_optimized.h in this directory) and any other host
 * binding call.  Plain pointers and sizes only; no exceptions cross it.
 *
 * Conventions
 *   O = occupied spin orbitals, V = virtual spin orbitals, N = O + V.
 *   Pair index for p<q: pidx(p,q) = q*(q-1)/2 + p; P2(n) = n*(n-1)/2.
 *   Integral blocks are antisymmetrised <pq||rs>, packed row-major:
 *     oooo [P2(O)][P2(O)]   <ij||kl>, i<j, k<l
 *     ooov [P2(O)][O*V]     <ij||ka>, i<j, column k*V+a
 *     oovv [P2(O)][P2(V)]   <ij||ab>, i<j, a<b
 *     ovov [O*V][O*V]       <ia||jb>, row i*V+a
 *     ovvv [O*V][P2(V)]     <ia||bc>, b<c
 *     vvvv [P2(V)][P2(V)]   <ab||cd>, a<b, c<d
 *   Orbitals are real, as in the reference: ovov must satisfy
 *   <ia||jb> = <jb||ia> (W_mbej and the T1 ring term use it).  oooo and
 *   vvvv are read as stored (the pp ladder sums tau_ij^cd <ab||cd> over row
 *   ab of vvvv), and the gradients of ccsd_backward are symmetrised over
 *   pair exchange.
 *   fock is the full N x N Fock matrix (row-major, occupied first).
 *   Amplitudes: t1[i*V+a]; t2 packed [P2(O)][P2(V)] holding t_ij^ab, i<j, a<b.
 *   Return value 0 = success, negative = error (ccsd_strerror).  Host buffers
 *   are caller-owned and copied; the context owns all device memory.
 *   One context per thread; calls on one context are serialised on its stream.
 */
#ifndef CCSD_AMD_H
#define CCSD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ccsd_ctx ccsd_ctx;

enum ccsd_status {
  CCSD_OK = 0,
  CCSD_ERR_ARG = -1,           /* null pointer, bad O/V, bad option */
  CCSD_ERR_HIP = -2,           /* HIP runtime error */
  CCSD_ERR_BUDGET = -3,        /* the plan exceeds the HBM budget (checked before allocating) */
  CCSD_ERR_NONFINITE = -4,     /* NaN/Inf in inputs or in the iteration */
  CCSD_ERR_NOT_CONVERGED = -5, /* max_iter reached; result is filled */
  CCSD_ERR_DENOMINATOR = -6,   /* |f_ii - f_aa| below 1e-10 */
  CCSD_ERR_NO_DEVICE = -7,     /* no HIP device */
  CCSD_ERR_STATE = -8          /* call order (e.g. backward before solve) */
};

enum ccsd_precision {
  CCSD_PREC_FP64 = 0,   /* all contractions in FP64 (reference mode) */
  CCSD_PREC_FP32 = 1,   /* ladder/ring/W_mbej/W_mbij GEMMs in FP32 MFMA */
  CCSD_PREC_BF16X3 = 2  /* same GEMMs as BF16 hi/lo split, 3 products, FP32 accumulate */
};

typedef struct ccsd_options {
  int precision;            /* ccsd_precision used before the switch to FP64 */
  double switch_threshold;  /* mixed modes: switch to FP64 once |dE| < max(this, switch_factor * tol, tol) (default 0) */
  int final_fp64_sweeps;    /* minimum FP64 iterations after the switch (default 3) */
  int diis_dim;             /* 0 or 1: plain Jacobi as the reference; >= 2: DIIS subspace (default 8) */
  int diis_start;           /* first iteration allowed to extrapolate (default 1) */
  int recompute_v4;         /* backward, four-virtual gradient block dE/d<ab||cd> =
                               1/2 (tau^T Lambda2 + Lambda2^T tau), P2(V)^2 doubles (16 GB at
                               O=40/V=300): 0 forms it whole in HBM (CCSD_ERR_BUDGET when it does not
                               fit in free device memory); 1 recomputes it in 512 MB row slabs through
                               two buffers, each slab streamed to the host while the next is formed;
                               < 0 (default) forms it whole when free memory holds it with 4 GB to
                               spare, else as 1.  Either way its copy to the host overlaps the rest of
                               the gradient pass. */
  int verbose;              /* print "Energi: E Steg: n Time/iter: s" lines like the reference */
  int device;               /* HIP device ordinal */
  size_t hbm_budget_bytes;  /* 0 = 90 % of the device's free memory */
  int diis_reset_on_switch; /* mixed modes: clear the DIIS subspace at the switch to FP64 (default 1) */
  double switch_factor;     /* mixed modes: switch once |dE| < switch_factor * tol (default 100: the FP64
                               sweeps then take |dE| down the last two decades) */
  double lambda_switch_factor; /* backward, BF16X3: switch the Lambda iterations to FP64 once
                                  max|dLambda| < this * lambda_tol (default 30; 0 = switch_factor), or once
                                  max|dLambda| has set no new minimum in 4 BF16x3 iterations
                                  (its rounding floor) */
  double lambda_tol;        /* backward: the Lambda iterations stop at max|dLambda| < lambda_tol;
                               0 (default) = 1e-2 x the tol of the last ccsd_solve, clamped to
                               [1e-11, 1e-8] (the gradient is then as accurate as the amplitudes it
                               is taken at) */
  int strassen_min_order;   /* FP64 and FP32 arithmetic: the square (OV)^3 products (the ring, W_mbej's
                               quadratic term and their FP64 reverse products in ccsd_backward) use one level of
                               Strassen-Winograd (seven half-order products; three (OV/2)^2 temporaries
                               and eight (OV/2)^2 operand sums counted in the plan) when OV is even and
                               at least this; 0 (default) =
                               8192, < 0 = never (classical products throughout) */
} ccsd_options;

typedef struct ccsd_problem {
  int n_occ;   /* O */
  int n_virt;  /* V */
  const double* fock;
  const double* oooo;
  const double* ooov;
  const double* oovv;
  const double* ovov;
  const double* ovvv;
  const double* vvvv;
} ccsd_problem;

typedef struct ccsd_result {
  double e_corr;          /* CCSD correlation energy */
  double e_mp2;           /* energy of the MP2 starting guess */
  double last_delta_e;    /* |E_n - E_{n-1}| at exit */
  double max_delta_t2;    /* max |t2_n - t2_{n-1}| at exit */
  int iterations;         /* amplitude updates performed */
  int converged;
  int diis_fallbacks;     /* DIIS solves rejected (plain Jacobi step taken) */
  int lowp_iterations;    /* iterations run in FP32 / BF16x3 */
  int fp64_iterations;    /* iterations run in FP64 */
  double seconds_total;   /* wall time of the iteration loop */
  double seconds_per_iter;
} ccsd_result;

typedef struct ccsd_grad_info {
  int lambda_iterations;
  int converged;
  double lambda_residual;     /* max |Lambda_n - Lambda_{n-1}| at exit */
  double seconds;
  int lambda_lowp_iterations; /* of lambda_iterations, run with BF16x3 reverse products */
} ccsd_grad_info;

/* Fill *o with defaults. */
int ccsd_default_options(ccsd_options* o);

/* Create a context from host buffers (copied into HBM). */
int ccsd_create(const ccsd_problem* p, const ccsd_options* o, ccsd_ctx** out);

/* Create a context whose integrals are generated in HBM by the counter-based
 * synthetic generator (SplitMix64 + Box-Muller, scale = standard deviation). */
int ccsd_create_synthetic(int n_occ, int n_virt, uint64_t seed, double scale, const ccsd_options* o,
                          ccsd_ctx** out);

/* Closed-shell (RHF) input in spatial orbitals, the data the reference's
 * CCSD_Memory_optimized::CCSD pulls from its Hartree_Fock_Solver
 * (ccsd_memory_optimized.cpp:49-74 and Return_Field_Q at :2987): orbital
 * energies, MO coefficients and the AO two-electron integrals. */
typedef struct ccsd_rhf_problem {
  int n_basis;        /* n: basis functions = spatial MOs */
  int n_occ;          /* doubly occupied spatial orbitals */
  int n_frozen;       /* lowest occupied spatial orbitals left out of the correlation (frozen core) */
  const double* eps;  /* canonical orbital energies [n] (the Fock matrix is diagonal) */
  const double* C;    /* MO coefficients C[mu*n + p]; NULL when eri is already in the MO basis */
  const double* eri;  /* (mu nu|la si), chemists' notation, [((mu*n + nu)*n + la)*n + si], n^4 */
} ccsd_rhf_problem;

/* Create a context from RHF data.  Replaces the reference's AO->MO step and
 * integral slicing (Prepear_AOs, ccsd_memory_optimized.cpp:2698-3522;
 * Fill_integ*_2D, :3574-4779): the integrals are copied to HBM, transformed
 * to the MO basis by four MFMA GEMMs there (when C is given), and packed into
 * the antisymmetrised spin-orbital blocks of ccsd_problem on the device
 * (spin orbital P: spatial n_frozen + P/2 or n_occ + (P-O)/2, spin P%2;
 * O = 2 (n_occ - n_frozen), V = 2 (n - n_occ)).  The transform needs 2 n^4
 * doubles of HBM during the call (CCSD_ERR_BUDGET when they are not free). */
int ccsd_create_rhf(const ccsd_rhf_problem* p, const ccsd_options* o, ccsd_ctx** out);

/* The AO->MO transform alone, on the device: mo_eri (pq|rs) =
 * sum C[mu][p] C[nu][q] C[la][r] C[si][s] (mu nu|la si), all n^4 arrays in the
 * ccsd_rhf_problem layout (host buffers). */
int ccsd_ao_to_mo(int n, const double* C, const double* ao_eri, double* mo_eri, int device);

/* Iterate from the MP2 guess (or from amplitudes set by ccsd_set_amplitudes)
 * until |dE| < tol or max_iter updates.  Returns CCSD_ERR_NOT_CONVERGED (and a
 * filled result) when max_iter is reached. */
int ccsd_solve(ccsd_ctx* c, double tol, int max_iter, ccsd_result* out);

/* Run exactly n amplitude updates in the given precision without a
 * convergence test (benchmarking); energies[k] receives E after update k. */
int ccsd_iterate(ccsd_ctx* c, int n, int precision, double* energies);

/* Reset amplitudes to the MP2 guess and clear the DIIS history. */
int ccsd_reset(ccsd_ctx* c);

/* Right-hand sides R1 (O*V) and packed R2 at the given amplitudes (FP64),
 * i.e. D*t = R without the division. */
int ccsd_residual(ccsd_ctx* c, const double* t1, const double* t2, double* r1, double* r2);

/* Energy E = f_ia t_i^a + sum_{i<j,a<b} <ij||ab> tau_ij^ab at given amplitudes. */
int ccsd_energy(ccsd_ctx* c, const double* t1, const double* t2, double* e);

/* Allocation-free variant of ccsd_energy.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Every pointer is a device pointer owned by
 * the caller: fov [O*V] (f_ia), oovv packed [P2(O)][P2(V)], t1 [O*V], t2
 * packed [P2(O)][P2(V)], e [1] (receives E); stream is a hipStream_t (NULL =
 * the null stream).  The kernels are enqueued on that stream and the call
 * returns without synchronising; it calls no hipMalloc / hipFree and no
 * allocating library routine.  Partial sums go to a fixed 64 KB device array
 * per GPU, held until the process ends, whose address is resolved on the
 * first call on each device: O above 128 (P2(O) + 1 > 8192 partials) returns
 * CCSD_ERR_BUDGET.  Calls that overlap on one GPU share that array and race;
 * order them (one stream, or events). */
int ccsd_energy_noalloc(int n_occ, int n_virt, const double* fov, const double* oovv, const double* t1,
                        const double* t2, double* e, void* stream);

int ccsd_get_amplitudes(ccsd_ctx* c, double* t1, double* t2);
int ccsd_set_amplitudes(ccsd_ctx* c, const double* t1, const double* t2);

/* Energy of every iteration of the last solve/iterate (E of the MP2 guess
 * first).  Returns the count; copies min(count, cap) values. */
int ccsd_energy_trace(ccsd_ctx* c, double* out, int cap);

/* Adjoint of the converged energy: solves the CCSD Lambda equations and
 * contracts the unrelaxed densities into dE/df (N*N, symmetrised) and
 * dE/d<pq||rs> in the packed block layout of ccsd_problem (symmetrised over
 * duplicate entries of oooo/ovov/vvvv).  Requires a converged solve.  With
 * precision BF16X3 the Lambda iterations follow the forward's mixed schedule
 * (BF16x3 reverse products, then at least final_fp64_sweeps FP64 iterations);
 * the gradient contraction is FP64. */
int ccsd_backward(ccsd_ctx* c, double* dE_dfock, double* dE_doooo, double* dE_dooov, double* dE_doovv,
                  double* dE_dovov, double* dE_dovvv, double* dE_dvvvv, ccsd_grad_info* info);

/* Add delta[k] to n entries of one input array in place (block 0 oooo,
 * 1 ooov, 2 oovv, 3 ovov, 4 ovvv, 5 vvvv: flat index into the packed block;
 * 6: flat index into the N x N Fock matrix).  Derived layouts are rebuilt and
 * the amplitudes are kept (warm start for the next solve).  The caller keeps
 * symmetric blocks symmetric. */
int ccsd_update_integrals(ccsd_ctx* c, int block, long n, const long* index, const double* delta);

/* Page-locked host memory (hipHostMalloc) for large outputs such as the
 * ccsd_backward gradients: copies into it run at DMA rate and overlap device
 * work.  No counterpart in the reference (its gradients do not exist). */
int ccsd_host_alloc(size_t bytes, void** out);
int ccsd_host_free(void* p);

/* Device bytes a context with the default options holds (or would hold for O, V before
 * create): DIIS(8), Strassen temporaries from OV = 8192. */
size_t ccsd_plan_bytes(int n_occ, int n_virt, int precision);

/* Per-kernel-class timing of the last iteration (milliseconds), HIP events on
 * the context's stream: [0] pp ladder, [1] ring, [2] W_mbej quadratic,
 * [3] W_mbij, [4] everything else; for a context made by ccsd_create_rhf,
 * [5] the AO->MO transform and [6] the spin-orbital packing at creation.
 * Returns the count written (at most 7). */
int ccsd_last_timings(ccsd_ctx* c, double* ms, int cap);

void ccsd_destroy(ccsd_ctx* c);
const char* ccsd_strerror(int code);

/* ---- Closed-shell (spin-adapted) CCSD for RHF inputs ----------------------
 * The spin-block mode of the reference's RHF path: CCSD_Memory_optimized::CCSD
 * (ccsd_memory_optimized.cpp:16-567) iterates spin-orbital amplitudes of an
 * RHF reference and skips the spin-forbidden blocks (:483-485); for a
 * closed-shell reference all spin blocks follow from the spatial amplitudes
 * t1[i][a] = t_{i alpha}^{a alpha} and t2[i][j][a][b] = t_{i alpha j beta}^{a
 * alpha b beta} (o = n_occ - n_frozen, v = n_basis - n_occ; t2 stored in full,
 * t2[i][j][a][b] = t2[j][i][b][a]), which this context iterates directly.
 * Same inputs as ccsd_create_rhf; the correlation energy equals that of the
 * spin-orbital context on the same problem.  FP64 only (CCSD_ERR_ARG for
 * another precision).  Set-up needs the n^4 integrals in HBM (twice with C). */
typedef struct ccsd_rhf_ctx ccsd_rhf_ctx;
int ccsd_rhf_create(const ccsd_rhf_problem* p, const ccsd_options* o, ccsd_rhf_ctx** out);
/* Device bytes ccsd_rhf_create needs at its peak (8-vector DIIS, AO input). */
size_t ccsd_rhf_plan_bytes(int n_basis, int n_occ, int n_frozen);
/* As ccsd_solve: from the MP2 guess (or ccsd_rhf_set_amplitudes) until |dE| < tol. */
int ccsd_rhf_solve(ccsd_rhf_ctx* c, double tol, int max_iter, ccsd_result* out);
/* n updates (with DIIS) from the current amplitudes; energies[n] may be NULL. */
int ccsd_rhf_iterate(ccsd_rhf_ctx* c, int n, double* energies);
/* Back to the MP2 guess, DIIS cleared. */
int ccsd_rhf_reset(ccsd_rhf_ctx* c);
/* Right-hand sides r1 [o][v], r2 [o][o][v][v] of D t = R at (t1, t2) (host arrays). */
int ccsd_rhf_residual(ccsd_rhf_ctx* c, const double* t1, const double* t2, double* r1, double* r2);
/* E = sum [2 (ia|jb) - (ib|ja)] (t2[i][j][a][b] + t1[i][a] t1[j][b]) at host amplitudes. */
int ccsd_rhf_energy(ccsd_rhf_ctx* c, const double* t1, const double* t2, double* e);
int ccsd_rhf_get_amplitudes(ccsd_rhf_ctx* c, double* t1, double* t2);
int ccsd_rhf_set_amplitudes(ccsd_rhf_ctx* c, const double* t1, const double* t2);
/* ms[0]: device time of the last right-hand side (ms). */
int ccsd_rhf_last_timings(ccsd_rhf_ctx* c, double* ms, int cap);
const char* ccsd_rhf_last_error(ccsd_rhf_ctx* c);
void ccsd_rhf_destroy(ccsd_rhf_ctx* c);

/* Test hook: copy the named FP64 intermediate of the last right-hand side
 * ("Fme", "Fmi", "Fae", "W1", "W2", "W3", "WT", "X", "Y", "Z", "Rp", "R1",
 * "tau", "Tring", "Tv", "T2u") into out; returns its element count or < 0. */
long ccsd_debug_get(ccsd_ctx* c, const char* name, double* out, long cap);

/* Test hook: run the MFMA GEMM engine on host matrices
 * C(MxN) = alpha*A(MxK)*B(KxN) + beta*C, A/B/C row-major, op flags transpose
 * the stored A (ta) / B (tb); precision selects FP64 / FP32 / BF16X3 arithmetic.
 * Bits 8..15 of precision force a tile configuration (0 = automatic): for
 * FP64 / FP32 the generic MFMA engine's configuration index + 1 (1..17); for
 * BF16X3 the hi/lo kernel's (1 = 128x128, 2 = 256x256 with the automatic
 * tile order, 3 = 256x256 plain order, 4 = 256x256 grouped order, 5 =
 * 160x320, 6 = 256x48).  Bit 16 (FP64) forces single-element global loads
 * where the engine would load element pairs.  Bit 17 (FP64 or FP32, M = N = K
 * even, beta = 0) runs the product through one level of Strassen's method
 * (seven half-size engine calls). */
int ccsd_selftest_gemm(int M, int N, int K, int ta, int tb, double alpha, const double* A, const double* B,
                       double beta, double* C, int precision);

#ifdef __cplusplus
}
#endif

#endif /* CCSD_AMD_H */
