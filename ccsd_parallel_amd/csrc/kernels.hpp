// Bandwidth-bound CCSD kernels (FP64) and their host launchers.
//
// Index conventions (shared with the packed C-ABI layouts):
//   O occupied, V virtual spin orbitals; pair index pidx(p,q) = q(q-1)/2 + p, p<q;
//   po = P2(O), pv = P2(V), ov = O*V.
//   t1[i*V+a];  packed t2[ij*pv + ab] (i<j, a<b)
//   T2u/tauu  [i][m][ab]      occupied pair unpacked (O*O*pv), antisymmetric in i,m
//   Tv        [ij][a][b]      virtual pair unpacked (po*V*V), antisymmetric in a,b
//   Tring     [(i,a),(m,e)]   t_im^ae as an ov x ov matrix
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ccsd {

struct Dims {
  int O, V;
  long po, pv, ov;
};

inline Dims make_dims(int O, int V) {
  Dims d;
  d.O = O; d.V = V;
  d.po = (long)O * (O - 1) / 2;
  d.pv = (long)V * (V - 1) / 2;
  d.ov = (long)O * V;
  return d;
}

// -- generic -------------------------------------------------------------------
void k_copy(double* dst, const double* src, long n, hipStream_t s);
void k_axpy(double* y, const double* x, double a, long n, hipStream_t s);  // y += a x
void k_fill(double* y, double v, long n, hipStream_t s);
void k_to_f32(float* dst, const double* src, long n, hipStream_t s);
void k_split_bf16(uint16_t* hi, uint16_t* lo, const double* src, long n, hipStream_t s);

// -- static integral re-layouts (setup) -------------------------------------------
void k_build_G(double* G, const double* ovov, Dims d, hipStream_t s);          // G[(m,e),(j,b)] = <mb||ej>
void k_build_Doovv(double* D, const double* oovv, Dims d, hipStream_t s);      // D[(m,e),(n,f)] = <mn||ef>
void k_build_ovvv_u(double* U, const double* ovvv, Dims d, hipStream_t s);     // U[m][b][e][f] = <mb||ef>
void k_build_oovv_ou(double* U, const double* oovv, Dims d, hipStream_t s);    // U[m][n][ef]
void k_build_oovv_vu(double* U, const double* oovv, Dims d, hipStream_t s);    // U[mn][e][f]

// -- amplitude layouts ------------------------------------------------------------
void k_build_tau(double* tau, const double* t2, const double* t1, Dims d, hipStream_t s);
void k_unpack_occ(double* T2u, double* tauu, const double* t2, const double* tau, Dims d, hipStream_t s);
void k_unpack_vir(double* Tv, const double* t2, Dims d, hipStream_t s);
void k_unpack_ring(double* Tr, const double* t2, Dims d, hipStream_t s);
// Tv and Tring (ring layout, with its zero diagonal blocks) in one pass; either may be null.
void k_unpack_vir_ring(double* Tv, double* Tr, const double* t2, Dims d, hipStream_t s);
// Destination of the ring layout: FP64 or FP32 rows of length ld (>= ov), or
// BF16 hi/lo rows interleaved per 8-element chunk (row r at r * 2 * ld);
// padding columns are left untouched.
struct RingOut {
  double* d = nullptr;
  float* f = nullptr;
  uint16_t* b = nullptr;
  long ld = 0;
};
// Low-precision iterations: FP32 occupied / virtual unpacks, ring layout as given
void k_unpack_occ_f32(float* T2u, float* tauu, const double* t2, const double* tau, Dims d, hipStream_t s);
void k_unpack_vir_ring_lp(float* Tv, const RingOut& tr, const double* t2, Dims d, hipStream_t s);

// -- intermediates ------------------------------------------------------------------
void k_fme_init(double* Fme, const double* f, Dims d, hipStream_t s);
void k_w1_finish(double* W1, const double* Q, Dims d, hipStream_t s);
void k_w3_init(double* W3, const double* ooov, Dims d, hipStream_t s);
void k_w3_unpack(double* W3u, const double* W3, Dims d, hipStream_t s);
// W2 += <mb||ij> + P(ij) U (W2 holds the tau.<mb||ef> product on entry)
void k_w2_finish(double* W2, const double* U, const double* ooov, Dims d, hipStream_t s);
// R1 -= sum_m P(i,m)[(m,a)] from the packed W_mbij product P (before k_w2_finish)
void k_r1_from_w2(double* R1, const double* P, Dims d, hipStream_t s);
void k_fae_t1_extract(double* Fae_t1, const double* WT, Dims d, hipStream_t s);
// S[(m,b,e)][j] = sum_f <mb||ef> t_j^f  ->  WT[(j,b),(m,e)] = S + <mb||ej>;  F_ae t1 part from S
void k_wt_from_s(double* WT, const double* S, const double* G, Dims d, hipStream_t s);
void k_wt_from_s_f32(float* WT, long ldw, const float* S, const float* G, Dims d, hipStream_t s);
void k_fae_t1_from_s(double* Fae_t1, const double* S, Dims d, hipStream_t s);
void k_fae_t1_from_s_f32(double* F, const float* S, Dims d, hipStream_t s);
void k_fmi_init(double* Fmi, const double* f, const double* ooov, const double* t1, Dims d, hipStream_t s);
void k_fmi_finish(double* Fmi, const double* Fme, const double* t1, Dims d, hipStream_t s);
void k_fae_init(double* Fae, const double* f, const double* Fae_t1, const double* Fme, const double* t1, Dims d,
                hipStream_t s);
void k_r1_init(double* R1, const double* f, const double* t1, Dims d, hipStream_t s);

// -- residual assembly / update ---------------------------------------------------------
// R2(ij,ab) = Rp + P(ij)P(ab) X + P(ab) Y + P(ij) Z.  mode 0: write R2 into out;
// mode 1: out = R2 / D (new amplitudes) and err = out - t2_old.
void k_assemble_t2(double* out, double* err, const double* Rp, const double* X, const double* Y, const double* Z,
                   const double* f, const double* t2_old, int mode, Dims d, hipStream_t s);
// t2 = <ij||ab> / D (MP2 guess, ccsd_memory_optimized.cpp:473-490); t1 = 0
void k_mp2_guess(double* amp, const double* oovv, const double* f, Dims d, hipStream_t s);
void k_update_t1(double* t1n, double* err, const double* R1, const double* f, const double* t1_old, Dims d,
                 hipStream_t s);

// -- reductions (deterministic two-pass, FP64) ------------------------------------------
// partial[b] = sum over block b of x.y ; host then reduces num_blocks values.
int reduce_blocks_for(long n);
void k_dot_partials(double* partial, const double* x, const double* y, long n, int nblocks, hipStream_t s);
// Partials of sum_{i<j,a<b} <ij||ab> (t2 + P t1 t1) (tau not stored); returns
// the number of partials written (at most 16384)
int k_energy_partials(double* partial, const double* oovv, const double* t2, const double* t1, Dims d, hipStream_t s);
void k_maxabs_partials(double* partial, const double* x, long n, int nblocks, hipStream_t s);
// partial[v*nblocks + b] for nv vectors x_v (stride ldx) against y
void k_multidot_partials(double* partial, const double* X, long ldx, int nv, const double* y, long n, int nblocks,
                         hipStream_t s);
// y = sum_v c[v] X_v  (c on device)
void k_lincomb(double* y, const double* X, long ldx, const double* c, int nv, long n, hipStream_t s);
void k_sum_partials(double* out, const double* partial, int nparts, int nsums, hipStream_t s);

// x[idx[k]] += delta[k]  (idx, delta on device); for ovov deltas applied to G:
// G[(m,e),(j,b)] -= delta at ovov index (m,b),(j,e) when to_g is set.
void k_scatter_add(double* x, const long* idx, const double* delta, long n, int to_g, Dims d, hipStream_t s);

// -- synthetic generator (counter-based; restates oracle/synth.py) ------------------------
// block id: 0 oooo, 1 ooov, 2 oovv, 3 ovov, 4 ovvv, 5 vvvv.  rows [r0, r1) of the block.
void k_synth_block(double* out, int block, long r0, long r1, Dims d, uint64_t seed, double scale, hipStream_t s);
// Whole packed block `block` (0 oooo, 1 ooov, 2 oovv, 3 ovov, 4 ovvv, 5 vvvv) of
// <PQ||RS> from RHF spatial MO integrals (pq|rs) [n^4] in HBM; nocc doubly
// occupied spatial orbitals of which the lowest nfrozen are not correlated.
void k_rhf_block(double* out, int block, Dims d, const double* mo, int n, int nocc, int nfrozen, hipStream_t s);

}  // namespace ccsd
