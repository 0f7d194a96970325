// CCSD right-hand side on the GPU: one amplitude update expressed as MFMA GEMMs
// over packed antisymmetric layouts plus bandwidth kernels.
//
// Factorisation follows CCSD_Memory_optimized (ccsd_memory_optimized.cpp):
//   F_me, F_mi (+F_me t_i^e, Fill_F2 :1591-1629), F_ae (-F_me t_m^a, Fill_F3 :1631-1667),
//   W_mnij with 1/2 tau (W_1, :659-1029), W_mnie (W_3), W_mbij with 1/2 <mb||ef> tau
//   (W_2, :1047-1295), W_mbej (W_4, Fill_W4_MPI :4781-5089), ring (Fill_W5 :6201-6882),
//   bare pp ladder 1/2 <ab||cd> tau and the rest of Fill_t2_new (:2001-2458).
// The reference evaluates each term as Armadillo dot products over compact RHF
// slices; here every O(N^5-6) term is one strided/batched GEMM call.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "engine.hpp"
#include "lowprec.hpp"
#include "strassen.hpp"

namespace ccsd {

namespace {

#define TRY(x)                         \
  do {                                 \
    int _rc = (x);                     \
    if (_rc != CCSD_OK) return _rc;    \
  } while (0)

#define HIPTRY(x)                                      \
  do {                                                 \
    hipError_t _e = (x);                               \
    if (_e != hipSuccess) {                            \
      c->last_error = hipGetErrorString(_e);           \
      return CCSD_ERR_HIP;                             \
    }                                                  \
  } while (0)

}  // namespace

// K splits of the FP64 pp ladder: whole-K tiles re-stream the tau panel set
// (P2(O) x P2(V) doubles, 280 MB at O=40/V=300) once per column tile on each
// XCD, and a set larger than half the 256 MB last-level cache comes from HBM
// each time.  The ladder is split into the fewest k-chunks whose tau panels fit
// (measured at O=40/V=300: 3 chunks 47.1 ms against 48.1 ms for whole-K tiles
// with the split tail; 2 chunks 49.6 ms); 0 (modelled choice) when the whole
// set fits, or when more than 4 chunks or 4 GB of partial slices would be
// needed (not measured there; the workspace stays small).
static int ladder_splits(const Dims& d) {
  const double set = (double)d.po * (double)d.pv * sizeof(double), half_llc = 128e6;
  if (set <= half_llc) return 0;
  const int s = (int)std::ceil(set / half_llc);
  return (s > 4 || s * set > 4e9) ? 0 : s;
}

// GEMM workspace of a context: at least kGemmWorkspaceBytes, and room for the
// pp ladder's k-split partial slices (P2(O) x P2(V) each) when it is split
// (which also gives W_mbij more k-splits: 3 -> 10 at O=40/V=300).
static size_t workspace_bytes(const Dims& d) {
  const int ls = ladder_splits(d);
  if (ls == 0) return kGemmWorkspaceBytes;
  const size_t ladder = (size_t)ls * (size_t)d.po * (size_t)d.pv * sizeof(double) + (16ull << 20);
  return std::max(kGemmWorkspaceBytes, (ladder + 255) / 256 * 256);
}

// Smallest OV at which the (OV)^3 FP64 products take one level of Strassen
// (ccsd_options.strassen_min_order: 0 = kStrassenMinN, < 0 = never).
static bool uses_strassen(const Dims& d, int min_order) {
  if (min_order < 0) return false;
  const long m = min_order == 0 ? kStrassenMinN : std::max(min_order, 2);
  return d.ov % 2 == 0 && d.ov >= m;
}

size_t plan_bytes(int O, int V, int precision, int diis_m, int strassen_min_order) {
  Dims d = make_dims(O, V);
  const long N = O + V;
  const long VV = (long)V * V;
  long n = 0;
  // static
  n += N * N + d.ov;
  n += d.po * d.po + d.po * d.ov + d.po * d.pv + d.ov * d.pv + d.pv * d.pv;
  n += 2 * d.ov * d.ov + d.ov * VV + (long)O * O * d.pv + d.po * VV;
  // amplitudes
  long nvec = d.ov + d.po * d.pv;
  n += 3 * nvec + d.po * d.pv + 2L * O * O * d.pv + d.po * VV + d.ov * d.ov;
  // intermediates
  n += d.ov + (long)O * O + 2 * VV + d.po * d.po + d.po * O * O + (long)O * d.po * V + (long)O * O * O * V;
  n += d.po * d.ov + (long)O * O * d.ov + 2 * d.ov * d.ov + d.po * VV + (long)O * O * d.pv + d.po * d.pv + d.ov;
  // DIIS
  if (diis_m >= 2) n += 2L * diis_m * nvec + diis_m;
  // temporary ovov at setup shares nothing: count it
  n += d.ov * d.ov;
  size_t bytes = (size_t)n * sizeof(double);
  bytes += workspace_bytes(d);  // GEMM workspace
  if (uses_strassen(d, strassen_min_order))  // temporaries, Tring's and Doovv's operand sums
    bytes += (strassen_temp_elems(d.ov) + 8 * (size_t)(d.ov / 2) * (size_t)(d.ov / 2)) * sizeof(double);
  bytes += 1ull << 20;   // reductions
  bytes += lowprec_plan_bytes(d, precision);
  return bytes;
}

static int dalloc(Ctx* c, double** p, long n) {
  if (n <= 0) n = 1;
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, (size_t)n * sizeof(double));
  if (e != hipSuccess) {
    c->last_error = hipGetErrorString(e);
    return CCSD_ERR_HIP;
  }
  c->allocs.push_back(q);
  c->bytes += (size_t)n * sizeof(double);
  *p = (double*)q;
  return CCSD_OK;
}

int plan_and_allocate(Ctx* c, int precision) {
  const Dims& d = c->d;
  const long O = d.O, V = d.V, N = O + V, VV = V * V;
  size_t need = plan_bytes(d.O, d.V, precision, c->opt.diis_dim, c->opt.strassen_min_order);
  size_t budget = c->opt.hbm_budget_bytes;
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    HIPTRY(hipMemGetInfo(&fr, &tot));
    budget = (size_t)(0.9 * (double)fr);
  }
  if (need > budget) {
    c->last_error = "plan needs " + std::to_string(need) + " bytes, budget " + std::to_string(budget);
    return CCSD_ERR_BUDGET;
  }
  TRY(dalloc(c, &c->f, N * N));
  TRY(dalloc(c, &c->fov, d.ov));
  TRY(dalloc(c, &c->oooo, d.po * d.po));
  TRY(dalloc(c, &c->ooov, d.po * d.ov));
  TRY(dalloc(c, &c->oovv, d.po * d.pv));
  TRY(dalloc(c, &c->ovvv, d.ov * d.pv));
  TRY(dalloc(c, &c->vvvv, d.pv * d.pv));
  TRY(dalloc(c, &c->G, d.ov * d.ov));
  TRY(dalloc(c, &c->Doovv, d.ov * d.ov));
  TRY(dalloc(c, &c->ovvv_u, d.ov * VV));
  TRY(dalloc(c, &c->oovv_ou, O * O * d.pv));
  TRY(dalloc(c, &c->oovv_vu, d.po * VV));
  c->nvec = d.ov + d.po * d.pv;
  TRY(dalloc(c, &c->amp, c->nvec));
  TRY(dalloc(c, &c->amp_new, c->nvec));
  TRY(dalloc(c, &c->err, c->nvec));
  TRY(dalloc(c, &c->tau, d.po * d.pv));
  TRY(dalloc(c, &c->T2u, O * O * d.pv));
  TRY(dalloc(c, &c->tauu, O * O * d.pv));
  TRY(dalloc(c, &c->Tv, d.po * VV));
  TRY(dalloc(c, &c->Tring, d.ov * d.ov));
  TRY(dalloc(c, &c->Fme, d.ov));
  TRY(dalloc(c, &c->Fmi, O * O));
  TRY(dalloc(c, &c->Fae, VV));
  TRY(dalloc(c, &c->Fae_t1, VV));
  TRY(dalloc(c, &c->W1, d.po * d.po));
  TRY(dalloc(c, &c->Q, d.po * O * O));
  TRY(dalloc(c, &c->W3, O * d.po * V));
  TRY(dalloc(c, &c->W3u, O * O * O * V));
  TRY(dalloc(c, &c->W2, d.po * d.ov));
  TRY(dalloc(c, &c->U, O * O * d.ov));
  TRY(dalloc(c, &c->WT, d.ov * d.ov));
  TRY(dalloc(c, &c->X, d.ov * d.ov));
  TRY(dalloc(c, &c->Y, d.po * VV));
  TRY(dalloc(c, &c->Z, O * O * d.pv));
  TRY(dalloc(c, &c->Rp, d.po * d.pv));
  TRY(dalloc(c, &c->R1, d.ov));
  c->diis_m = c->opt.diis_dim >= 2 ? c->opt.diis_dim : 0;
  if (c->diis_m) {
    TRY(dalloc(c, &c->dvec, (long)c->diis_m * c->nvec));
    TRY(dalloc(c, &c->derr, (long)c->diis_m * c->nvec));
    TRY(dalloc(c, &c->dcoef, c->diis_m));
    c->B.assign((size_t)c->diis_m * c->diis_m, 0.0);
  }
  TRY(dalloc(c, &c->partial, 1 << 16));
  TRY(dalloc(c, &c->scal, 64));
  double* ws = nullptr;
  const size_t wsb = workspace_bytes(d);
  TRY(dalloc(c, &ws, (long)(wsb / 8)));
  if (uses_strassen(d, c->opt.strassen_min_order)) {
    const long hh = (d.ov / 2) * (d.ov / 2);
    TRY(dalloc(c, &c->strassen_tmp, (long)strassen_temp_elems(d.ov)));
    TRY(dalloc(c, &c->strassen_tring_s, 4 * hh));
    TRY(dalloc(c, &c->strassen_doovv_t, 4 * hh));
  }
  c->ws.ptr = ws;
  c->ws.bytes = wsb;
  // Low-precision operand copies are made on first use (lowprec_amplitudes),
  // after the integrals are in HBM; plan_bytes above already reserved them.
  for (auto& e : c->ev) HIPTRY(hipEventCreate(&e));
  (void)N;
  return CCSD_OK;
}

// Derived static layouts from the packed blocks (ovov consumed into G).
static int finish_static(Ctx* c, const double* ovov_dev) {
  const Dims& d = c->d;
  hipStream_t s = c->stream;
  k_build_G(c->G, ovov_dev, d, s);
  k_build_Doovv(c->Doovv, c->oovv, d, s);
  TRY(strassen_doovv_sums(c));
  k_build_ovvv_u(c->ovvv_u, c->ovvv, d, s);
  k_build_oovv_ou(c->oovv_ou, c->oovv, d, s);
  k_build_oovv_vu(c->oovv_vu, c->oovv, d, s);
  std::vector<double> fov((size_t)d.ov);
  const long N = d.O + d.V;
  for (long i = 0; i < d.O; ++i)
    for (long a = 0; a < d.V; ++a) fov[i * d.V + a] = c->h_f[i * N + d.O + a];
  HIPTRY(hipMemcpyAsync(c->fov, fov.data(), fov.size() * 8, hipMemcpyHostToDevice, s));
  HIPTRY(hipMemcpyAsync(c->f, c->h_f.data(), c->h_f.size() * 8, hipMemcpyHostToDevice, s));
  HIPTRY(hipGetLastError());
  HIPTRY(hipStreamSynchronize(s));
  return CCSD_OK;
}

int setup_from_host(Ctx* c, const ccsd_problem* p) {
  const Dims& d = c->d;
  hipStream_t s = c->stream;
  auto up = [&](double* dst, const double* src, long n) -> int {
    HIPTRY(hipMemcpyAsync(dst, src, (size_t)n * 8, hipMemcpyHostToDevice, s));
    return CCSD_OK;
  };
  TRY(up(c->oooo, p->oooo, d.po * d.po));
  TRY(up(c->ooov, p->ooov, d.po * d.ov));
  TRY(up(c->oovv, p->oovv, d.po * d.pv));
  TRY(up(c->ovvv, p->ovvv, d.ov * d.pv));
  TRY(up(c->vvvv, p->vvvv, d.pv * d.pv));
  // ovov is staged in X (same ov x ov shape) and consumed into G.
  TRY(up(c->X, p->ovov, d.ov * d.ov));
  return finish_static(c, c->X);
}

int setup_from_rhf(Ctx* c, int n, int nocc, int nfrozen, const double* C, const double* eri) {
  const Dims& d = c->d;
  hipStream_t s = c->stream;
  const size_t n4 = (size_t)n * n * n * n;
  const size_t need = n4 * 8 * (C ? 2 : 1) + (size_t)n * n * 8;
  size_t free_b = 0, total_b = 0;
  HIPTRY(hipMemGetInfo(&free_b, &total_b));
  if (need > free_b) return CCSD_ERR_BUDGET;
  double *d_eri = nullptr, *d_tmp = nullptr, *d_C = nullptr;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(s);
    (void)hipFree(d_eri);
    (void)hipFree(d_tmp);
    (void)hipFree(d_C);
  };
  int rc = CCSD_OK;
  if (hipMalloc(&d_eri, n4 * 8) != hipSuccess || (C && hipMalloc(&d_tmp, n4 * 8) != hipSuccess) ||
      (C && hipMalloc(&d_C, (size_t)n * n * 8) != hipSuccess)) {
    cleanup();
    return CCSD_ERR_BUDGET;
  }
  if (hipMemcpyAsync(d_eri, eri, n4 * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
      (C && hipMemcpyAsync(d_C, C, (size_t)n * n * 8, hipMemcpyHostToDevice, s) != hipSuccess))
    rc = CCSD_ERR_HIP;
  // HIP events on the stream: [5] transform, [6] spin-orbital packing
  if (rc == CCSD_OK && hipEventRecord(c->ev[10], s) != hipSuccess) rc = CCSD_ERR_HIP;
  if (rc == CCSD_OK && C) rc = ao_to_mo_device(d_eri, d_tmp, d_C, n, c->ws, s);
  if (rc == CCSD_OK && hipEventRecord(c->ev[11], s) != hipSuccess) rc = CCSD_ERR_HIP;
  if (rc == CCSD_OK) {
    // ovov is staged in X (same ov x ov shape) and consumed into G
    double* dst[6] = {c->oooo, c->ooov, c->oovv, c->X, c->ovvv, c->vvvv};
    for (int b = 0; b < 6; ++b) k_rhf_block(dst[b], b, d, d_eri, n, nocc, nfrozen, s);
    if (hipGetLastError() != hipSuccess) rc = CCSD_ERR_HIP;
  }
  if (rc == CCSD_OK && hipEventRecord(c->ev[9], s) != hipSuccess) rc = CCSD_ERR_HIP;
  if (rc == CCSD_OK) {
    float t_tr = 0.f, t_pack = 0.f;
    if (hipEventSynchronize(c->ev[9]) != hipSuccess || hipEventElapsedTime(&t_tr, c->ev[10], c->ev[11]) != hipSuccess ||
        hipEventElapsedTime(&t_pack, c->ev[11], c->ev[9]) != hipSuccess)
      rc = CCSD_ERR_HIP;
    c->timings_ms[5] = t_tr;
    c->timings_ms[6] = t_pack;
  }
  cleanup();
  if (rc != CCSD_OK) return rc;
  return finish_static(c, c->X);
}

int setup_synthetic(Ctx* c, uint64_t seed, double scale) {
  const Dims& d = c->d;
  hipStream_t s = c->stream;
  k_synth_block(c->oooo, 0, 0, d.po, d, seed, scale, s);
  k_synth_block(c->ooov, 1, 0, d.po, d, seed, scale, s);
  k_synth_block(c->oovv, 2, 0, d.po, d, seed, scale, s);
  k_synth_block(c->X, 3, 0, d.ov, d, seed, scale, s);
  k_synth_block(c->ovvv, 4, 0, d.ov, d, seed, scale, s);
  k_synth_block(c->vvvv, 5, 0, d.pv, d, seed, scale, s);
  HIPTRY(hipGetLastError());
  return finish_static(c, c->X);
}

// FP32 product into an FP64 destination (the low-precision iterations).
static int run_f32(Ctx* c, const GemmDesc<float>& g) {
  hipError_t e = gemm<float>(g, c->ws, c->stream);
  if (e != hipSuccess) {
    c->last_error = hipGetErrorString(e);
    return CCSD_ERR_HIP;
  }
  return CCSD_OK;
}

// The (OV)^3 FP64 products: one level of Strassen (strassen.hip) when the
// context holds its temporaries (sized for order OV), else the classical
// engine call.  With
// beta == 1 the Strassen product goes to `scratch` (M x N, contiguous) first
// and is added to C after.  ws: the engine workspace (the context's if null).
int square_product(Ctx* c, const GemmDesc<double>& g, double* scratch, const GemmWorkspace* ws,
                   const StrassenSums* sums) {
  GemmDesc<double> h = g;
  if (g.beta != 0.0) {
    h.C = scratch;
    h.beta = 0.0;
  }
  if (!c->strassen_tmp || g.M != c->d.ov || (g.beta != 0.0 && (!scratch || g.beta != 1.0 || g.sCm != g.N)) ||
      !strassen_applies(h, 2))
    return run_gemm(c, g, ws);
  hipError_t e = gemm_strassen(h, c->strassen_tmp, ws ? *ws : c->ws, c->stream, sums);
  if (e != hipSuccess) {
    c->last_error = hipGetErrorString(e);
    return CCSD_ERR_HIP;
  }
  if (g.beta != 0.0) k_axpy(g.C, scratch, 1.0, (long)g.M * g.N, c->stream);
  return CCSD_OK;
}

// The W_mbej quadratic term's product Tring . Doovv (B = Doovv, k-major rows):
// the descriptor the operand sums are formed for and used with.
static GemmDesc<double> wmbej_quad_desc(Ctx* c, double beta) {
  const long ov = c->d.ov;
  return G64(c->Tring, ov, 1, c->Doovv, ov, 1, c->WT, ov, 1, ov, ov, ov, 0.5, beta);
}

int strassen_doovv_sums(Ctx* c) {
  if (!c->strassen_doovv_t) return CCSD_OK;
  hipError_t e = strassen_sums(wmbej_quad_desc(c, 0.0), nullptr, c->strassen_doovv_t, c->stream);
  if (e != hipSuccess) {
    c->last_error = hipGetErrorString(e);
    return CCSD_ERR_HIP;
  }
  return CCSD_OK;
}

// All right-hand-side pieces at amplitudes (t1, t2) (device pointers):
//   Rp  = <ab||ij> + pp ladder + hh ladder           (packed)
//   X   = ring GEMM (P(ij)P(ab) applied at assembly)
//   Y   = P(ab) pieces, Z = P(ij) pieces, R1 = T1 right-hand side.
// FP64 throughout when prec is FP64;
// otherwise the four dominant contractions run in `prec` (lowprec.hip) and all
// other contractions on FP32 copies of their operands, each accumulating into
// the same FP64 intermediates.
int compute_rhs(Ctx* c, const double* t1, const double* t2, int prec, bool time_it) {
  const Dims& d = c->d;
  const long O = d.O, V = d.V, VV = V * V, po = d.po, pv = d.pv, ov = d.ov;
  hipStream_t s = c->stream;
  const bool lowp = prec != CCSD_PREC_FP64;
  float ms_pp = 0, ms_ring = 0, ms_wq = 0, ms_w2 = 0, ms_all = 0;
  if (time_it) HIPTRY(hipEventRecord(c->ev[0], s));
  LowPrec* L = nullptr;
  // contraction g in FP64, or on the FP32 operands A32 / B32 (g's strides)
  auto mm = [&](const GemmDesc<double>& g, const float* A32, const float* B32) -> int {
    return lowp ? run_f32(c, f32_desc(g, A32, B32)) : run_gemm(c, g);
  };

  // -- amplitude layouts
  k_build_tau(c->tau, t2, t1, d, s);
  if (lowp) {
    TRY(lowprec_allocate(c, prec));
    L = c->lp;
    k_unpack_occ_f32(L->T2u_f, nullptr, t2, c->tau, d, s);
    k_unpack_vir_ring_lp(L->Tv_f, lowprec_ring_out(c, prec), t2, d, s);
    k_to_f32(L->t1_f, t1, ov, s);
    TRY(lowprec_amplitudes(c, prec));
  } else {
    k_unpack_occ(c->T2u, c->tauu, t2, c->tau, d, s);
    k_unpack_vir_ring(c->Tv, c->Tring, t2, d, s);
  }
  const float* t1f = lowp ? L->t1_f : nullptr;

  // -- F_me = f_me + <mn||ef> t_n^f
  k_fme_init(c->Fme, c->f, d, s);
  if (lowp) TRY(lowprec_fme(c, prec, t1));
  else TRY(run_gemm(c, G64(c->Doovv, ov, 1, t1, 1, 0, c->Fme, 1, 0, ov, 1, ov, 1.0, 1.0)));

  // -- W_mnij = <mn||ij> + 1/2 tau <mn||ef> + P(ij) t_j^e <mn||ie>
  k_copy(c->W1, c->oooo, po * po, s);
  {
    GemmDesc<double> g = G64(c->oovv, pv, 1, c->tau, 1, pv, c->W1, po, 1, po, po, pv, 1.0, 1.0);
    if (prec == CCSD_PREC_BF16X3) {
      TRY(lowprec_w1_bf16(c));
    } else if (lowp) {
      GemmDesc<float> f = f32_desc(g, L->oovv_f, L->tau32);
      f.sBn = L->ld_pv;
      TRY(run_f32(c, f));
    } else {
      TRY(run_gemm(c, g));
    }
  }
  TRY(run_gemm(c, G64(c->ooov, V, 1, t1, 1, V, c->Q, O, 1, po * O, O, V, 1.0, 0.0)));
  k_w1_finish(c->W1, c->Q, d, s);
  if (lowp) k_to_f32(L->W1_f, c->W1, po * po, s);

  // -- W_mnie (layout [i][mn][e]) = <mn||ie> + t_i^f <mn||fe>
  k_w3_init(c->W3, c->ooov, d, s);
  {
    GemmDesc<double> g = G64(t1, V, 1, c->oovv_vu, V, 1, c->W3, po * V, 1, O, V, V, 1.0, 1.0);
    g.nb1 = (int)po; g.sB1 = VV; g.sC1 = V;
    TRY(mm(g, t1f, lowp ? L->oovv_vu_f : nullptr));
  }
  k_w3_unpack(c->W3u, c->W3, d, s);
  if (lowp) {
    k_to_f32(L->W3_f, c->W3, O * po * V, s);
    k_to_f32(L->W3u_f, c->W3u, O * O * O * V, s);
  }

  // -- W_mbij (layout [ij][(m,b)]) = <mb||ij> + 1/2 tau <mb||ef> + P(ij) t_i^e <mb||ej>.
  // The product P = 1/2 tau <mb||ef> lands in W2 alone first: the T1 term
  // -1/2 tau_im^ef <ma||ef> is its partial trace over m (k_r1_from_w2), so that
  // term costs no second pass over <ma||ef>.
  k_r1_init(c->R1, c->f, t1, d, s);
  if (time_it) HIPTRY(hipEventRecord(c->ev[1], s));
  if (lowp) {
    TRY(lowprec_w2(c, prec));
  } else {
    TRY(run_gemm(c, G64(c->tau, pv, 1, c->ovvv, 1, pv, c->W2, ov, 1, po, ov, pv, 1.0, 0.0)));
  }
  if (time_it) HIPTRY(hipEventRecord(c->ev[2], s));
  k_r1_from_w2(c->R1, c->W2, d, s);
  {
    GemmDesc<double> g = G64(t1, V, 1, c->G, ov, 1, c->U, O * ov, 1, O, V, V, 1.0, 0.0);
    g.nb1 = (int)O; g.nb2 = (int)O;  // b1 = m, b2 = j
    g.sB1 = V * ov; g.sB2 = V;
    g.sC1 = V; g.sC2 = ov;
    TRY(mm(g, t1f, lowp ? L->G_f : nullptr));
  }
  k_w2_finish(c->W2, c->U, c->ooov, d, s);
  if (lowp) k_to_f32(L->W2_f, c->W2, po * ov, s);

  // -- W_mbej stored transposed: WT[(j,b),(m,e)].  In the low-precision modes
  // it accumulates in FP32 instead (lowprec_wmbej_acc: leading dimension ldw),
  // since its only consumer there is the ring GEMM's FP32 / BF16 operand.
  long ldw = ov;
  float* Wf = lowp ? lowprec_wmbej_acc(c, prec, &ldw) : nullptr;
  {
    // S[(m,b,e)][j] = sum_f <mb||ef> t_j^f as one tall GEMM (M = O V^2), staged in X
    // (free until the ring GEMM); then WT = S^(permuted) + <mb||ej>.
    double* S = c->X;
    GemmDesc<double> g = G64(c->ovvv_u, V, 1, t1, 1, V, S, O, 1, O * VV, O, V, 1.0, 0.0);
    if (lowp) {  // S in FP32 (half the traffic of the two consumers)
      float* S32 = reinterpret_cast<float*>(c->X);
      if (prec == CCSD_PREC_BF16X3) {
        TRY(lowprec_wmbej_s_bf16(c, t1, S32));
      } else {
        GemmDesc<float> f = f32_desc(g, L->ovvv_u_f, t1f);
        f.Cd = nullptr;
        f.C = S32;
        TRY(run_f32(c, f));
      }
      k_fae_t1_from_s_f32(c->Fae_t1, S32, d, s);
      k_wt_from_s_f32(Wf, ldw, S32, L->G_f, d, s);
    } else {
      TRY(run_gemm(c, g));
      k_fae_t1_from_s(c->Fae_t1, S, d, s);  // F_ae t1 part = -sum_m S[(m,a,e)][m]
      k_wt_from_s(c->WT, S, c->G, d, s);
    }
  }
  {
    // + t_n^b W_mnje, batched over (j, m)
    GemmDesc<double> g = G64(t1, 1, V, c->W3u, V, 1, c->WT, ov, 1, V, V, O, 1.0, 1.0);
    g.nb1 = (int)O; g.nb2 = (int)O;  // b1 = j, b2 = m
    g.sB1 = O * O * V; g.sB2 = O * V;
    g.sC1 = V * ov; g.sC2 = V;
    if (lowp) {
      GemmDesc<float> f = f32_desc(g, t1f, L->W3u_f);
      f.Cd = nullptr;
      f.C = Wf;
      f.sCm = ldw; f.sC1 = V * ldw;
      TRY(run_f32(c, f));
    } else {
      TRY(run_gemm(c, g));
    }
  }
  if (time_it) HIPTRY(hipEventRecord(c->ev[3], s));
  if (lowp) {
    TRY(lowprec_wmbej(c, prec));
  } else {
    // + 1/2 t_jn^bf <mn||ef>: one-level Strassen into X (free here: S is
    // consumed), then WT += X; the classical product accumulates directly.
    // Tring's operand sums serve this product and the ring's; Doovv's are
    // formed with Doovv.
    StrassenSums sw;
    if (c->strassen_tring_s) {
      hipError_t e = strassen_sums(wmbej_quad_desc(c, 0.0), c->strassen_tring_s, nullptr, s);
      if (e != hipSuccess) {
        c->last_error = hipGetErrorString(e);
        return CCSD_ERR_HIP;
      }
      sw.S = c->strassen_tring_s;
      sw.T = c->strassen_doovv_t;
    }
    TRY(square_product(c, wmbej_quad_desc(c, 1.0), c->X, nullptr, &sw));
  }
  if (time_it) HIPTRY(hipEventRecord(c->ev[4], s));

  // -- F_mi
  k_fmi_init(c->Fmi, c->f, c->ooov, t1, d, s);
  TRY(mm(G64(c->oovv_ou, O * pv, 1, c->T2u, 1, O * pv, c->Fmi, O, 1, O, O, O * pv, 1.0, 1.0),
         lowp ? L->oovv_ou_f : nullptr, lowp ? L->T2u_f : nullptr));
  k_fmi_finish(c->Fmi, c->Fme, t1, d, s);
  if (lowp) k_to_f32(L->Fmi_f, c->Fmi, O * O, s);
  // -- F_ae
  k_fae_init(c->Fae, c->f, c->Fae_t1, c->Fme, t1, d, s);
  TRY(mm(G64(c->Tv, 1, V, c->oovv_vu, V, 1, c->Fae, V, 1, V, V, po * V, -1.0, 1.0), lowp ? L->Tv_f : nullptr,
         lowp ? L->oovv_vu_f : nullptr));
  if (lowp) k_to_f32(L->Fae_f, c->Fae, VV, s);

  // -- packed residual: <ab||ij> + pp ladder + hh ladder
  k_copy(c->Rp, c->oovv, po * pv, s);
  if (time_it) HIPTRY(hipEventRecord(c->ev[5], s));
  // pp ladder Rp[ij][ab] += sum_{e<f} tau[ij][ef] <ab||ef> (Fill_t2_new :2093):
  // row ab of the packed vvvv block is k-contiguous, so B is read as rows
  // (which also admits the 8-wave 112 x 128 tile; an n-contiguous B does not)
  if (lowp) {
    TRY(lowprec_ladder(c, prec));
  } else {
    TRY(run_gemm(c, G64(c->tau, pv, 1, c->vvvv, 1, pv, c->Rp, pv, 1, po, pv, pv, 1.0, 1.0), nullptr,
                 ladder_splits(d)));
  }
  if (time_it) HIPTRY(hipEventRecord(c->ev[6], s));
  {
    GemmDesc<double> g = G64(c->W1, 1, po, c->tau, pv, 1, c->Rp, pv, 1, po, pv, po, 1.0, 1.0);
    if (prec == CCSD_PREC_BF16X3) {
      TRY(lowprec_hh_bf16(c));
    } else if (lowp) {
      GemmDesc<float> f = f32_desc(g, L->W1_f, L->tau32);
      f.sBk = L->ld_pv;
      TRY(run_f32(c, f));
    } else {
      TRY(run_gemm(c, g));
    }
  }

  // -- ring X = Tring * W_mbej
  if (time_it) HIPTRY(hipEventRecord(c->ev[7], s));
  if (lowp) {
    TRY(lowprec_ring(c, prec));
  } else {
    // A = Tring with the W_mbej product's strides: its operand sums, formed there
    StrassenSums sr;
    sr.S = c->strassen_tring_s;
    TRY(square_product(c, G64(c->Tring, ov, 1, c->WT, 1, ov, c->X, ov, 1, ov, ov, ov, 1.0, 0.0), nullptr, nullptr,
                       &sr));
  }
  if (time_it) HIPTRY(hipEventRecord(c->ev[8], s));

  // -- Y[ij][a][b] = t_ij^ae F_be - t_m^a W_mbij
  if (prec == CCSD_PREC_BF16X3) {
    TRY(lowprec_y_bf16(c));
  } else {
    TRY(mm(G64(c->Tv, V, 1, c->Fae, 1, V, c->Y, V, 1, po * V, V, V, 1.0, 0.0), lowp ? L->Tv_f : nullptr,
           lowp ? L->Fae_f : nullptr));
  }
  {
    GemmDesc<double> g = G64(t1, 1, V, c->W2, V, 1, c->Y, V, 1, V, V, O, -1.0, 1.0);
    g.nb1 = (int)po; g.sB1 = ov; g.sC1 = VV;
    TRY(mm(g, t1f, lowp ? L->W2_f : nullptr));
  }
  // -- Z[i][j][ab] = -t_im^ab F_mj - t_i^e <je||ab>
  {
    GemmDesc<double> g = G64(c->Fmi, 1, O, c->T2u, pv, 1, c->Z, pv, 1, O, pv, O, -1.0, 0.0);
    g.nb1 = (int)O; g.sB1 = O * pv; g.sC1 = O * pv;
    TRY(mm(g, lowp ? L->Fmi_f : nullptr, lowp ? L->T2u_f : nullptr));
  }
  {
    GemmDesc<double> g = G64(t1, V, 1, c->ovvv, pv, 1, c->Z, O * pv, 1, O, pv, V, -1.0, 1.0);
    g.nb1 = (int)O; g.sB1 = V * pv; g.sC1 = pv;
    if (lowp) {
      GemmDesc<float> f = f32_desc(g, t1f, L->ovvv32);
      f.sBk = L->ld_pv; f.sB1 = V * L->ld_pv;
      TRY(run_f32(c, f));
    } else {
      TRY(run_gemm(c, g));
    }
  }

  // -- T1 right-hand side (R1 holds its f and tau.<ma||ef> terms already)
  TRY(mm(G64(c->W3, po * V, 1, c->Tv, V, 1, c->R1, V, 1, O, V, po * V, 1.0, 1.0), lowp ? L->W3_f : nullptr,
         lowp ? L->Tv_f : nullptr));                                                                   // -1/2 t_mn^ae W_mnie
  if (lowp) TRY(lowprec_r1_ring(c, prec));                                                       // t_im^ae F_me
  else TRY(run_gemm(c, G64(c->Tring, ov, 1, c->Fme, 1, 0, c->R1, 1, 0, ov, 1, ov, 1.0, 1.0)));
  TRY(mm(G64(c->G, ov, 1, t1, 1, 0, c->R1, 1, 0, ov, 1, ov, 1.0, 1.0), lowp ? L->G_f : nullptr, t1f));  // -t_n^f <na||if>
  TRY(run_gemm(c, G64(c->Fmi, 1, O, t1, V, 1, c->R1, V, 1, O, V, O, -1.0, 1.0)));               // -t_m^a F_mi

  if (time_it) {
    HIPTRY(hipEventRecord(c->ev[9], s));
    HIPTRY(hipEventSynchronize(c->ev[9]));
    HIPTRY(hipEventElapsedTime(&ms_w2, c->ev[1], c->ev[2]));
    HIPTRY(hipEventElapsedTime(&ms_wq, c->ev[3], c->ev[4]));
    HIPTRY(hipEventElapsedTime(&ms_pp, c->ev[5], c->ev[6]));
    HIPTRY(hipEventElapsedTime(&ms_ring, c->ev[7], c->ev[8]));
    HIPTRY(hipEventElapsedTime(&ms_all, c->ev[0], c->ev[9]));
    c->timings_ms[0] = ms_pp;
    c->timings_ms[1] = ms_ring;
    c->timings_ms[2] = ms_wq;
    c->timings_ms[3] = ms_w2;
    c->timings_ms[4] = ms_all - ms_pp - ms_ring - ms_wq - ms_w2;
  }
  HIPTRY(hipGetLastError());
  return CCSD_OK;
}

// E = f_ia t_i^a + sum_{i<j,a<b} <ij||ab> tau_ij^ab  (Calc_Energy, :592-648)
int device_energy(Ctx* c, const double* t1, const double* t2, double* e) {
  const Dims& d = c->d;
  hipStream_t s = c->stream;
  const int nb2 = k_energy_partials(c->partial, c->oovv, t2, t1, d, s), nb1 = reduce_blocks_for(d.ov);
  k_dot_partials(c->partial + nb2, c->fov, t1, d.ov, nb1, s);
  k_sum_partials(c->scal, c->partial, nb2, 1, s);
  k_sum_partials(c->scal + 1, c->partial + nb2, nb1, 1, s);
  double h[2];
  HIPTRY(hipMemcpyAsync(h, c->scal, 16, hipMemcpyDeviceToHost, s));
  HIPTRY(hipStreamSynchronize(s));
  *e = h[0] + h[1];
  return CCSD_OK;
}

}  // namespace ccsd
