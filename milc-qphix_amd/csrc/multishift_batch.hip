// multishift_batch.hip -- multi-shift CG on several right-hand sides at once,
// FP32 iteration with FP64 reliable updates and per-shift FP64 refinement.
//
// No counterpart in the reference, whose multi-shift solver
// (ks_multicg_offset.cpp:205-689) takes one source in one precision.  Per
// right-hand side this is the same Jegerlehner recursion (relative shifts,
// base system = smallest shift, zeta/beta/alpha recurrences, lazy solution
// updates), with:
//  * up to NR = 8 sources advancing together: one multi-RHS Dslash per
//    half-step (links read once for all NR), one fused update of r and one of
//    every (x_j, p_j) per iteration, each source with its own device-side
//    scalars and gate;
//  * FP32 vectors and links; FP64 solutions x_j, flushed from the FP32
//    accumulators at every reliable update;
//  * reliable updates on the base system: when |r| has dropped 10x (or looks
//    converged, or the step budget is spent) that source pauses, its base true
//    residual b - (M + sigma_low) x_low is recomputed with the FP64 Dslash and
//    replaces r (residual replacement); it stops once that true residual meets
//    the target;
//  * refinement: every shifted solution is then checked in FP64 and, where its
//    own true residual misses the target, finished by mixed-precision CG from
//    its multi-shift value.
#include "blas.h"
#include "dslash_mrhs.h"
#include "handles.h"
#include "qphix_mi355x.h"

#include <algorithm>
#include <chrono>
#include <map>
#include <cmath>
#include <cstring>
#include <vector>

namespace qphix {

template <int NR>
__global__ void __launch_bounds__(FINALIZE_BLOCK)
    k_finalize_multi(const double *partials, int nb, SolverState *st, int op) {
  double s[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    s[r] = 0;
    for (int b = threadIdx.x; b < nb; b += FINALIZE_BLOCK) s[r] += partials[(size_t)r * nb + b];
  }
  block_sum<NR>(s);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (st[r].active) apply_scalar_op(st + r, op, s + r);
  }
}

// r_r += beta_low,r Ap_r ; reduce |r_r|^2 per source.  One work-item per
// (site, source), resident grid walking XCD slabs of nlog logical blocks;
// one partial per block and source.
template <int NR>
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_msx_update_r(float *rs, const float *Ap, int vh, const SolverState *st, double *partials, int nlog) {
  const unsigned mask = rhs_mask<NR>(st);
  if (mask == 0) return;
  const XcdWalk wk(nlog);
  double v = 0.0;
  for (int k = 0;; ++k) {
    const int blk = wk.at(k);
    if (blk < 0) break;
    const int t = blk * BLAS_BLOCK + threadIdx.x;
    const int i = t / NR, r = t % NR;
    if (i >= vh || !((mask >> r) & 1u)) continue;
    const float bl = (float)st[r].beta_i[st[r].j_low];
    cplx<float> rv[3], av[3];
    Spin<float>::load(rs, t, rv);
    Spin<float>::load(Ap, t, av);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      rv[c].re += bl * av[c].re;
      rv[c].im += bl * av[c].im;
      v += (double)rv[c].re * rv[c].re + (double)rv[c].im * rv[c].im;
    }
    Spin<float>::store(rs, t, rv);
  }
  const double sum = lane_block_sum<NR, BLAS_BLOCK>(v);
  if (threadIdx.x < NR) partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = sum;
}

// x_j += beta_j p_j ; p_j = zeta_j r + alpha_j p_j  for every shift j < num_now,
// one work-item per (site, source); layout [shift][site][source][6].
template <int NR>
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_msx_update_pm(float *xs, float *p, const float *rs, int vh, int nshift, const SolverState *st) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= vh * NR) return;
  const int r = t % NR;
  const SolverState &s = st[r];
  if (!s.active) return;
  const size_t stride = (size_t)vh * NR;
  cplx<float> rv[3];
  Spin<float>::load(rs, t, rv);
  const int nn = s.num_now;
  for (int j = 0; j < nn && j < nshift; ++j) {
    if (s.sdone[j] == 1) continue;
    const float bj = (float)s.beta_i[j], aj = (float)s.alpha[j], zj = (float)s.zeta_ip1[j];
    cplx<float> xv[3], pv[3];
    Spin<float>::load(xs + 6 * stride * j, t, xv);
    Spin<float>::load(p + 6 * stride * j, t, pv);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      xv[c].re += bj * pv[c].re;
      xv[c].im += bj * pv[c].im;
      pv[c].re = zj * rv[c].re + aj * pv[c].re;
      pv[c].im = zj * rv[c].im + aj * pv[c].im;
    }
    Spin<float>::store(xs + 6 * stride * j, t, xv);
    Spin<float>::store(p + 6 * stride * j, t, pv);
  }
}

// End of a step: sources that asked for a true-residual check pause, and
// shifts that took their last update retire.
__global__ void k_msx_commit(SolverState *st, int nr) {
  const int r = threadIdx.x;
  if (r >= nr || !st[r].active) return;
  for (int j = 0; j < st[r].num_now; ++j)
    if (st[r].sdone[j] == 2) st[r].sdone[j] = 1;
  if (st[r].ru) st[r].active = 0;
}

// Source slot r: r_s = p_j = (float) b for all j, x_s = 0.
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_msx_init(float *rs, float *p, float *xs, const double *b, int vh, int NR, int r, int nshift) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  const size_t stride = (size_t)vh * NR;
  for (int e = 0; e < 6; ++e) {
    const float v = (float)b[6 * (size_t)i + e];
    rs[6 * ((size_t)i * NR + r) + e] = v;
    for (int j = 0; j < nshift; ++j) {
      p[6 * (stride * j + (size_t)i * NR + r) + e] = v;
      xs[6 * (stride * j + (size_t)i * NR + r) + e] = 0.f;
    }
  }
}

// x64_j += x_s[j] (slot r); x_s[j] = 0
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_msx_flush(double *const *x64, float *xs, int vh, int NR, int r, int nshift) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  const size_t stride = (size_t)vh * NR;
  for (int j = 0; j < nshift; ++j) {
    float *s = xs + 6 * (stride * j + (size_t)i * NR + r);
    double *d = x64[j] + 6 * (size_t)i;
    for (int e = 0; e < 6; ++e) {
      d[e] += (double)s[e];
      s[e] = 0.f;
    }
  }
}

// r_s (slot r) = (float) r64
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_msx_set_r(float *rs, const double *r64, int vh, int NR, int r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  for (int e = 0; e < 6; ++e) rs[6 * ((size_t)i * NR + r) + e] = (float)r64[6 * (size_t)i + e];
}

// Resume a paused source with the FP64 true residual's norm.
__global__ void k_msx_resume(SolverState *s, const SolverState *t) {
  s->rsq = t->rsq;
  s->mx_maxr = sqrt(t->rsq);
  s->ru = 0;
  s->active = 1;
}

// Resident grid of a lane-mapped kernel: CUs x resident blocks per CU (from
// the occupancy calculator), a multiple of N_XCD, no larger than needed.
template <typename K> static int resident_grid(K kernel, int nlog) {
  int bpc = 0;
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, kernel, DSLASH_BLOCK, 0));
  if (bpc < 1) bpc = 1;
  const int need = (nlog + N_XCD - 1) / N_XCD * N_XCD;
  const int g = ctx().n_cu * bpc / N_XCD * N_XCD;
  return std::max(N_XCD, std::min(need, g));
}

template <int NR, bool LANE, int NL, int EPI> static int mrhs_grid_of(int vh) {
  if (!LANE) return (vh + DSLASH_BLOCK - 1) / DSLASH_BLOCK;
  static int g = 0;  // per instantiation; the lattice is fixed between QPHIX_init and QPHIX_finalize
  static int g_vh = -1;
  if (g_vh != vh) {
    g = resident_grid(k_dslash_mrhs_lane<NL, NR, EPI>, (int)(((size_t)vh * NR + DSLASH_BLOCK - 1) / DSLASH_BLOCK));
    g_vh = vh;
  }
  return g;
}

// Blocks of a multi-RHS Dslash launch (for EPI_PKP, the partial count per source).
template <int NR> static int mrhs_blocks(int vh, bool lane, int recon, int epi) {
  if (!lane) return (vh + DSLASH_BLOCK - 1) / DSLASH_BLOCK;
  if (recon == 12)
    return epi == EPI_PLAIN ? mrhs_grid_of<NR, true, 12, EPI_PLAIN>(vh) : mrhs_grid_of<NR, true, 12, EPI_PKP>(vh);
  return epi == EPI_PLAIN ? mrhs_grid_of<NR, true, 18, EPI_PLAIN>(vh) : mrhs_grid_of<NR, true, 18, EPI_PKP>(vh);
}

template <int NR, bool LANE, int NL, int EPI> static void mrhs_launch1(const MrhsArgs &a) {
  const int nb = mrhs_grid_of<NR, LANE, NL, EPI>(a.g.vh);
  if (LANE) hipLaunchKernelGGL((k_dslash_mrhs_lane<NL, NR, EPI>), dim3(nb), dim3(DSLASH_BLOCK), 0, ctx().stream, a);
  else hipLaunchKernelGGL((k_dslash_mrhs<NL, NR, EPI>), dim3(nb), dim3(DSLASH_BLOCK), 0, ctx().stream, a);
}

template <int NR, bool LANE>
static void mrhs_launch(const LinksT<float> &Lf, int epi, const MrhsArgs &a) {
  if (Lf.recon == 12) {
    if (epi == EPI_PLAIN) mrhs_launch1<NR, LANE, 12, EPI_PLAIN>(a);
    else mrhs_launch1<NR, LANE, 12, EPI_PKP>(a);
  } else {
    if (epi == EPI_PLAIN) mrhs_launch1<NR, LANE, 18, EPI_PLAIN>(a);
    else mrhs_launch1<NR, LANE, 18, EPI_PKP>(a);
  }
}

// Multi-RHS Dslash on the [vh][NR][6] interleaved layout.  lane selects the
// thread mapping; EPI_PKP leaves NR x mrhs_blocks<NR>(...) partial sums.
template <int NR>
static void mrhs_dslash(const LinksT<float> &Lf, int cb, float *out, const float *in, int epi,
                        float s, const float *aux, const SolverState *st, bool lane) {
  const Geometry &gg = ctx().g;
  MrhsArgs a{};
  a.fat = reinterpret_cast<const float4 *>(Lf.fat[cb]);
  a.lng = reinterpret_cast<const float4 *>(Lf.lng[cb]);
  a.lsign = Lf.lsign[cb];
  a.in = in;
  a.out = out;
  a.aux = aux;
  a.g = Geo{gg.nxh, gg.ny, gg.nz, gg.nt, gg.vh, gg.nxy, gg.nxyz};
  a.cb = cb;
  a.s = s;
  a.partials = ctx().partials;
  a.states = st;
  a.nlog = (int)(((size_t)gg.vh * NR + DSLASH_BLOCK - 1) / DSLASH_BLOCK);
  if (lane) mrhs_launch<NR, true>(Lf, epi, a);
  else mrhs_launch<NR, false>(Lf, epi, a);
  HIP_CHECK(hipGetLastError());
}

static double wall_now() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

#include "multishift_driver.h"

// ----------------------------------------------------------------- multi-RHS Dslash API
__global__ void k_mrhs_pack(float *dst, const float *const *src, int vh, int n, int NR) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= vh * NR) return;
  const int i = t / NR, r = t % NR;
  for (int e = 0; e < 6; ++e) dst[6 * (size_t)t + e] = r < n ? src[r][6 * (size_t)i + e] : 0.f;
}

__global__ void k_mrhs_unpack(float *const *dst, const float *src, int vh, int n, int NR) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= vh * NR) return;
  const int i = t / NR, r = t % NR;
  if (r >= n) return;
  for (int e = 0; e < 6; ++e) dst[r][6 * (size_t)i + e] = src[6 * (size_t)t + e];
}

template <int NR>
static double dslash_multi(LinksT<float> *L, QPHIX_F3_ColorVector **out, QPHIX_F3_ColorVector **in,
                           int n, int cb, int reps, bool lane) {
  const int vh = ctx().g.vh;
  const size_t nvr = 6 * (size_t)vh * NR;
  char *w = (char *)ws2_get(2 * nvr * sizeof(float) + NR * sizeof(SolverState) + 2 * NR * sizeof(float *) + 1024);
  float *pin = (float *)w, *pout = pin + nvr;
  SolverState *st = (SolverState *)(((uintptr_t)(pout + nvr) + 255) & ~(uintptr_t)255);
  float **ptrs = (float **)(st + NR);
  std::vector<SolverState> hs(NR);
  for (int r = 0; r < NR; ++r) hs[r].active = r < n;
  std::memcpy(ctx().pinned, hs.data(), NR * sizeof(SolverState));
  float *hp[16];
  for (int r = 0; r < n; ++r) {
    hp[r] = in[r]->half[1 - cb];
    hp[NR + r] = out[r]->half[cb];
  }
  std::memcpy((char *)ctx().pinned + NR * sizeof(SolverState), hp, 2 * NR * sizeof(float *));
  HIP_CHECK(hipMemcpyAsync(st, ctx().pinned, NR * sizeof(SolverState) + 2 * NR * sizeof(float *),
                           hipMemcpyHostToDevice, ctx().stream));
  const int nbt = (vh * NR + BLAS_BLOCK - 1) / BLAS_BLOCK;
  hipLaunchKernelGGL(k_mrhs_pack, dim3(nbt), dim3(BLAS_BLOCK), 0, ctx().stream, pin, ptrs, vh, n, NR);
  double ms = 0;
  if (reps > 0) {
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, ctx().stream));
    for (int k = 0; k < reps; ++k) mrhs_dslash<NR>(*L, cb, pout, pin, EPI_PLAIN, 0.f, nullptr, st, lane);
    HIP_CHECK(hipEventRecord(e1, ctx().stream));
    HIP_CHECK(hipEventSynchronize(e1));
    float f = 0;
    HIP_CHECK(hipEventElapsedTime(&f, e0, e1));
    ms = (double)f / reps;
    HIP_CHECK(hipEventDestroy(e0));
    HIP_CHECK(hipEventDestroy(e1));
  } else {
    mrhs_dslash<NR>(*L, cb, pout, pin, EPI_PLAIN, 0.f, nullptr, st, lane);
  }
  hipLaunchKernelGGL(k_mrhs_unpack, dim3(nbt), dim3(BLAS_BLOCK), 0, ctx().stream, ptrs + NR, pout, vh, n, NR);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(ctx().stream));
  return ms;
}

static double dslash_multi_api(QPHIX_F3_FermionLinksAsqtad *L, QPHIX_F3_ColorVector **out,
                               QPHIX_F3_ColorVector **in, int n, int parity, int reps) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called: asqtad_dslash_multi", -1.0);
  QPHIX_REQUIRE(L && out && in && n >= 1 && n <= 8, "asqtad_dslash_multi: 1..8 vectors", -1.0);
  QPHIX_REQUIRE(parity == QPHIX_EVEN || parity == QPHIX_ODD, "asqtad_dslash_multi: parity must be EVEN or ODD", -1.0);
  const int cb = parity == QPHIX_EVEN ? 0 : 1;
  QPHIX_REQUIRE(L->fat[cb], "asqtad_dslash_multi: links of the output parity missing", -1.0);
  for (int r = 0; r < n; ++r)
    QPHIX_REQUIRE(out[r] && in[r] && out[r]->half[cb] && in[r]->half[1 - cb],
                  "asqtad_dslash_multi: vectors lack the needed halves", -1.0);
  const MrhsTune &tune = autotune_mrhs(*L);
  if (tune.nr == 0) return -1.0;
  const int nr = n == 1 ? 1 : n == 2 ? 2 : n <= 4 ? 4 : 8;
  const bool lane = tune.lane[nr];
  if (nr == 1) return dslash_multi<1>(L, out, in, n, cb, reps, lane);
  if (nr == 2) return dslash_multi<2>(L, out, in, n, cb, reps, lane);
  if (nr == 4) return dslash_multi<4>(L, out, in, n, cb, reps, lane);
  return dslash_multi<8>(L, out, in, n, cb, reps, lane);
}

} // namespace qphix

using namespace qphix;

extern "C" {

void QPHIX_F3_asqtad_dslash_multi(QPHIX_F3_FermionLinksAsqtad *L, QPHIX_F3_ColorVector *out[],
                                  QPHIX_F3_ColorVector *in[], int nvec, QPHIX_evenodd_t parity) {
  dslash_multi_api(L, out, in, nvec, parity, 0);
}

int QPHIX_F3_asqtad_mrhs_tune(QPHIX_F3_FermionLinksAsqtad *L, double *us) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called: asqtad_mrhs_tune", 0);
  QPHIX_REQUIRE(L && L->fat[0] && L->fat[1], "asqtad_mrhs_tune: both link parities are required", 0);
  const MrhsTune &t = autotune_mrhs(*L);
  if (us) {
    const int nrs[4] = {1, 2, 4, 8};
    for (int c = 0; c < 4; ++c)
      for (int lane = 0; lane < 2; ++lane) us[2 * c + lane] = t.us[nrs[c]][lane];
  }
  return t.nr == 0 ? 0 : t.nr * 2 + (t.lane[t.nr] ? 1 : 0);
}

double QPHIX_F3_asqtad_dslash_multi_time(QPHIX_F3_FermionLinksAsqtad *L, QPHIX_F3_ColorVector *out[],
                                         QPHIX_F3_ColorVector *in[], int nvec,
                                         QPHIX_evenodd_t parity, int reps) {
  return dslash_multi_api(L, out, in, nvec, parity, reps < 1 ? 1 : reps);
}

int QPHIX_D3_asqtad_invert_multi_rhs(QPHIX_info_t *info, QPHIX_D3_FermionLinksAsqtad *L,
                                     QPHIX_invert_arg_t *inv, QPHIX_resid_arg_t *res[],
                                     double *mass, int nshift, int nrhs, QPHIX_D3_ColorVector *out[],
                                     QPHIX_D3_ColorVector *in[], int rhs_block) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called: asqtad_invert_multi_rhs", 0);
  QPHIX_REQUIRE(info && L && inv && res && mass && out && in, "asqtad_invert_multi_rhs: null argument", 0);
  info->status = QPHIX_FAIL;
  QPHIX_REQUIRE(nshift > 0 && nshift <= MAXNDMASS && nrhs > 0, "asqtad_invert_multi_rhs: bad counts", 0);
  QPHIX_REQUIRE(inv->max > 0 && inv->nrestart > 0, "asqtad_invert_multi_rhs: bad iteration limits", 0);
  QPHIX_REQUIRE(L->fat[0] && L->fat[1], "asqtad_invert_multi_rhs: both link parities are required", 0);
  for (int r = 0; r < nrhs; ++r) {
    QPHIX_REQUIRE(in[r], "asqtad_invert_multi_rhs: null source", 0);
    for (int cb = 0; cb < 2; ++cb) {
      if (!(inv->parity & (cb ? QPHIX_ODD : QPHIX_EVEN))) continue;
      QPHIX_REQUIRE(in[r]->half[cb], "asqtad_invert_multi_rhs: source lacks a solved parity", 0);
      for (int j = 0; j < nshift; ++j)
        QPHIX_REQUIRE(out[(size_t)r * nshift + j] && out[(size_t)r * nshift + j]->half[cb],
                      "asqtad_invert_multi_rhs: solution lacks a solved parity", 0);
    }
    for (int j = 0; j < nshift; ++j) QPHIX_REQUIRE(res[(size_t)r * nshift + j], "asqtad_invert_multi_rhs: null res_arg", 0);
  }
  QPHIX_REQUIRE(rhs_block == 0 || rhs_block == 1 || rhs_block == 2 || rhs_block == 4 || rhs_block == 8,
                "asqtad_invert_multi_rhs: rhs_block must be 0 (autotune), 1, 2, 4 or 8", 0);
  if (!L->sloppy) {
    L->sloppy = new LinksT<float>;
    links_to_float(*L, *L->sloppy);
  }
  const MrhsTune &tune = autotune_mrhs(*L->sloppy);
  if (tune.nr == 0) return 0;                 // workspace allocation failed (error recorded)
  int nb = rhs_block > 0 ? rhs_block : tune.nr;
  while (nb > 1 && nb / 2 >= nrhs) nb /= 2;   // no block wider than the sources need
  const bool lane = tune.lane[nb];
  switch (nb) {
  case 1: return invert_multi_rhs<1>(info, L, inv, res, mass, nshift, nrhs, out, in, lane);
  case 2: return invert_multi_rhs<2>(info, L, inv, res, mass, nshift, nrhs, out, in, lane);
  case 4: return invert_multi_rhs<4>(info, L, inv, res, mass, nshift, nrhs, out, in, lane);
  default: return invert_multi_rhs<8>(info, L, inv, res, mass, nshift, nrhs, out, in, lane);
  }
}

} // extern "C"
