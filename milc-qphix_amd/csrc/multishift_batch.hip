// multishift_batch.hip -- multi-shift CG on several right-hand sides at once:
// an FP64 first phase, then FP32 iteration with FP64 reliable updates, then
// per-shift FP64-checked refinement.
//
// No counterpart in the reference, whose multi-shift solver
// (ks_multicg_offset.cpp:205-689) takes one source in one precision.  Per
// right-hand side this is the same Jegerlehner recursion (relative shifts,
// base system = smallest shift, zeta/beta/alpha recurrences, lazy solution
// updates), with:
//  * up to NR = 8 sources advancing together: one multi-RHS Dslash per
//    half-step (links read once for all NR), one fused update of r and one of
//    every (x_j, p_j) per iteration, each source with its own device-side
//    scalars and gate; shifts retire once their zeta-scaled residual meets the
//    target;
//  * precision: FP64 vectors and links until every source's |r| is below
//    QPHIX_MSX_SWITCH (default 3e-4) times |b| -- the shifted systems cannot
//    be corrected later, and the rounding of the large early steps is what
//    limits their accuracy -- then FP32 vectors and links, with FP64 solutions
//    x_j flushed from FP32 accumulators at every reliable update;
//  * reliable updates on the base system: when |r| has dropped 10x (or looks
//    converged, or the step budget is spent) that source pauses, its base true
//    residual b - (M + sigma_low) x_low is recomputed with the FP64 Dslash and
//    replaces r (residual replacement); it stops once that true residual meets
//    the target;
//  * refinement: every shifted solution is then checked in FP64 and, where its
//    own true residual misses the target, finished by the same batched solver
//    on its correction equation, sources of one shift together.
#include "blas.h"
#include "dslash_mrhs.h"
#include "handles.h"
#include "qphix_mi355x.h"

#include <algorithm>
#include <chrono>
#include <map>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

namespace qphix {

template <int NR>
__global__ void __launch_bounds__(FINALIZE_BLOCK)
    k_finalize_multi(const double *partials, int nb, SolverState *st, int op, int phase) {
  double s[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    s[r] = 0;
    for (int b = threadIdx.x; b < nb; b += FINALIZE_BLOCK) s[r] += partials[(size_t)r * nb + b];
  }
  block_sum<NR>(s);  // totals in thread 0
  // one thread per source runs its scalar recurrences (a chain of dependent
  // state reads and writes each; in one thread for all sources they queued)
  __shared__ double tot[NR];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r) tot[r] = s[r];
  }
  __syncthreads();
  const int r = threadIdx.x;
  if (r < NR && st[r].active) {
    st[r].phase = phase;
    apply_scalar_op(st + r, op, tot + r);
  }
}

// r_r += beta_low,r Ap_r ; reduce |r_r|^2 per source.  One work-item per
// (site, source), resident grid walking XCD slabs of nlog logical blocks;
// one partial per block and source.
template <typename T, int NR>
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_msx_update_r(T *rs, const T *Ap, int vh, const SolverState *st, double *partials, int nlog) {
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
    const T bl = (T)st[r].beta_i[st[r].j_low];
    cplx<T> rv[3], av[3];
    Spin<T>::load(rs, t, rv);
    Spin<T>::load(Ap, t, av);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      rv[c].re += bl * av[c].re;
      rv[c].im += bl * av[c].im;
      v += (double)rv[c].re * rv[c].re + (double)rv[c].im * rv[c].im;
    }
    Spin<T>::store(rs, t, rv);
  }
  const double sum = lane_block_sum<NR, BLAS_BLOCK>(v);
  if (threadIdx.x < NR) partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = sum;
}

// One step of shift j: x += beta p ; p = zeta r + alpha p (explicit fmas: the
// single-step and grouped paths below round identically).
template <typename T>
__device__ __forceinline__ void msx_one(cplx<T> *xv, cplx<T> *pv, const cplx<T> *rv, T bj, T aj, T zj) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    xv[c].re = fma(bj, pv[c].re, xv[c].re);
    xv[c].im = fma(bj, pv[c].im, xv[c].im);
    pv[c].re = fma(zj, rv[c].re, aj * pv[c].re);
    pv[c].im = fma(zj, rv[c].im, aj * pv[c].im);
  }
}

// x_j += beta_j p_j ; p_j = zeta_j r + alpha_j p_j  for every shift j < num_now
// not yet retired, one work-item per (site, source); layout
// [shift][site][source][6].  Only the base system's p feeds the operator, so
// the shifted systems go up to st.fuse steps at a time: on a deferred step
// (st.defer) only the base system is updated and r is saved in rprev slot
// st.pend; on the next step that is not deferred each shifted system does
// every pending step and then this one in registers, reading and writing x_j
// and p_j once for all of them (same arithmetic and order as single steps).
#ifndef MSX_GROUP
#define MSX_GROUP 2
#endif
// Per-block copy of the coefficients the update reads, for every source r of
// the chunk and shift j: (beta, alpha, zeta) of the pending steps 0 .. 2 and of
// this step, and sdone[j] -- read from LDS instead of each lane loading its own
// source's state for every shift.
constexpr int MSX_NC = 3 * MSX_MAXDEFER + 1;
constexpr int MSX_STAGE_BYTES = 48 * 1024;  // larger (NR x nshift): coefficients read from the states
template <int NR> inline bool msx_staged(int nshift) {
  return (size_t)NR * nshift * MSX_NC * sizeof(double) <= (size_t)MSX_STAGE_BYTES;
}

// Blocks of k_msx_update_pm: as many as are resident at once (occupancy API x
// CUs), each stepping through the slots, so the coefficient copy is made once
// per resident block.
template <typename K> inline int msx_pm_grid(K kernel, int nslots, size_t lds) {
  int bpc = 0;
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, kernel, BLAS_BLOCK, lds));
  if (bpc < 1) bpc = 1;
  const int need = (nslots + BLAS_BLOCK - 1) / BLAS_BLOCK;
  const int res = bpc * ctx().n_cu;
  return need < res ? need : res;
}

template <typename T, int NR, bool STAGED>
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_msx_update_pm(T *xs, T *p, const T *rs, T *rprev, int vh, int nshift, const SolverState *st) {
  extern __shared__ double msx_cf[];  // STAGED: [NR][nshift][MSX_NC]
  if constexpr (STAGED) {
    const int n = NR * nshift * MSX_NC;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      const int rr = k / (nshift * MSX_NC), rem = k - rr * nshift * MSX_NC;
      const int j = rem / MSX_NC, c = rem - j * MSX_NC;
      const SolverState &sr = st[rr];
      const int q = c / 3, w = c - 3 * q;
      double v;
      if (c == MSX_NC - 1) v = (double)sr.sdone[j];
      else if (q < MSX_MAXDEFER - 1) v = w == 0 ? sr.pbeta[q][j] : (w == 1 ? sr.palpha[q][j] : sr.pzeta[q][j]);
      else v = w == 0 ? sr.beta_i[j] : (w == 1 ? sr.alpha[j] : sr.zeta_ip1[j]);
      msx_cf[k] = v;
    }
    __syncthreads();
  }
  // grid-stride over the (site, source) slots: the stride is a multiple of NR,
  // so a work-item's source r is the same for all its slots
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = t0 % NR;
  const SolverState &s = st[r];
  if (!s.active) return;
  // coefficient w (0 beta, 1 alpha, 2 zeta) of shift j for pending step q
  // (q = MSX_MAXDEFER - 1: this step)
  auto coef = [&](int q, int j, int w) -> T {
    if constexpr (STAGED) return (T)msx_cf[((size_t)r * nshift + j) * MSX_NC + 3 * q + w];
    else if (q < MSX_MAXDEFER - 1) return (T)(w == 0 ? s.pbeta[q][j] : (w == 1 ? s.palpha[q][j] : s.pzeta[q][j]));
    else return (T)(w == 0 ? s.beta_i[j] : (w == 1 ? s.alpha[j] : s.zeta_ip1[j]));
  };
  auto retired = [&](int j) -> bool {
    if constexpr (STAGED) return msx_cf[((size_t)r * nshift + j) * MSX_NC + MSX_NC - 1] == 1.0;
    else return s.sdone[j] == 1;
  };
  constexpr int NOW = MSX_MAXDEFER - 1;
  const size_t stride = (size_t)vh * NR;
  const int nn = s.num_now, jl = s.j_low;
  const int np = s.pend;
  // num_now only shrinks, so the first pending step covers every shift of the later ones
  const int lim = min(np ? s.pend_nn[0] : nn, nshift);
  for (int t = t0; t < vh * NR; t += gridDim.x * blockDim.x) {
  cplx<T> rv[3];
  Spin<T>::load(rs, t, rv);
  auto step1 = [&](int j) {
    cplx<T> xv[3], pv[3];
    Spin<T>::load(xs + 6 * stride * j, t, xv);
    Spin<T>::load(p + 6 * stride * j, t, pv);
    msx_one<T>(xv, pv, rv, coef(NOW, j, 0), coef(NOW, j, 1), coef(NOW, j, 2));
    Spin<T>::store(xs + 6 * stride * j, t, xv);
    Spin<T>::store(p + 6 * stride * j, t, pv);
  };
  if (s.defer) {
    Spin<T>::store(rprev + 6 * stride * s.pend, t, rv);
    step1(jl);
    continue;
  }
  // Shifts MSX_GROUP at a time: the group's x_j, p_j loads are issued before
  // any store, so a work-item's shifts are not one dependent load-store chain.
  cplx<T> rq[MSX_MAXDEFER - 1][3];
#pragma unroll
  for (int q = 0; q < MSX_MAXDEFER - 1; ++q)
    if (q < np) Spin<T>::load(rprev + 6 * stride * q, t, rq[q]);
  for (int j0 = 0; j0 < lim; j0 += MSX_GROUP) {
    bool on[MSX_GROUP];
    cplx<T> xv[MSX_GROUP][3], pv[MSX_GROUP][3];
#pragma unroll
    for (int g = 0; g < MSX_GROUP; ++g) {
      const int j = j0 + g;
      on[g] = j < lim && !retired(j);  // retired earlier: no update
      if (on[g]) {
        Spin<T>::load(xs + 6 * stride * j, t, xv[g]);
        Spin<T>::load(p + 6 * stride * j, t, pv[g]);
      }
    }
#pragma unroll
    for (int g = 0; g < MSX_GROUP; ++g) {
      if (!on[g]) continue;
      const int j = j0 + g;
      if (j != jl) {  // the pending steps in order (the base system took them already)
#pragma unroll
        for (int q = 0; q < MSX_MAXDEFER - 1; ++q)
          if (q < np && j < s.pend_nn[q]) msx_one<T>(xv[g], pv[g], rq[q], coef(q, j, 0), coef(q, j, 1), coef(q, j, 2));
      }
      if (j < nn || j == jl)  // then this one
        msx_one<T>(xv[g], pv[g], rv, coef(NOW, j, 0), coef(NOW, j, 1), coef(NOW, j, 2));
      Spin<T>::store(xs + 6 * stride * j, t, xv[g]);
      Spin<T>::store(p + 6 * stride * j, t, pv[g]);
    }
  }
  }
}

// End of a step: sources that asked for a true-residual check pause, and
// shifts that took their last update retire.
__global__ void k_msx_commit(SolverState *st, int nr) {
  const int r = threadIdx.x;
  if (r >= nr || !st[r].active) return;
  SolverState &s = st[r];
  for (int j = 0; j < s.num_now; ++j)
    if (s.sdone[j] == 2) s.sdone[j] = 1;
  if (s.defer) {  // the shifted updates of this step wait for a later one
    const int q = s.pend++;
    s.pend_nn[q] = s.num_now;
    for (int j = 0; j < s.num_now; ++j) {
      s.pbeta[q][j] = s.beta_i[j];
      s.palpha[q][j] = s.alpha[j];
      s.pzeta[q][j] = s.zeta_ip1[j];
    }
  } else {
    s.pend = 0;
  }
  s.defer = 0;
  if (s.ru) s.active = 0;
}

// Source slot r: r_s = p_j = (T) b for all j, x_s = 0.
template <typename T>
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_msx_init(T *rs, T *p, T *xs, const double *b, int vh, int NR, int r, int nshift) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  const size_t stride = (size_t)vh * NR;
  const int t = i * NR + r;
  cplx<double> bd[3];
  Spin<double>::load(b, i, bd);
  cplx<T> v[3];
  const cplx<T> z[3] = {{0, 0}, {0, 0}, {0, 0}};
  for (int c = 0; c < 3; ++c) v[c] = cplx<T>{(T)bd[c].re, (T)bd[c].im};
  Spin<T>::store(rs, t, v);
  for (int j = 0; j < nshift; ++j) {
    Spin<T>::store(p + 6 * stride * j, t, v);
    Spin<T>::store(xs + 6 * stride * j, t, z);
  }
}

// x64_j += x_s[j] (slot r); x_s[j] = 0
template <typename T>
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_msx_flush(double *const *x64, T *xs, int vh, int NR, int r, int nshift) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  const size_t stride = (size_t)vh * NR;
  const int t = i * NR + r;
  const cplx<T> z[3] = {{0, 0}, {0, 0}, {0, 0}};
  for (int j = 0; j < nshift; ++j) {
    cplx<T> sv[3];
    cplx<double> dv[3];
    Spin<T>::load(xs + 6 * stride * j, t, sv);
    Spin<double>::load(x64[j], i, dv);
    for (int c = 0; c < 3; ++c) {
      dv[c].re += (double)sv[c].re;
      dv[c].im += (double)sv[c].im;
    }
    Spin<double>::store(x64[j], i, dv);
    Spin<T>::store(xs + 6 * stride * j, t, z);
  }
}

// r_s (slot r) = (float) r64
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_msx_set_r(float *rs, const double *r64, int vh, int NR, int r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  cplx<double> rd[3];
  Spin<double>::load(r64, i, rd);
  cplx<float> v[3];
  for (int c = 0; c < 3; ++c) v[c] = cplx<float>{(float)rd[c].re, (float)rd[c].im};
  Spin<float>::store(rs, i * NR + r, v);
}

// Resume a paused source with the FP64 true residual's norm; delta > 0 also
// sets the next update's drop factor.
__global__ void k_msx_resume(SolverState *s, const SolverState *t, double delta) {
  s->rsq = t->rsq;
  s->mx_maxr = sqrt(t->rsq);
  s->ru = 0;
  s->active = 1;
  if (delta > 0) s->mx_delta = delta;
}

// FP64 -> FP32 copy of n consecutive spinor slots (the direction vectors at
// the end of the FP64 phase), slot by slot: the FP64 and FP32 tiles order
// their reals differently.
__global__ void __launch_bounds__(BLAS_BLOCK) k_msx_demote(float *dst, const double *src, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cplx<double> d[3];
  Spin<double>::load(src, (int)i, d);
  cplx<float> f[3];
  for (int c = 0; c < 3; ++c) f[c] = cplx<float>{(float)d[c].re, (float)d[c].im};
  Spin<float>::store(dst, (int)i, f);
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

// QPHIX_MRHS_K (1, 2 or 4), when set, replaces the tuned sources-per-work-item.
static int mrhs_k_override(int tuned) {
  const char *e = std::getenv("QPHIX_MRHS_K");
  if (!e) return tuned;
  const int k = std::atoi(e);
  return k == 1 || k == 2 || k == 4 ? k : tuned;
}

// Logical 256-item blocks of a multi-RHS launch with K sources per work-item.
inline int mrhs_nlog(int vh, int NR, int K) {
  return (int)(((size_t)vh * (NR / K) + DSLASH_BLOCK - 1) / DSLASH_BLOCK);
}

// One source per work-item with NR >= 2 lanes per site uses the LDS-staged
// link kernel (k_dslash_mrhs_lds; bit-identical results) unless switched off
// (QPHIX_MI355X_set_mrhs_lds(0), or QPHIX_MRHS_LDS=0 before the first call).
static int g_mrhs_lds = -1;
// Steps per update of the shifted systems in the batched multi-shift solver
// (k_msx_update_pm; identical results for every value): QPHIX_MI355X_set_msx_fuse.
static int g_msx_fuse = MSX_MAXDEFER;
int msx_fuse() { return g_msx_fuse; }
static bool mrhs_lds_on() {
  if (g_mrhs_lds < 0) {
    const char *e = std::getenv("QPHIX_MRHS_LDS");
    g_mrhs_lds = (e && std::atoi(e) == 0) ? 0 : 1;
  }
  return g_mrhs_lds == 1;
}

// K sources per work-item with links staged through LDS: K = 1 (NR >= 2,
// k_dslash_mrhs_lds) and K = 2, 4 (NR >= 2 K, k_dslash_mrhs_ldsk).
template <int NR, int K> constexpr bool mrhs_has_lds() { return (K == 1 && NR >= 2) || (K >= 2 && NR >= 2 * K); }

// Resident grid of the kernel a launch uses (cached per kernel; the lattice is
// fixed between QPHIX_init and QPHIX_finalize).
template <typename T, int NR, int K, int NL, int EPI> static int mrhs_grid_of(int vh) {
  const int v = (mrhs_has_lds<NR, K>() && mrhs_lds_on()) ? 1 : 0;
  static int g[2] = {0, 0};
  static int g_vh[2] = {-1, -1};
  if (g_vh[v] != vh) {
    if constexpr (K == 1 && NR >= 2) {
      if (v) g[v] = resident_grid(k_dslash_mrhs_lds<T, NL, NR, EPI>, mrhs_nlog(vh, NR, K));
      else g[v] = resident_grid(k_dslash_mrhs<T, NL, NR, K, EPI>, mrhs_nlog(vh, NR, K));
    } else if constexpr (K >= 2 && NR >= 2 * K) {
      if (v) g[v] = resident_grid(k_dslash_mrhs_ldsk<T, NL, NR, K, EPI>, mrhs_nlog(vh, NR, K));
      else g[v] = resident_grid(k_dslash_mrhs<T, NL, NR, K, EPI>, mrhs_nlog(vh, NR, K));
    } else {
      g[v] = resident_grid(k_dslash_mrhs<T, NL, NR, K, EPI>, mrhs_nlog(vh, NR, K));
    }
    g_vh[v] = vh;
  }
  return g[v];
}

template <typename T, int NR, int K> static int mrhs_blocks_k(int vh, int recon, int epi) {
  if (recon == 12)
    return epi == EPI_PLAIN ? mrhs_grid_of<T, NR, K, 12, EPI_PLAIN>(vh) : mrhs_grid_of<T, NR, K, 12, EPI_PKP>(vh);
  return epi == EPI_PLAIN ? mrhs_grid_of<T, NR, K, 18, EPI_PLAIN>(vh) : mrhs_grid_of<T, NR, K, 18, EPI_PKP>(vh);
}

// Blocks of a multi-RHS Dslash launch (for EPI_PKP, the partial count per source).
template <int NR, typename T = float> static int mrhs_blocks(int vh, int K, int recon, int epi) {
  if (K >= 4 && NR % 4 == 0) return mrhs_blocks_k<T, NR, (NR % 4 == 0 ? 4 : 1)>(vh, recon, epi);
  if (K >= 2 && NR % 2 == 0) return mrhs_blocks_k<T, NR, (NR % 2 == 0 ? 2 : 1)>(vh, recon, epi);
  return mrhs_blocks_k<T, NR, 1>(vh, recon, epi);
}

template <typename T, int NR, int K>
static void mrhs_launch(const LinksT<T> &L, int epi, MrhsArgsT<T> a) {
  const int vh = a.g.vh;
  a.nlog = mrhs_nlog(vh, NR, K);
  if constexpr (K >= 2 && NR >= 2 * K) {
    if (mrhs_lds_on()) {
      const dim3 blk(DSLASH_BLOCK);
      if (L.recon == 12) {
        if (epi == EPI_PLAIN)
          hipLaunchKernelGGL((k_dslash_mrhs_ldsk<T, 12, NR, K, EPI_PLAIN>), dim3(mrhs_grid_of<T, NR, K, 12, EPI_PLAIN>(vh)),
                             blk, 0, ctx().stream, a);
        else
          hipLaunchKernelGGL((k_dslash_mrhs_ldsk<T, 12, NR, K, EPI_PKP>), dim3(mrhs_grid_of<T, NR, K, 12, EPI_PKP>(vh)),
                             blk, 0, ctx().stream, a);
      } else {
        if (epi == EPI_PLAIN)
          hipLaunchKernelGGL((k_dslash_mrhs_ldsk<T, 18, NR, K, EPI_PLAIN>), dim3(mrhs_grid_of<T, NR, K, 18, EPI_PLAIN>(vh)),
                             blk, 0, ctx().stream, a);
        else
          hipLaunchKernelGGL((k_dslash_mrhs_ldsk<T, 18, NR, K, EPI_PKP>), dim3(mrhs_grid_of<T, NR, K, 18, EPI_PKP>(vh)),
                             blk, 0, ctx().stream, a);
      }
      return;
    }
  }
  if constexpr (K == 1 && NR >= 2) {
    if (mrhs_lds_on()) {
      const dim3 blk(DSLASH_BLOCK);
      if (L.recon == 12) {
        if (epi == EPI_PLAIN)
          hipLaunchKernelGGL((k_dslash_mrhs_lds<T, 12, NR, EPI_PLAIN>), dim3(mrhs_grid_of<T, NR, K, 12, EPI_PLAIN>(vh)),
                             blk, 0, ctx().stream, a);
        else
          hipLaunchKernelGGL((k_dslash_mrhs_lds<T, 12, NR, EPI_PKP>), dim3(mrhs_grid_of<T, NR, K, 12, EPI_PKP>(vh)),
                             blk, 0, ctx().stream, a);
      } else {
        if (epi == EPI_PLAIN)
          hipLaunchKernelGGL((k_dslash_mrhs_lds<T, 18, NR, EPI_PLAIN>), dim3(mrhs_grid_of<T, NR, K, 18, EPI_PLAIN>(vh)),
                             blk, 0, ctx().stream, a);
        else
          hipLaunchKernelGGL((k_dslash_mrhs_lds<T, 18, NR, EPI_PKP>), dim3(mrhs_grid_of<T, NR, K, 18, EPI_PKP>(vh)),
                             blk, 0, ctx().stream, a);
      }
      return;
    }
  }
  if (L.recon == 12) {
    if (epi == EPI_PLAIN)
      hipLaunchKernelGGL((k_dslash_mrhs<T, 12, NR, K, EPI_PLAIN>), dim3(mrhs_grid_of<T, NR, K, 12, EPI_PLAIN>(vh)),
                         dim3(DSLASH_BLOCK), 0, ctx().stream, a);
    else
      hipLaunchKernelGGL((k_dslash_mrhs<T, 12, NR, K, EPI_PKP>), dim3(mrhs_grid_of<T, NR, K, 12, EPI_PKP>(vh)),
                         dim3(DSLASH_BLOCK), 0, ctx().stream, a);
  } else {
    if (epi == EPI_PLAIN)
      hipLaunchKernelGGL((k_dslash_mrhs<T, 18, NR, K, EPI_PLAIN>), dim3(mrhs_grid_of<T, NR, K, 18, EPI_PLAIN>(vh)),
                         dim3(DSLASH_BLOCK), 0, ctx().stream, a);
    else
      hipLaunchKernelGGL((k_dslash_mrhs<T, 18, NR, K, EPI_PKP>), dim3(mrhs_grid_of<T, NR, K, 18, EPI_PKP>(vh)),
                         dim3(DSLASH_BLOCK), 0, ctx().stream, a);
  }
}

// Traversal order of the multi-RHS kernels.  Default (nz a multiple of 8):
// XCD x takes the z-planes [x nz / 8, (x + 1) nz / 8) and walks them t-slice
// by t-slice (SiteOrder on = 2), so the 8 XCDs advance through t together and
// the +-3 t-slice gather windows stay small enough for the Infinity Cache.
// Measured at 32^3x64, 8 sources x 12 shifts (batched solve 3.13 -> 3.00 s):
// FP64 plain / p.Ap Dslash 898 / 943 -> 797 / 810 us, FP32 350 -> 346 us.
// QPHIX_SITE_ORDER=t: each XCD sweeps a t-slab in site order (the single-
// source kernels' order); QPHIX_SITE_ORDER="py,pz": (py x pz) patches through
// each XCD's t-slab (fewer bytes fetched, no faster: it breaks the sequential
// link stream).  Every order is a permutation of which lane computes which
// site; per-block partial sums (and so rounding) follow the order.
static SiteOrder site_order(int NR) {
  const Geometry &g = ctx().g;
  SiteOrder o{0, 4, 4, 0};
  (void)NR;
  const char *e = std::getenv("QPHIX_SITE_ORDER");
  if (!e || std::strcmp(e, "z") == 0) {
    if (g.nz % N_XCD == 0) o.on = 2;
    return o;
  }
  if (std::strchr(e, ',') == nullptr) return o;
  std::sscanf(e, "%d,%d", &o.py, &o.pz);
  if (o.py < 1 || o.pz < 1 || g.ny % o.py || g.nz % o.pz || g.nt % N_XCD) return o;
  o.tslab = g.nt / N_XCD;
  o.on = 1;
  return o;
}

// Multi-RHS Dslash on the [vh][NR][6] interleaved layout, K sources per
// work-item (1, 2 or 4; reduced to a divisor of NR); EPI_PKP leaves
// NR x mrhs_blocks<NR, T>(vh, K, ...) partial sums.
template <int NR, typename T>
static void mrhs_dslash(const LinksT<T> &L, int cb, T *out, const T *in, int epi, T s, const T *aux,
                        const SolverState *st, int K) {
  const Geometry &gg = ctx().g;
  MrhsArgsT<T> a{};
  a.fat = reinterpret_cast<const typename vec16<T>::type *>(L.fat[cb]);
  a.lng = reinterpret_cast<const typename vec16<T>::type *>(L.lng[cb]);
  a.lsign = L.lsign[cb];
  a.in = in;
  a.out = out;
  a.aux = aux;
  a.g = Geo{gg.nxh, gg.ny, gg.nz, gg.nt, gg.vh, gg.nxy, gg.nxyz};
  a.cb = cb;
  a.s = s;
  a.partials = ctx().partials;
  a.states = st;
  a.order = site_order(NR);
  if (K >= 4 && NR % 4 == 0) mrhs_launch<T, NR, (NR % 4 == 0 ? 4 : 1)>(L, epi, a);
  else if (K >= 2 && NR % 2 == 0) mrhs_launch<T, NR, (NR % 2 == 0 ? 2 : 1)>(L, epi, a);
  else mrhs_launch<T, NR, 1>(L, epi, a);
  HIP_CHECK(hipGetLastError());
}

static double wall_now() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// FP64 true residual of one source from its D D x (dd, overwritten with r):
// r = b + (dd + s x), the arithmetic and order of the Dslash's restart
// epilogue; |r|^2 and the relative-residue term reduced through rc.
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_msx_true_res(double *dd, const double *x, const double *b, double s, int vh, ReduceCtl rc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < vh) {
    cplx<double> acc[3], xv[3], bv[3];
    Spin<double>::load(dd, i, acc);
    Spin<double>::load(x, i, xv);
    Spin<double>::load(b, i, bv);
    double num = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double tr = acc[c].re + s * xv[c].re;
      const double ti = acc[c].im + s * xv[c].im;
      acc[c].re = bv[c].re + tr;
      acc[c].im = bv[c].im + ti;
      num += acc[c].re * acc[c].re + acc[c].im * acc[c].im;
    }
    Spin<double>::store(dd, i, acc);
    v[0] = num;
    v[1] = rel_term(site_norm2_rounded(acc), site_norm2_rounded(xv));
  }
  grid_reduce<2>(v, rc);
}

#include "multishift_driver.h"

// ----------------------------------------------------------------- multi-RHS Dslash API
__global__ void k_mrhs_pack(float *dst, const float *const *src, int vh, int n, int NR) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= vh * NR) return;
  const int i = t / NR, r = t % NR;
  cplx<float> v[3] = {{0, 0}, {0, 0}, {0, 0}};
  if (r < n) Spin<float>::load(src[r], i, v);
  Spin<float>::store(dst, t, v);
}

__global__ void k_mrhs_unpack(float *const *dst, const float *src, int vh, int n, int NR) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= vh * NR) return;
  const int i = t / NR, r = t % NR;
  if (r >= n) return;
  cplx<float> v[3];
  Spin<float>::load(src, t, v);
  Spin<float>::store(dst[r], i, v);
}

// Sources passed by value (no device pointer array to stage).
template <typename T> struct PackSrc {
  const T *p[8];
};

template <typename T>
__global__ void k_mrhs_pack_args(T *dst, PackSrc<T> src, int vh, int n, int NR) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= vh * NR) return;
  const int i = t / NR, r = t % NR;
  cplx<T> v[3] = {{0, 0}, {0, 0}, {0, 0}};
  if (r < n) {
    const T *s = src.p[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) s = r == k ? src.p[k] : s;  // selects, no indexed private array
    Spin<T>::load(s, i, v);
  }
  Spin<T>::store(dst, t, v);
}

__global__ void k_mrhs_gates(SolverState *st, int n, int nr) {
  const int r = threadIdx.x;
  if (r < nr) st[r].active = r < n ? 1 : 0;
}

template <typename T>
int mrhs_dslash_packed(const LinksT<T> &L, int cb, const T *const *src, int n, T *pin, T *pout, SolverState *st) {
  const int vh = ctx().g.vh;
  const int nr = n <= 2 ? 2 : (n <= 4 ? 4 : 8);
  PackSrc<T> ps{};
  for (int r = 0; r < n; ++r) ps.p[r] = src[r];
  hipLaunchKernelGGL(k_mrhs_gates, dim3(1), dim3(64), 0, ctx().stream, st, n, nr);
  hipLaunchKernelGGL(k_mrhs_pack_args<T>, dim3((vh * nr + BLAS_BLOCK - 1) / BLAS_BLOCK), dim3(BLAS_BLOCK), 0,
                     ctx().stream, pin, ps, vh, n, nr);
  if (nr == 2) mrhs_dslash<2, T>(L, cb, pout, pin, EPI_PLAIN, (T)0, nullptr, st, 1);
  else if (nr == 4) mrhs_dslash<4, T>(L, cb, pout, pin, EPI_PLAIN, (T)0, nullptr, st, 1);
  else mrhs_dslash<8, T>(L, cb, pout, pin, EPI_PLAIN, (T)0, nullptr, st, 1);
  return nr;
}
template <typename T>
__global__ void k_mrhs_unpack_strided(T *dst, size_t nv, const T *src, int vh, int n, int NR) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= vh * NR) return;
  const int i = t / NR, r = t % NR;
  if (r >= n) return;
  cplx<T> v[3];
  Spin<T>::load(src, t, v);
  Spin<T>::store(dst + (size_t)r * nv, i, v);
}

template <typename T> void mrhs_unpack(const T *src, T *dst, int n, int nr) {
  const int vh = ctx().g.vh;
  hipLaunchKernelGGL(k_mrhs_unpack_strided<T>, dim3((vh * nr + BLAS_BLOCK - 1) / BLAS_BLOCK), dim3(BLAS_BLOCK), 0,
                     ctx().stream, dst, 6 * (size_t)vh, src, vh, n, nr);
  HIP_CHECK(hipGetLastError());
}
template void mrhs_unpack<double>(const double *, double *, int, int);
template void mrhs_unpack<float>(const float *, float *, int, int);

template int mrhs_dslash_packed<double>(const LinksT<double> &, int, const double *const *, int, double *, double *,
                                        SolverState *);
template int mrhs_dslash_packed<float>(const LinksT<float> &, int, const float *const *, int, float *, float *,
                                       SolverState *);

template <int NR>
static double dslash_multi(LinksT<float> *L, QPHIX_F3_ColorVector **out, QPHIX_F3_ColorVector **in,
                           int n, int cb, int reps, int K) {
  const int vh = ctx().g.vh;
  const size_t nvr = 6 * (size_t)vh * NR;
  char *w = (char *)ws2_get(2 * nvr * sizeof(float) + NR * sizeof(SolverState) + 2 * NR * sizeof(float *) + 1024);
  if (!w) return -1.0;  // out of device memory, error recorded
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
    for (int k = 0; k < reps; ++k) mrhs_dslash<NR, float>(*L, cb, pout, pin, EPI_PLAIN, 0.f, nullptr, st, K);
    HIP_CHECK(hipEventRecord(e1, ctx().stream));
    HIP_CHECK(hipEventSynchronize(e1));
    float f = 0;
    HIP_CHECK(hipEventElapsedTime(&f, e0, e1));
    ms = (double)f / reps;
    HIP_CHECK(hipEventDestroy(e0));
    HIP_CHECK(hipEventDestroy(e1));
  } else {
    mrhs_dslash<NR, float>(*L, cb, pout, pin, EPI_PLAIN, 0.f, nullptr, st, K);
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
  const int K = mrhs_k_override(tune.k[nr]);
  if (nr == 1) return dslash_multi<1>(L, out, in, n, cb, reps, K);
  if (nr == 2) return dslash_multi<2>(L, out, in, n, cb, reps, K);
  if (nr == 4) return dslash_multi<4>(L, out, in, n, cb, reps, K);
  return dslash_multi<8>(L, out, in, n, cb, reps, K);
}

} // namespace qphix

using namespace qphix;

extern "C" {

void QPHIX_MI355X_set_mrhs_lds(int on) { g_mrhs_lds = on ? 1 : 0; }
void QPHIX_MI355X_set_msx_fuse(int steps) { g_msx_fuse = std::max(1, std::min(steps, MSX_MAXDEFER)); }

void QPHIX_F3_asqtad_dslash_multi(QPHIX_F3_FermionLinksAsqtad *L, QPHIX_F3_ColorVector *out[],
                                  QPHIX_F3_ColorVector *in[], int nvec, QPHIX_evenodd_t parity) {
  dslash_multi_api(L, out, in, nvec, parity, 0);
}

int QPHIX_F3_asqtad_mrhs_tune(QPHIX_F3_FermionLinksAsqtad *L, double *us, int *best_k) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called: asqtad_mrhs_tune", 0);
  QPHIX_REQUIRE(L && L->fat[0] && L->fat[1], "asqtad_mrhs_tune: both link parities are required", 0);
  const MrhsTune &t = autotune_mrhs(*L);
  if (us) {
    const int nrs[4] = {1, 2, 4, 8};
    for (int c = 0; c < 4; ++c)
      for (int q = 0; q < 3; ++q) us[3 * c + q] = t.us[nrs[c]][q];
  }
  if (best_k) *best_k = t.nr ? t.k[t.nr] : 0;
  return t.nr;
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
    for (int j = 0; j < nshift; ++j) {
      QPHIX_REQUIRE(res[(size_t)r * nshift + j], "asqtad_invert_multi_rhs: null res_arg", 0);
      const double t = res[(size_t)r * nshift + j]->resid > 0 ? res[(size_t)r * nshift + j]->resid : res[0]->resid;
      QPHIX_REQUIRE(t > 0, "asqtad_invert_multi_rhs: every solution needs a positive resid target", 0);
    }
  }
  QPHIX_REQUIRE(rhs_block == 0 || rhs_block == 1 || rhs_block == 2 || rhs_block == 4 || rhs_block == 8,
                "asqtad_invert_multi_rhs: rhs_block must be 0 (autotune), 1, 2, 4 or 8", 0);
  const LinksT<float> *Lf = sloppy_links(L);
  if (!Lf) {  // out of device memory, error recorded
    info->status = QPHIX_MEM_ERROR;
    return 0;
  }
  const MrhsTune &tune = autotune_mrhs(*Lf);
  if (tune.nr == 0) return 0;                 // workspace allocation failed (error recorded)
  int nb = rhs_block > 0 ? rhs_block : tune.nr;
  while (nb > 1 && nb / 2 >= nrhs) nb /= 2;   // no block wider than the sources need
  const int K = mrhs_k_override(tune.k[nb]);
  switch (nb) {
  case 1: return invert_multi_rhs<1>(info, L, inv, res, mass, nshift, nrhs, out, in, K);
  case 2: return invert_multi_rhs<2>(info, L, inv, res, mass, nshift, nrhs, out, in, K);
  case 4: return invert_multi_rhs<4>(info, L, inv, res, mass, nshift, nrhs, out, in, K);
  default: return invert_multi_rhs<8>(info, L, inv, res, mass, nshift, nrhs, out, in, K);
  }
}

} // extern "C"
