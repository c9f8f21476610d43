// reduce.h -- deterministic grid reductions and the on-device solver scalars.
//
// Every reducing kernel writes one partial per block; a one-block kernel
// launched right after it sums the partials in a fixed order and runs the
// solver's scalar update in place.  The reference does the same arithmetic on
// the host after an OpenMP reduction + g_doublesum (ks_d_congrad_fn.cpp:486-566,
// ks_multicg_offset.cpp:463-598); keeping it on the device lets the host queue
// many iterations without reading anything back.
#pragma once
#include "device.h"

namespace qphix {

constexpr int MAXNDMASS = 150;  // ks_d_congrad_fn.h:4
constexpr int MSX_MAXDEFER = 8;  // batched multi-shift: most steps per shifted update

struct SolverState {
  double rsq, oldrsq, pkp, relrsq, src_norm;
  double size_r, size_relr, final_rsq, final_rel;
  double a, b, msq_x4, rsqmin, relrsqmin, rsqstop, vol;
  int iteration, niter, max_cg, active;
  int single_coef;  // 1: scalars rounded to float as the F build's fptype does
  int converged, nshift, j_low, num_now, pad_;
  double offset_low;
  // Single-reduction (FEWSUMS) steps: the p.Ap Dslash's block 0 copies the
  // gate, the iteration and the |r|^2 estimates here before the fused update
  // kernel runs, so that kernel reads values its own block 0 may overwrite.
  int fs_gate, fs_iter;
  double fs_rsq, fs_mxrsq;
  double shifts[MAXNDMASS], zeta_i[MAXNDMASS], zeta_im1[MAXNDMASS], zeta_ip1[MAXNDMASS];
  double beta_i[MAXNDMASS], beta_im1[MAXNDMASS], alpha[MAXNDMASS];
  int finished[MAXNDMASS];
  // mixed-precision solver scalars
  double mx_alpha, mx_beta, mx_gamma, mx_rsq, mx_rsq_old, mx_maxr, mx_delta, mx_rp;
  int ru;      // batched multi-shift: reliable update / convergence check requested
  int pad2_;
  // batched multi-shift: 0 shift still iterating, 2 converged (this step's
  // update is its last), 1 retired; sdone_at = the step it converged at
  int sdone[MAXNDMASS];
  int sdone_at[MAXNDMASS];
  // batched multi-shift, shifted systems updated up to `fuse` steps at a time:
  // `defer` (set by this step's finalize) = the shifted updates of this step
  // are left to a later one; `pend` (kept by the commit) = how many steps are
  // pending, step q with its num_now and coefficients in pend_nn[q], pbeta[q],
  // palpha[q], pzeta[q]; `fuse` = 1 updates every step.  A step that requests
  // a pause, or retires a shift, is never deferred, so a paused or finished
  // source has nothing pending.  `phase` = the step's place in its batch:
  // every source closes its group at phase fuse - 1 (mod fuse), so the
  // sources of a wavefront (interleaved per site) do their full updates on
  // the same steps; a wavefront with one source updating every shift costs
  // as much as one where all do.
  int defer, pend, fuse, phase;
  int pend_nn[MSX_MAXDEFER - 1];
  double pbeta[MSX_MAXDEFER - 1][MAXNDMASS], palpha[MSX_MAXDEFER - 1][MAXNDMASS];
  double pzeta[MSX_MAXDEFER - 1][MAXNDMASS];
};

enum ScalarOp {
  SOP_NONE = 0,
  SOP_SRC_NORM,   // src_norm = v0
  SOP_CG_PKP,     // ks_d_congrad_fn.cpp:486-511
  SOP_CG_RSQ2,    // ks_d_congrad_fn.cpp:524-572
  SOP_CG_RESTART, // ks_d_congrad_fn.cpp:372-419
  SOP_MS_PKP,     // ks_multicg_offset.cpp:463-496
  SOP_MS_RSQ2,    // ks_multicg_offset.cpp:504-598
  SOP_STORE_RSQ,  // rsq = v0
  SOP_MX_PAP,     // mixed CG: alpha from p.(-Mp)
  SOP_MX_RSQ,     // mixed CG: sloppy |r|^2, reliable-update test, beta
  SOP_MX_TRUE,    // mixed CG: true |b - Mx|^2 after a reliable update
  SOP_MX_RP,      // mixed CG: r.p for re-orthogonalising p after an update
  SOP_MSX_RSQ2    // batched mixed multi-shift: |r|^2, update request, alpha
};

// (fptype) cast of the reference: round to float when the solver is the F build.
__device__ __forceinline__ double rnd(const SolverState *s, double v) {
  return s->single_coef ? (double)(float)v : v;
}

template <typename R> __device__ void ms_pkp_recurrence(SolverState *s) {
  const int jl = s->j_low;
  s->beta_i[jl] = (double)(R)(-s->rsq / s->pkp);
  s->zeta_ip1[jl] = 1.0;
  for (int j = 0; j < s->num_now; ++j) {
    if (j == jl) continue;
    R zi = (R)s->zeta_i[j], zim1 = (R)s->zeta_im1[j], bim1 = (R)s->beta_im1[jl];
    R bl = (R)s->beta_i[jl], al = (R)s->alpha[jl], sh = (R)s->shifts[j];
    R zip1 = zi * zim1 * bim1;
    const R c1 = bl * al * (zim1 - zi);
    const R c2 = (R)(zim1 * bim1 * (1.0 + sh * bl));
    R bj;
    if (c1 + c2 != (R)0.0) zip1 /= c1 + c2;
    else { zip1 = 0; s->finished[j] = 1; }
    if (zi != (R)0.0) bj = bl * zip1 / zi;
    else { zip1 = 0; bj = 0; s->finished[j] = 1; }
    s->zeta_ip1[j] = zip1;
    s->beta_i[j] = bj;
    if (s->finished[j] == 1 && j == s->num_now - 1 && j > jl) s->num_now--;
  }
}

template <typename R> __device__ void ms_alpha_recurrence(SolverState *s) {
  const int jl = s->j_low;
  s->alpha[jl] = (double)(R)(s->rsq / s->oldrsq);
  for (int j = 0; j < s->num_now; ++j) {
    if (j == jl) continue;
    const R zi = (R)s->zeta_i[j], bl = (R)s->beta_i[jl];
    if (zi * bl != (R)0.0)
      s->alpha[j] = (R)s->alpha[jl] * (R)s->zeta_ip1[j] * (R)s->beta_i[j] / (zi * bl);
    else { s->alpha[j] = 0; s->finished[j] = 1; }
  }
  for (int j = 0; j < s->num_now; ++j) {
    s->beta_im1[j] = s->beta_i[j];
    s->zeta_im1[j] = s->zeta_i[j];
    s->zeta_i[j] = s->zeta_ip1[j];
  }
}

// Runs in one thread of the last block.  v[0], v[1] are the reduced sums.
__device__ inline void apply_scalar_op(SolverState *s, int op, const double *v) {
  switch (op) {
  case SOP_SRC_NORM:
    s->src_norm = v[0];
    break;
  case SOP_STORE_RSQ:
    s->rsq = v[0];
    break;
  case SOP_CG_PKP:
    s->oldrsq = s->rsq;
    s->pkp = v[0];
    s->iteration++;
    s->a = rnd(s, -s->rsq / s->pkp);
    break;
  case SOP_CG_RSQ2: {
    s->rsq = v[0];
    if (s->relrsqmin > 0) s->relrsq = rnd(s, sqrt(2.0 * v[1] / s->vol));
    s->size_r = rnd(s, s->rsq) / s->src_norm;
    s->size_relr = rnd(s, s->relrsq);
    s->final_rsq = s->size_r;
    s->b = rnd(s, rnd(s, s->rsq) / s->oldrsq);
    const bool conv = (s->rsqmin <= 0 || s->rsqmin > s->size_r) &&
                      (s->relrsqmin <= 0 || s->relrsqmin > s->size_relr);
    if (s->iteration % s->niter == 0 || conv) s->active = 0;
    break;
  }
  case SOP_CG_RESTART:
    s->rsq = v[0];
    if (s->relrsqmin > 0) s->relrsq = rnd(s, sqrt(2.0 * v[1] / s->vol));
    s->final_rsq = rnd(s, s->rsq) / s->src_norm;
    s->final_rel = rnd(s, s->relrsq);
    s->iteration++;
    break;
  case SOP_MS_PKP:
    s->oldrsq = s->rsq;
    s->pkp = v[0];
    s->iteration++;
    if (s->single_coef) ms_pkp_recurrence<float>(s);
    else ms_pkp_recurrence<double>(s);
    break;
  case SOP_MS_RSQ2:
    s->rsq = v[0];
    if (s->rsq <= s->rsqstop) {
      s->converged = 1;
      s->active = 0;
      break;
    }
    // The host never queues more than max_cg steps, so no cap is needed here.
    if (s->single_coef) ms_alpha_recurrence<float>(s);
    else ms_alpha_recurrence<double>(s);
    // shifted updates of this step wait for a later one (k_ms_update_pm_def)
    s->defer = (s->fuse > 1 && s->pend + 1 < s->fuse) ? 1 : 0;
    break;
  // Mixed precision: CG in the sloppy precision with reliable updates of the
  // FP64 solution and residual (Sleijpen & van der Vorst, Computing 56 (1996)
  // 141; the lattice form of Clark et al., Comput. Phys. Commun. 181 (2010)
  // 1517, sec. 4).  The sloppy step: pkp = p.(-Mp), alpha = |r|^2/(p.Mp).
  case SOP_MX_PAP:
    s->pkp = v[0];
    s->iteration++;
    s->mx_alpha = -s->mx_rsq / v[0];
    break;
  case SOP_MX_RSQ: {
    // Flag an update once |r| dropped by delta below its largest value since
    // the last update, or the sloppy residual meets the target, or the step
    // budget is spent; the host then recomputes r = b - Mx in FP64.
    const double rnew = v[0];
    s->mx_rsq_old = s->mx_rsq;
    s->mx_beta = rnew / s->mx_rsq;
    s->mx_rsq = rnew;
    const double r = sqrt(rnew);
    if (r > s->mx_maxr) s->mx_maxr = r;
    if (r < s->mx_delta * s->mx_maxr || rnew < s->rsqstop || s->iteration >= s->max_cg) s->active = 0;
    break;
  }
  case SOP_MX_TRUE:
    s->rsq = v[0];
    s->final_rsq = v[0] / s->src_norm;
    if (s->relrsqmin > 0) s->final_rel = sqrt(2.0 * v[1] / s->vol);
    s->iteration++;
    break;
  case SOP_MSX_RSQ2: {
    // As SOP_MS_RSQ2, but the step always completes (update_pm runs) and a
    // pause for an FP64 true-residual check is requested instead of stopping:
    // on apparent convergence, when |r| fell by delta below its maximum since
    // the last check, or when the step budget is spent.
    s->rsq = v[0];
    const double r = sqrt(v[0]);
    if (r > s->mx_maxr) s->mx_maxr = r;
    if (s->rsq <= s->rsqstop || r < s->mx_delta * s->mx_maxr || s->iteration >= s->max_cg) s->ru = 1;
    // A shifted system's residual is zeta_{i+1,j} r: once that meets the
    // target the shift takes this step's update and then retires.
    bool retire = false;
    for (int j = 0; j < s->num_now; ++j)
      if (j != s->j_low && s->sdone[j] == 0 && s->zeta_ip1[j] * s->zeta_ip1[j] * s->rsq < s->rsqstop) {
        s->sdone[j] = 2;
        s->sdone_at[j] = s->iteration;
        retire = true;
      }
    ms_alpha_recurrence<double>(s);
    s->defer = (s->fuse > 1 && s->pend + 1 < s->fuse && s->phase % s->fuse != s->fuse - 1 && !s->ru &&
                !retire) ? 1 : 0;
    break;
  }
  case SOP_MX_RP:
    // p <- r + beta (p - gamma r): gamma = r.p/|r|^2 restores p's
    // orthogonality to the new residual, beta = |r_true|^2 / |r_sloppy,old|^2.
    s->mx_rp = v[0];
    s->mx_gamma = v[0] / s->rsq;
    s->mx_beta = s->rsq / s->mx_rsq_old;
    s->mx_rsq = s->rsq;
    s->mx_maxr = sqrt(s->rsq);
    s->active = 1;
    break;
  default:
    break;
  }
}

// Block-wide sum of NV values; result valid in thread 0.  blockDim.x <= 1024.
template <int NV> __device__ __forceinline__ void block_sum(double *v) {
#pragma unroll
  for (int k = 0; k < NV; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  }
  __shared__ double sh[NV][16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k][w] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0;
      for (int j = 0; j < nw; ++j) s += sh[k][j];
      v[k] = s;
    }
  }
  __syncthreads();
}

// Two-stage reduction.  Stage 1 (inside the reducing kernel): each block
// writes its partial sums with plain stores.  Stage 2 (k_finalize, one block,
// launched next on the same stream): sums the partials in a fixed order and
// runs the scalar update.  A single-pass "last block finishes" scheme needs a
// device-scope release per block, which on the 8-XCD MI355X writes back L2
// and serialises an atomic per block -- measured at 60-80 us per reducing
// kernel at 32^4, against ~2 us for the extra launch.
struct ReduceCtl {
  double *partials;    // [NV][gridDim.x]
  SolverState *state;  // scalar update target, or null
  int op;
  double *out;         // when state is null: the NV sums are stored here
};

// Every thread of every block calls this once with its contribution.
template <int NV> __device__ __forceinline__ void grid_reduce(double *v, const ReduceCtl &rc) {
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
    const int nb = gridDim.x;
#pragma unroll
    for (int k = 0; k < NV; ++k) rc.partials[(size_t)k * nb + blockIdx.x] = v[k];
  }
}

constexpr int FINALIZE_BLOCK = 1024;

// Consumer-side finalize.  A kernel that needs the previous kernel's
// reduction sums its per-block partials ([NV][nb]) itself, in one fixed order,
// so every block obtains the same value bit for bit and no separate finalize
// launch (and its dependency gap) is needed.  Result valid in every thread.
// The loads of ROWS rows (one row = one partial per thread and sum) are all
// issued before the first is added, so a block pays the L2 latency once per
// ROWS rows; each thread still adds its partials in index order.  8 loads in
// flight per thread whatever NV (more spilled in the 4-sum update kernels).
template <int NV> __device__ __forceinline__ void sum_partials(const double *partials, int nb, double *v) {
  constexpr int SUM_ROWS = NV >= 8 ? 1 : 8 / NV;
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = 0.0;
  const int bd = blockDim.x;
  int b = threadIdx.x;
  for (; b + (SUM_ROWS - 1) * bd < nb; b += SUM_ROWS * bd) {
    double t[NV][SUM_ROWS];
#pragma unroll
    for (int u = 0; u < SUM_ROWS; ++u)
#pragma unroll
      for (int k = 0; k < NV; ++k) t[k][u] = partials[(size_t)k * nb + b + u * bd];
#pragma unroll
    for (int u = 0; u < SUM_ROWS; ++u)
#pragma unroll
      for (int k = 0; k < NV; ++k) s[k] += t[k][u];
  }
  for (; b < nb; b += bd) {
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] += partials[(size_t)k * nb + b];
  }
  block_sum<NV>(s);
  __shared__ double bc[NV];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) bc[k] = s[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = bc[k];
}

// Block-uniform read of a gate that block 0 of the same kernel may clear:
// thread 0 reads it once and the block follows that value.
__device__ __forceinline__ bool gate_open(const int *gate) {
  __shared__ int g;
  if (threadIdx.x == 0) g = *gate;
  __syncthreads();
  return g != 0;
}

template <int NV>
__global__ void __launch_bounds__(FINALIZE_BLOCK)
    k_finalize(const double *partials, int nb, ReduceCtl rc, const int *gate) {
  if (gate && *gate == 0) return;
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    s[k] = 0;
    for (int b = threadIdx.x; b < nb; b += FINALIZE_BLOCK) s[k] += partials[(size_t)k * nb + b];
  }
  block_sum<NV>(s);
  if (threadIdx.x == 0) {
    if (rc.state) apply_scalar_op(rc.state, rc.op, s);
    else if (rc.out) {
#pragma unroll
      for (int k = 0; k < NV; ++k) rc.out[k] = s[k];
    }
  }
}

// Enqueues stage 2 for a reducing kernel that ran with `nb` blocks.
void reduce_finish(int nv, int nb, const ReduceCtl &rc, const int *gate);

} // namespace qphix
