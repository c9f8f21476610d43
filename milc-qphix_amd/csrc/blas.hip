// blas.hip -- fused vector updates and reductions of the CG and multi-shift
// solvers.  One work-item per site (6 reals); FP64 accumulation; the grid
// reduction's last block runs the solver's scalar update on the device.
//
// Reference loops: calc_source_norm / calc_rsq2 / axpy (ks_d_congrad_fn.cpp:62-188),
// calc_rsq2 / update_dest / update_pm / copyKS (ks_multicg_offset.cpp:69-184).
#include "blas.h"

#include <cstdlib>

namespace qphix {

template <typename T>
__global__ void __launch_bounds__(BLAS_BLOCK) k_norm2(const T *v, int vh, ReduceCtl rc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s[1] = {0.0};
  if (i < vh) {
    cplx<typename Spin<T>::C> a[3];
    Spin<T>::load(v, i, a);
    for (int c = 0; c < 3; ++c) s[0] += (double)a[c].re * a[c].re + (double)a[c].im * a[c].im;
  }
  grid_reduce<1>(s, rc);
}

// x += a p ; r += a ttt ; reduce |r|^2 and the per-site relative residue.
template <typename T>
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_cg_update(T *x, const T *p, T *r, const T *ttt, int vh, ReduceCtl rc, const int *gate) {
  if (*gate == 0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s[2] = {0.0, 0.0};
  if (i < vh) {
    using C = typename Spin<T>::C;
    const C a = (C)rc.state->a;
    cplx<C> xv[3], pv[3], rv[3], tv[3];
    Spin<T>::load(x, i, xv);
    Spin<T>::load(p, i, pv);
    Spin<T>::load(r, i, rv);
    Spin<T>::load(ttt, i, tv);
    double num = 0;
    for (int c = 0; c < 3; ++c) {
      xv[c].re = xv[c].re + a * pv[c].re;
      xv[c].im = xv[c].im + a * pv[c].im;
      rv[c].re = rv[c].re + a * tv[c].re;
      rv[c].im = rv[c].im + a * tv[c].im;
      num += (double)rv[c].re * rv[c].re + (double)rv[c].im * rv[c].im;
    }
    Spin<T>::store(x, i, xv);
    Spin<T>::store(r, i, rv);
    s[0] = num;
    s[1] = rel_term(site_norm2_rounded(rv), site_norm2_rounded(xv));
  }
  grid_reduce<2>(s, rc);
}

// p = r + b p
template <typename T>
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_cg_axpy(const T *r, T *p, int vh, const SolverState *st, const int *gate) {
  if (*gate == 0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  using C = typename Spin<T>::C;
  const C b = (C)st->b;
  cplx<C> rv[3], pv[3];
  Spin<T>::load(r, i, rv);
  Spin<T>::load(p, i, pv);
  for (int c = 0; c < 3; ++c) {
    pv[c].re = rv[c].re + b * pv[c].re;
    pv[c].im = rv[c].im + b * pv[c].im;
  }
  Spin<T>::store(p, i, pv);
}

// CG step with consumer-side reductions (the arithmetic of SOP_CG_PKP and
// k_cg_update): every block derives a = -rsq / pkp from the p.Ap partials of
// the preceding Dslash; x += a p, r += a Ap; block 0 records pkp, a, oldrsq
// and the iteration (fields no other block of this kernel reads); |r|^2 and
// the relative-residue sum go to out_partials.
template <typename T>
__global__ void __launch_bounds__(CONS_BLOCK_MAX)
    k_cg_update_c(T *x, const T *p, T *r, const T *ttt, int vh, SolverState *st,
                  const double *pkp_partials, int nb_pkp, double *out_partials, const int *gate) {
  if (!gate_open(gate)) return;
  using C = typename Spin<T>::C;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  cplx<C> xv[3], pv[3], rv[3], tv[3];
  if (i < vh) {  // requested before the partials are summed
    Spin<T>::load(x, i, xv);
    Spin<T>::load(p, i, pv);
    Spin<T>::load(r, i, rv);
    Spin<T>::load(ttt, i, tv);
  }
  double pk[1];
  sum_partials<1>(pkp_partials, nb_pkp, pk);
  const double rsq = st->rsq;
  const double a = rnd(st, -rsq / pk[0]);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->oldrsq = rsq;
    st->pkp = pk[0];
    st->iteration++;
    st->a = a;
  }
  double s[2] = {0.0, 0.0};
  if (i < vh) {
    const C ac = (C)a;
    double num = 0;
    for (int c = 0; c < 3; ++c) {
      xv[c].re = xv[c].re + ac * pv[c].re;
      xv[c].im = xv[c].im + ac * pv[c].im;
      rv[c].re = rv[c].re + ac * tv[c].re;
      rv[c].im = rv[c].im + ac * tv[c].im;
      num += (double)rv[c].re * rv[c].re + (double)rv[c].im * rv[c].im;
    }
    Spin<T>::store(x, i, xv);
    Spin<T>::store(r, i, rv);
    s[0] = num;
    s[1] = rel_term(site_norm2_rounded(rv), site_norm2_rounded(xv));
  }
  block_sum<2>(s);
  if (threadIdx.x == 0) {
    out_partials[blockIdx.x] = s[0];
    out_partials[(size_t)gridDim.x + blockIdx.x] = s[1];
  }
}

// p = r + b p, with b and the stopping rule of SOP_CG_RSQ2 evaluated in every
// block from the update's partials; block 0 records them and closes the gate
// when the reference's inner loop would end (every block then skips the
// update, whichever value of the gate it read).
template <typename T>
__global__ void __launch_bounds__(CONS_BLOCK_MAX)
    k_cg_axpy_c(const T *r, T *p, int vh, SolverState *st, const double *partials, int nb, int *gate) {
  if (!gate_open(gate)) return;
  using C = typename Spin<T>::C;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  cplx<C> rv[3], pv[3];
  if (i < vh) {  // requested before the partials are summed
    Spin<T>::load(r, i, rv);
    Spin<T>::load(p, i, pv);
  }
  double v[2];
  sum_partials<2>(partials, nb, v);
  const double rsq = v[0];
  // relrsq keeps its previous value when the relative residue is off
  const double relrsq = st->relrsqmin > 0 ? rnd(st, sqrt(2.0 * v[1] / st->vol)) : st->relrsq;
  const double size_r = rnd(st, rsq) / st->src_norm;
  const double size_relr = rnd(st, relrsq);
  const double b = rnd(st, rnd(st, rsq) / st->oldrsq);
  const bool conv = (st->rsqmin <= 0 || st->rsqmin > size_r) &&
                    (st->relrsqmin <= 0 || st->relrsqmin > size_relr);
  const bool stop = st->iteration % st->niter == 0 || conv;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rsq = rsq;
    st->relrsq = relrsq;
    st->size_r = size_r;
    st->size_relr = size_relr;
    st->final_rsq = size_r;
    st->b = b;
    if (stop) *gate = 0;
  }
  if (stop) return;
  if (i >= vh) return;
  const C bc = (C)b;
  for (int c = 0; c < 3; ++c) {
    pv[c].re = rv[c].re + bc * pv[c].re;
    pv[c].im = rv[c].im + bc * pv[c].im;
  }
  Spin<T>::store(p, i, pv);
}

// Single-reduction CG step, the FEWSUMS variant of the reference's loop
// (ks_d_congrad_fn.cpp:450-541; MILC's algorithm, whose sums the C++ port
// leaves uncomputed, :150-158).  The preceding EPI_PKP_FS Dslash reduced
// pkp = p.ttt, c_tr = Re<ttt|r>, c_tt = <ttt|ttt> and the actual |r|^2 of the
// residual before this step, and snapshotted the gate, the iteration and rsq
// (the previous step's estimate, or the restart's true value) into fs_*.
// Then a = -rsq / pkp; x += a p; r += a ttt; the new |r|^2 follows without a
// second reduction, rsq' = actual + 2 a c_tr + a^2 c_tt; b = rsq' / actual;
// and unless the inner loop ends here (restart interval or convergence on
// rsq', ks_d_congrad_fn.cpp:566-572) p = r + b p, all in one pass.  Every
// block derives the same scalars from the same partials; block 0 records them.
// The relative residue needs a reduction of the updated vectors, so this step
// is used only when relresid is off.
template <typename T>
__global__ void __launch_bounds__(CONS_BLOCK_MAX)
    k_cg_step_fs(T *x, T *p, T *r, const T *ttt, int vh, SolverState *st, const double *partials, int nb,
                 int *gate) {
  if (st->fs_gate == 0) return;  // written by the preceding kernel: uniform
  using C = typename Spin<T>::C;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  cplx<C> xv[3], pv[3], rv[3], tv[3];
  if (i < vh) {  // requested before the partials are summed
    Spin<T>::load(x, i, xv);
    Spin<T>::load(p, i, pv);
    Spin<T>::load(r, i, rv);
    Spin<T>::load(ttt, i, tv);
  }
  double v[4];
  sum_partials<4>(partials, nb, v);
  const double pkp = v[0], c_tr = v[1], c_tt = v[2], actual = v[3];
  const double a = rnd(st, -st->fs_rsq / pkp);
  const double rsq = actual + 2.0 * a * c_tr + a * a * c_tt;
  const int it = st->fs_iter + 1;
  const double size_r = rnd(st, rsq) / st->src_norm;
  const double b = rnd(st, rnd(st, rsq) / actual);
  const bool conv = st->rsqmin <= 0 || st->rsqmin > size_r;
  const bool stop = it % st->niter == 0 || conv;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->iteration = it;
    st->oldrsq = actual;
    st->pkp = pkp;
    st->a = a;
    st->rsq = rsq;
    st->size_r = size_r;
    st->final_rsq = size_r;
    st->b = b;
    if (stop) *gate = 0;
  }
  if (i >= vh) return;
  const C ac = (C)a, bc = (C)b;
  for (int c = 0; c < 3; ++c) {
    xv[c].re = xv[c].re + ac * pv[c].re;
    xv[c].im = xv[c].im + ac * pv[c].im;
    rv[c].re = rv[c].re + ac * tv[c].re;
    rv[c].im = rv[c].im + ac * tv[c].im;
  }
  Spin<T>::store(x, i, xv);
  Spin<T>::store(r, i, rv);
  if (stop) return;
  for (int c = 0; c < 3; ++c) {
    pv[c].re = rv[c].re + bc * pv[c].re;
    pv[c].im = rv[c].im + bc * pv[c].im;
  }
  Spin<T>::store(p, i, pv);
}

// multi-shift: r += beta_low ttt ; reduce |r|^2
template <typename T>
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_ms_update_r(T *r, const T *ttt, int vh, ReduceCtl rc, const int *gate) {
  if (*gate == 0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s[1] = {0.0};
  if (i < vh) {
    using C = typename Spin<T>::C;
    const C bl = (C)rc.state->beta_i[rc.state->j_low];
    cplx<C> rv[3], tv[3];
    Spin<T>::load(r, i, rv);
    Spin<T>::load(ttt, i, tv);
    for (int c = 0; c < 3; ++c) {
      rv[c].re = rv[c].re + bl * tv[c].re;
      rv[c].im = rv[c].im + bl * tv[c].im;
      s[0] += (double)rv[c].re * rv[c].re + (double)rv[c].im * rv[c].im;
    }
    Spin<T>::store(r, i, rv);
  }
  grid_reduce<1>(s, rc);
}

// One step of shift j: x_j += beta_j p_j ; p_j = zeta_j r + alpha_j p_j, with
// zeta_j r formed in FP64 and rounded to the vector type (explicit fmas: every
// kernel below rounds identically).
template <typename C>
__device__ __forceinline__ void ms_one(cplx<C> *xv, cplx<C> *pv, const cplx<C> *rv, C bj, C aj, double zj) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    xv[c].re = fma(bj, pv[c].re, xv[c].re);
    xv[c].im = fma(bj, pv[c].im, xv[c].im);
    const C tr = (C)((double)rv[c].re * zj), ti = (C)((double)rv[c].im * zj);
    pv[c].re = fma(aj, pv[c].re, tr);
    pv[c].im = fma(aj, pv[c].im, ti);
  }
}

// multi-shift update_pm: x_j += beta_j p_j ; p_j = zeta_j r + alpha_j p_j,
// every shift j < num_now in one launch (blockIdx.y = j).
template <typename T>
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_ms_update_pm(T *const *x, T *const *p, const T *r, int vh, const SolverState *st,
                   const int *gate) {
  if (*gate == 0) return;
  const int j = blockIdx.y;
  if (j >= st->num_now) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  using C = typename Spin<T>::C;
  cplx<C> xv[3], pv[3], rv[3];
  Spin<T>::load(x[j], i, xv);
  Spin<T>::load(p[j], i, pv);
  Spin<T>::load(r, i, rv);
  ms_one<C>(xv, pv, rv, (C)st->beta_i[j], (C)st->alpha[j], st->zeta_ip1[j]);
  Spin<T>::store(x[j], i, xv);
  Spin<T>::store(p[j], i, pv);
}

// The same with the shifted systems updated up to st->fuse steps at a time.
// Only the base system's p (j_low) feeds the operator, so on a deferred step
// (st->defer, decided by this step's finalize) only j_low is updated and r is
// kept in rprev slot st->pend; the next step that is not deferred applies, for
// every other shift, the pending steps (their coefficients kept by
// k_ms_commit) and then its own, reading and writing x_j and p_j once.  The
// arithmetic and its order are those of k_ms_update_pm step by step.
// flush = 1: the pending steps only (after the iteration stopped).
//
// The grid is resident-sized (a multiple of 8): XCD xc = blockIdx % 8 (blocks
// are dispatched round-robin over the 8 XCDs) takes a contiguous run of site
// blocks and its blocks stride through it.  On a deferred step they update
// the base system only; otherwise each work-item loads r and the pending
// residuals once and updates every shift of its site.
// Measured at 32^3x64, 12 shifts, FP64, per step: 204 us with one block per
// slot and the shift in blockIdx.y (each shift re-read the pending
// residuals, 0.83 GB per step), 171 us with one block per slot in this order
// (0.34 GB; a deferred step still dispatched every shift's blocks, 107 us for
// 0.3 GB), 124 us with the resident grid (48 us per deferred step, 656 us per
// full step; the solve 1.04 -> 0.96 s).  One, three or four shifts at a time
// instead of two: 664 / 646 / 638 us against 657 us per full step (168-239
// VGPRs, no spills), under 0.3% of the solve.
template <typename T>
__device__ __forceinline__ void ms_def_one(T *const *x, T *const *p, const T *r, T *rprev, size_t nv,
                                           const SolverState *st, int j, int i, int flush) {
  const int jl = st->j_low, np = st->pend, nn = st->num_now;
  using C = typename Spin<T>::C;
  cplx<C> xv[3], pv[3], rv[3];
  if (j == jl) {  // the base system: every step, never pending
    if (flush) return;
    Spin<T>::load(r, i, rv);
    if (st->defer) Spin<T>::store(rprev + (size_t)np * nv, i, rv);
    Spin<T>::load(x[j], i, xv);
    Spin<T>::load(p[j], i, pv);
    ms_one<C>(xv, pv, rv, (C)st->beta_i[j], (C)st->alpha[j], st->zeta_ip1[j]);
    Spin<T>::store(x[j], i, xv);
    Spin<T>::store(p[j], i, pv);
    return;
  }
  // num_now only shrinks, so the first pending step covers the later ones
  if (j >= (np ? st->pend_nn[0] : nn)) return;
  Spin<T>::load(x[j], i, xv);
  Spin<T>::load(p[j], i, pv);
  for (int q = 0; q < np; ++q) {
    if (j >= st->pend_nn[q]) break;
    Spin<T>::load(rprev + (size_t)q * nv, i, rv);
    ms_one<C>(xv, pv, rv, (C)st->pbeta[q][j], (C)st->palpha[q][j], st->pzeta[q][j]);
  }
  if (!flush && j < nn) {
    Spin<T>::load(r, i, rv);
    ms_one<C>(xv, pv, rv, (C)st->beta_i[j], (C)st->alpha[j], st->zeta_ip1[j]);
  }
  Spin<T>::store(x[j], i, xv);
  Spin<T>::store(p[j], i, pv);
}

template <typename T>
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_ms_update_pm_def(T *const *x, T *const *p, const T *r, T *rprev, int vh, int nshift,
                       const SolverState *st, const int *gate, int flush) {
  if (!flush && *gate == 0) return;
  if (flush && st->pend == 0) return;  // nothing pending
  const size_t nv = 6 * (size_t)vh;
  const int nsb = (vh + BLAS_BLOCK - 1) / BLAS_BLOCK;
  const int xc = blockIdx.x % 8, l = blockIdx.x / 8, per = gridDim.x / 8;
  if (!flush && st->defer) {  // this step's shifted updates wait: the base system only
    const int lo = (int)((long long)xc * nsb / 8), hi = (int)((long long)(xc + 1) * nsb / 8);
    for (int sb = lo + l; sb < hi; sb += per) {
      const int i = sb * BLAS_BLOCK + threadIdx.x;
      if (i < vh) ms_def_one<T>(x, p, r, rprev, nv, st, st->j_low, i, flush);
    }
    return;
  }
  // every shift: a work-item keeps r and the pending residuals in registers
  // and takes its site's shifts two at a time (both loaded before either is
  // stored); per (shift, site) the steps and their order are as above
  using C = typename Spin<T>::C;
  const int jl = st->j_low, np = st->pend, nn = st->num_now;
  const int jmax = min(np ? st->pend_nn[0] : nn, nshift);
  const int lo = (int)((long long)xc * nsb / 8), hi = (int)((long long)(xc + 1) * nsb / 8);
  for (int sb = lo + l; sb < hi; sb += per) {
    const int i = sb * BLAS_BLOCK + threadIdx.x;
    if (i >= vh) continue;
    cplx<C> rv[3], rq[MSX_MAXDEFER - 1][3];
    if (!flush) Spin<T>::load(r, i, rv);
#pragma unroll
    for (int q = 0; q < MSX_MAXDEFER - 1; ++q)
      if (q < np) Spin<T>::load(rprev + (size_t)q * nv, i, rq[q]);
    for (int j0 = 0; j0 < jmax; j0 += 2) {
      cplx<C> xv[2][3], pv[2][3];
      bool on[2];
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        const int j = j0 + g;
        on[g] = j < jmax && j != jl;
        if (on[g]) {
          Spin<T>::load(x[j], i, xv[g]);
          Spin<T>::load(p[j], i, pv[g]);
        }
      }
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        if (!on[g]) continue;
        const int j = j0 + g;
#pragma unroll
        for (int q = 0; q < MSX_MAXDEFER - 1; ++q)
          if (q < np && j < st->pend_nn[q])
            ms_one<C>(xv[g], pv[g], rq[q], (C)st->pbeta[q][j], (C)st->palpha[q][j], st->pzeta[q][j]);
        if (!flush && j < nn) ms_one<C>(xv[g], pv[g], rv, (C)st->beta_i[j], (C)st->alpha[j], st->zeta_ip1[j]);
        Spin<T>::store(x[j], i, xv[g]);
        Spin<T>::store(p[j], i, pv[g]);
      }
    }
    if (!flush) {  // the base system (this step is not deferred: nothing to keep)
      cplx<C> xv[3], pv[3];
      Spin<T>::load(x[jl], i, xv);
      Spin<T>::load(p[jl], i, pv);
      ms_one<C>(xv, pv, rv, (C)st->beta_i[jl], (C)st->alpha[jl], st->zeta_ip1[jl]);
      Spin<T>::store(x[jl], i, xv);
      Spin<T>::store(p[jl], i, pv);
    }
  }
}

// End of a step of the deferring multi-shift update: keep this step's
// coefficients when its shifted updates wait, else nothing is pending.
// flush = 1: after k_ms_update_pm_def(flush) -- nothing pending.
__global__ void k_ms_commit(SolverState *s, const int *gate, int flush) {
  if (!flush && *gate == 0) return;
  if (!flush && s->defer) {
    const int q = s->pend++;
    s->pend_nn[q] = s->num_now;
    for (int j = 0; j < s->num_now; ++j) {
      s->pbeta[q][j] = s->beta_i[j];
      s->palpha[q][j] = s->alpha[j];
      s->pzeta[q][j] = s->zeta_ip1[j];
    }
  } else {
    s->pend = 0;
  }
  s->defer = 0;
}

// multi-shift update_dest: x_j += beta_j p_j, j < num_now
template <typename T>
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_ms_update_dest(T *const *x, T *const *p, int vh, const SolverState *st) {
  const int j = blockIdx.y;
  if (j >= st->num_now) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  using C = typename Spin<T>::C;
  const C bj = (C)st->beta_i[j];
  cplx<C> xv[3], pv[3];
  Spin<T>::load(x[j], i, xv);
  Spin<T>::load(p[j], i, pv);
  for (int c = 0; c < 3; ++c) {
    xv[c].re = xv[c].re + bj * pv[c].re;
    xv[c].im = xv[c].im + bj * pv[c].im;
  }
  Spin<T>::store(x[j], i, xv);
}

__global__ void k_set_int(int *p, int v) { *p = v; }

// --------------------------------------------------------------- launchers
static inline int nb(int vh) { return (vh + BLAS_BLOCK - 1) / BLAS_BLOCK; }

ReduceCtl reduce_ctl(SolverState *st, int op) {
  ReduceCtl rc;
  rc.partials = ctx().partials;
  rc.state = st;
  rc.op = op;
  rc.out = st ? nullptr : reduce_result_ptr();
  return rc;
}

double *reduce_result_ptr() { return ctx().partials + 4 * (size_t)(ctx().max_blocks - 1); }

void reduce_finish(int nv, int nblocks, const ReduceCtl &rc, const int *gate) {
  if (nv == 1)
    hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(FINALIZE_BLOCK), 0, ctx().stream, rc.partials,
                       nblocks, rc, gate);
  else
    hipLaunchKernelGGL(k_finalize<2>, dim3(1), dim3(FINALIZE_BLOCK), 0, ctx().stream, rc.partials,
                       nblocks, rc, gate);
  HIP_CHECK(hipGetLastError());
}

template <typename T> void blas_norm2(const T *v, SolverState *st, int op) {
  const int vh = ctx().g.vh;
  const ReduceCtl rc = reduce_ctl(st, op);
  hipLaunchKernelGGL(k_norm2<T>, dim3(nb(vh)), dim3(BLAS_BLOCK), 0, ctx().stream, v, vh, rc);
  HIP_CHECK(hipGetLastError());
  reduce_finish(1, nb(vh), rc, nullptr);
}

template <typename T>
void blas_cg_update(T *x, const T *p, T *r, const T *ttt, SolverState *st, const int *gate) {
  const int vh = ctx().g.vh;
  const ReduceCtl rc = reduce_ctl(st, SOP_CG_RSQ2);
  hipLaunchKernelGGL(k_cg_update<T>, dim3(nb(vh)), dim3(BLAS_BLOCK), 0, ctx().stream, x, p, r,
                     ttt, vh, rc, gate);
  HIP_CHECK(hipGetLastError());
  reduce_finish(2, nb(vh), rc, gate);
}

template <typename T> void blas_cg_axpy(const T *r, T *p, const SolverState *st, const int *gate) {
  const int vh = ctx().g.vh;
  hipLaunchKernelGGL(k_cg_axpy<T>, dim3(nb(vh)), dim3(BLAS_BLOCK), 0, ctx().stream, r, p, vh, st,
                     gate);
  HIP_CHECK(hipGetLastError());
}

int cons_block() {
  static const int b = [] {
    const char *e = std::getenv("QPHIX_CONS_BLOCK");
    const int v = e ? std::atoi(e) : 512;
    return (v == 256 || v == 1024) ? v : 512;
  }();
  return b;
}

// Second partials region: consumer-side reductions read one region while
// the consuming kernel writes its own partials into the other.
double *partials_b() { return ctx().partials + 2 * (size_t)ctx().max_blocks; }

template <typename T>
void blas_cg_step_c(T *x, T *p, T *r, const T *ttt, SolverState *st, int *gate, int nb_pkp) {
  const int vh = ctx().g.vh;
  hipLaunchKernelGGL(k_cg_update_c<T>, dim3(cons_nb(vh)), dim3(cons_block()), 0, ctx().stream, x, p, r, ttt, vh,
                     st, ctx().partials, nb_pkp, partials_b(), gate);
  hipLaunchKernelGGL(k_cg_axpy_c<T>, dim3(cons_nb(vh)), dim3(cons_block()), 0, ctx().stream, r, p, vh, st,
                     partials_b(), cons_nb(vh), gate);
  HIP_CHECK(hipGetLastError());
}

template <typename T>
void blas_cg_step_fs(T *x, T *p, T *r, const T *ttt, SolverState *st, int *gate, int nb_pkp) {
  const int vh = ctx().g.vh;
  hipLaunchKernelGGL(k_cg_step_fs<T>, dim3(cons_nb(vh)), dim3(cons_block()), 0, ctx().stream, x, p, r, ttt, vh,
                     st, ctx().partials, nb_pkp, gate);
  HIP_CHECK(hipGetLastError());
}

template <typename T> void blas_ms_update_r(T *r, const T *ttt, SolverState *st, const int *gate) {
  const int vh = ctx().g.vh;
  const ReduceCtl rc = reduce_ctl(st, SOP_MS_RSQ2);
  hipLaunchKernelGGL(k_ms_update_r<T>, dim3(nb(vh)), dim3(BLAS_BLOCK), 0, ctx().stream, r, ttt, vh,
                     rc, gate);
  HIP_CHECK(hipGetLastError());
  reduce_finish(1, nb(vh), rc, gate);
}

template <typename T>
void blas_ms_update_pm(T *const *x, T *const *p, const T *r, int nshift, const SolverState *st,
                       const int *gate) {
  const int vh = ctx().g.vh;
  hipLaunchKernelGGL(k_ms_update_pm<T>, dim3(nb(vh), nshift), dim3(BLAS_BLOCK), 0, ctx().stream,
                     x, p, r, vh, st, gate);
  HIP_CHECK(hipGetLastError());
}

template <typename T>
void blas_ms_update_pm_def(T *const *x, T *const *p, const T *r, T *rprev, int nshift, SolverState *st,
                           const int *gate, bool flush) {
  const int vh = ctx().g.vh;
  // resident-sized grid, a multiple of 8 (one run of site blocks per XCD)
  const int grid = 8 * std::max(ctx().n_cu, 1);
  hipLaunchKernelGGL(k_ms_update_pm_def<T>, dim3(grid), dim3(BLAS_BLOCK), 0, ctx().stream, x, p, r, rprev, vh,
                     nshift, st, gate, flush ? 1 : 0);
  hipLaunchKernelGGL(k_ms_commit, dim3(1), dim3(1), 0, ctx().stream, st, gate, flush ? 1 : 0);
  HIP_CHECK(hipGetLastError());
}

template <typename T>
void blas_ms_update_dest(T *const *x, T *const *p, int nshift, const SolverState *st) {
  const int vh = ctx().g.vh;
  hipLaunchKernelGGL(k_ms_update_dest<T>, dim3(nb(vh), nshift), dim3(BLAS_BLOCK), 0, ctx().stream,
                     x, p, vh, st);
  HIP_CHECK(hipGetLastError());
}

void set_int(int *p, int v) {
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, ctx().stream, p, v);
  HIP_CHECK(hipGetLastError());
}

#define INST(T)                                                                                \
  template void blas_norm2<T>(const T *, SolverState *, int);                                  \
  template void blas_cg_update<T>(T *, const T *, T *, const T *, SolverState *, const int *); \
  template void blas_cg_step_c<T>(T *, T *, T *, const T *, SolverState *, int *, int);         \
  template void blas_cg_step_fs<T>(T *, T *, T *, const T *, SolverState *, int *, int);        \
  template void blas_cg_axpy<T>(const T *, T *, const SolverState *, const int *);             \
  template void blas_ms_update_r<T>(T *, const T *, SolverState *, const int *);               \
  template void blas_ms_update_pm<T>(T *const *, T *const *, const T *, int,                   \
                                     const SolverState *, const int *);                        \
  template void blas_ms_update_dest<T>(T *const *, T *const *, int, const SolverState *);      \
  template void blas_ms_update_pm_def<T>(T *const *, T *const *, const T *, T *, int, SolverState *, \
                                         const int *, bool);
INST(double)
INST(float)
#undef INST

} // namespace qphix
