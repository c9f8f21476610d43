// mixed.hip -- mixed-precision CG with reliable updates.
//
// No counterpart in the reference (its solvers run in one precision per
// translation unit, qcd_data_types.h:4-20).  Solves the same system as
// QPHIX_D3_asqtad_invert, (4 m^2 - D D) x = b to a squared relative true
// residual `resid`, with every CG step in FP32 (or with FP16-stored spinors and
// FP32 arithmetic) on FP32 links, and FP64 kept for the solution, the true
// residual recomputed with the FP64 Dslash at each reliable update, and all
// reductions.  Each reliable update and each sloppy step counts as one
// iteration (the reference's counting, C9).
#include "blas.h"
#include "dslash_kernel.h"
#include "handles.h"
#include "qphix_mi355x.h"

#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>

namespace qphix {

// The sloppy step with consumer-side reductions (the arithmetic of SOP_MX_PAP
// and k_mx_step): every block derives alpha = -|r|^2 / p.(-Mp) from the
// preceding Dslash's partials; block 0 records pkp, alpha, the iteration and
// the old |r|^2 (fields no other block of this kernel reads).
template <typename P>
__global__ void __launch_bounds__(CONS_BLOCK_MAX)
    k_mx_step_c(float *xs, const P *p, float *rs, const float *Ap, int vh, SolverState *st,
                const double *pkp_partials, int nb_pkp, double *out_partials, const int *gate) {
  if (!gate_open(gate)) return;
  // the site's vectors are requested before the partials are summed, so both
  // latencies overlap
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  cplx<float> xv[3], pv[3], rv[3], av[3];
  if (i < vh) {
    Spin<float>::load(xs, i, xv);
    Spin<P>::load(p, i, pv);
    Spin<float>::load(rs, i, rv);
    Spin<float>::load(Ap, i, av);
  }
  double pk[1];
  sum_partials<1>(pkp_partials, nb_pkp, pk);
  const double rsq = st->mx_rsq;
  const double alpha = -rsq / pk[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->pkp = pk[0];
    st->iteration++;
    st->mx_alpha = alpha;
    st->mx_rsq_old = rsq;
  }
  double s[1] = {0.0};
  if (i < vh) {
    const float al = (float)alpha;
    for (int c = 0; c < 3; ++c) {
      xv[c].re += al * pv[c].re;
      xv[c].im += al * pv[c].im;
      rv[c].re += al * av[c].re;
      rv[c].im += al * av[c].im;
      s[0] += (double)rv[c].re * rv[c].re + (double)rv[c].im * rv[c].im;
    }
    Spin<float>::store(xs, i, xv);
    Spin<float>::store(rs, i, rv);
  }
  block_sum<1>(s);
  if (threadIdx.x == 0) out_partials[blockIdx.x] = s[0];
}

// p = r_s + beta p, with beta and the reliable-update test of SOP_MX_RSQ
// evaluated in every block from the step's partials; block 0 records them and
// closes the gate when an update is due (every block then skips, whichever
// value of the gate it read; the running maximum is updated idempotently).
template <typename P>
__global__ void __launch_bounds__(CONS_BLOCK_MAX)
    k_mx_axpy_c(const float *rs, P *p, int vh, SolverState *st, const double *partials, int nb, int *gate) {
  if (!gate_open(gate)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  cplx<float> rv[3], pv[3];
  if (i < vh) {  // requested before the partials are summed
    Spin<float>::load(rs, i, rv);
    Spin<P>::load(p, i, pv);
  }
  double v[1];
  sum_partials<1>(partials, nb, v);
  const double rnew = v[0];
  const double beta = rnew / st->mx_rsq_old;
  const double r = sqrt(rnew);
  const double maxr = fmax(st->mx_maxr, r);
  const bool update = r < st->mx_delta * maxr || rnew < st->rsqstop || st->iteration >= st->max_cg;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->mx_beta = beta;
    st->mx_rsq = rnew;
    st->mx_maxr = maxr;
    if (update) *gate = 0;
  }
  if (update) return;
  if (i >= vh) return;
  const float be = (float)beta;
  for (int c = 0; c < 3; ++c) {
    pv[c].re = rv[c].re + be * pv[c].re;
    pv[c].im = rv[c].im + be * pv[c].im;
  }
  Spin<P>::store(p, i, pv);
}

// The sloppy step as a single-reduction (FEWSUMS) step: the preceding
// EPI_PKP_FS Dslash reduced pkp = p.Ap (Ap = -Mp), c_tr = r.Ap, c_tt = Ap.Ap
// and rr = |r|^2, and snapshotted the gate, the iteration and |r|^2 (mx_rsq:
// the previous step's estimate, or the true value after a reliable update).
// alpha = -|r|^2 / pkp; x_s += alpha p; r_s += alpha Ap; the new |r|^2 follows
// as rr + 2 alpha c_tr + alpha^2 c_tt (ks_d_congrad_fn.cpp:531); beta = that /
// rr; the reliable-update test of k_mx_axpy_c; p = r_s + beta p unless an
// update is due.  One pass over the vectors and no second reduction per step.
template <typename P>
__global__ void __launch_bounds__(CONS_BLOCK_MAX)
    k_mx_step_fs(float *xs, P *p, float *rs, const float *Ap, int vh, SolverState *st, const double *partials,
                 int nb, int *gate) {
  if (st->fs_gate == 0) return;  // written by the preceding kernel: uniform
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  cplx<float> xv[3], pv[3], rv[3], av[3];
  if (i < vh) {  // requested before the partials are summed
    Spin<float>::load(xs, i, xv);
    Spin<P>::load(p, i, pv);
    Spin<float>::load(rs, i, rv);
    Spin<float>::load(Ap, i, av);
  }
  double v[4];
  sum_partials<4>(partials, nb, v);
  const double pkp = v[0], c_tr = v[1], c_tt = v[2], rr = v[3];
  const double alpha = -st->fs_mxrsq / pkp;
  const double rnew = rr + 2.0 * alpha * c_tr + alpha * alpha * c_tt;
  const double beta = rnew / rr;
  const int it = st->fs_iter + 1;
  const double r = sqrt(fmax(rnew, 0.0));
  const double maxr = fmax(st->mx_maxr, r);  // idempotent: block 0 may have stored it already
  const bool update = !(rnew > 0) || r < st->mx_delta * maxr || rnew < st->rsqstop || it >= st->max_cg;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->pkp = pkp;
    st->iteration = it;
    st->mx_alpha = alpha;
    st->mx_rsq_old = rr;
    st->mx_beta = beta;
    st->mx_rsq = rnew;
    st->mx_maxr = maxr;
    if (update) *gate = 0;
  }
  if (i >= vh) return;
  const float al = (float)alpha, be = (float)beta;
  for (int c = 0; c < 3; ++c) {
    xv[c].re += al * pv[c].re;
    xv[c].im += al * pv[c].im;
    rv[c].re += al * av[c].re;
    rv[c].im += al * av[c].im;
  }
  Spin<float>::store(xs, i, xv);
  Spin<float>::store(rs, i, rv);
  if (update) return;
  for (int c = 0; c < 3; ++c) {
    pv[c].re = rv[c].re + be * pv[c].re;
    pv[c].im = rv[c].im + be * pv[c].im;
  }
  Spin<P>::store(p, i, pv);
}

// x += x_s ; x_s = 0 (per site: the FP64 and FP32 tiles order their reals
// differently)
__global__ void __launch_bounds__(BLAS_BLOCK) k_mx_accum(double *x, float *xs, int vh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  cplx<double> xv[3];
  cplx<float> sv[3];
  Spin<double>::load(x, i, xv);
  Spin<float>::load(xs, i, sv);
  for (int c = 0; c < 3; ++c) {
    xv[c].re += (double)sv[c].re;
    xv[c].im += (double)sv[c].im;
    sv[c] = cplx<float>{0.f, 0.f};
  }
  Spin<double>::store(x, i, xv);
  Spin<float>::store(xs, i, sv);
}

// r_s = (float) r ; p = r_s ; x_s = 0
template <typename P>
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_mx_init(float *rs, const double *r, P *p, float *xs, int vh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  cplx<double> rd[3];
  Spin<double>::load(r, i, rd);
  cplx<float> v[3];
  for (int c = 0; c < 3; ++c) {
    v[c].re = (float)rd[c].re;
    v[c].im = (float)rd[c].im;
  }
  Spin<float>::store(rs, i, v);
  Spin<P>::store(p, i, v);
  const cplx<float> z[3] = {{0, 0}, {0, 0}, {0, 0}};
  Spin<float>::store(xs, i, z);
}

// r_s = (float) r ; reduce r_s.p
template <typename P>
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_mx_rp(float *rs, const double *r, const P *p, int vh, ReduceCtl rc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s[1] = {0.0};
  if (i < vh) {
    cplx<float> v[3], pv[3];
    cplx<double> rd[3];
    Spin<P>::load(p, i, pv);
    Spin<double>::load(r, i, rd);
    for (int c = 0; c < 3; ++c) {
      v[c].re = (float)rd[c].re;
      v[c].im = (float)rd[c].im;
      s[0] += (double)v[c].re * pv[c].re + (double)v[c].im * pv[c].im;
    }
    Spin<float>::store(rs, i, v);
  }
  grid_reduce<1>(s, rc);
}

// p = r_s + beta (p - gamma r_s)
template <typename P>
__global__ void __launch_bounds__(BLAS_BLOCK)
    k_mx_newp(const float *rs, P *p, int vh, const SolverState *st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  const float be = (float)st->mx_beta, ga = (float)st->mx_gamma;
  cplx<float> rv[3], pv[3];
  Spin<float>::load(rs, i, rv);
  Spin<P>::load(p, i, pv);
  for (int c = 0; c < 3; ++c) {
    pv[c].re = rv[c].re + be * (pv[c].re - ga * rv[c].re);
    pv[c].im = rv[c].im + be * (pv[c].im - ga * rv[c].im);
  }
  Spin<P>::store(p, i, pv);
}

__global__ void k_mx_start(SolverState *s) {
  s->mx_rsq = s->rsq;
  s->mx_rsq_old = s->rsq;
  s->mx_maxr = sqrt(s->rsq);
  s->active = 1;
}

static double wall_now() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static const size_t HEAD = offsetof(SolverState, shifts);

static void read_state(const SolverState *d, SolverState &h) {
  HIP_CHECK(hipMemcpyAsync(ctx().pinned, d, sizeof(SolverState), hipMemcpyDeviceToHost, ctx().stream));
  HIP_CHECK(hipStreamSynchronize(ctx().stream));
  std::memcpy(&h, ctx().pinned, sizeof(SolverState));
}

static void write_state(SolverState *d, const SolverState &h) {
  std::memcpy(ctx().pinned, &h, sizeof(SolverState));
  HIP_CHECK(hipMemcpyAsync(d, ctx().pinned, sizeof(SolverState), hipMemcpyHostToDevice, ctx().stream));
  HIP_CHECK(hipStreamSynchronize(ctx().stream));
}

// Sloppy operator application Ap = -(4m^2 - D D) p, reducing p.Ap (and, for
// a single-reduction step, r.Ap, Ap.Ap and |r|^2 with r = rs).  The partials
// are left to the following step kernel (consumer_finalize).
static DslashExtra pkp_extra(double s, const void *p, float *rs, SolverState *st, const int *gate) {
  DslashExtra e;
  e.gate = gate;
  e.s = s;
  e.aux = p;
  e.r = rs;
  e.state = st;
  e.scalar_op = SOP_MX_PAP;
  e.consumer_finalize = true;
  return e;
}

static void sloppy_apply(const LinksT<float> &Lf, int parity, float *p, float *tmp, float *Ap, float *rs,
                         double s, SolverState *st, const int *gate, int epi) {
  DslashExtra e1;
  e1.gate = gate;
  launch_dslash<float>(Lf, 1 - parity, tmp, p, EPI_PLAIN, e1);
  launch_dslash<float>(Lf, parity, Ap, tmp, epi, pkp_extra(s, p, rs, st, gate));
}

static void sloppy_apply(const LinksT<float> &Lf, int parity, HalfSite *p, HalfSite *tmp, float *Ap,
                         float *rs, double s, SolverState *st, const int *gate, int epi) {
  DslashExtra e1;
  e1.gate = gate;
  launch_dslash_half(Lf, 1 - parity, tmp, nullptr, p, EPI_PLAIN, e1);
  launch_dslash_half(Lf, parity, nullptr, Ap, tmp, epi, pkp_extra(s, p, rs, st, gate));
}

// The same with 16-bit fixed-point links.
static void sloppy_apply(const LinksT<Q16> &Lq, int parity, float *p, float *tmp, float *Ap, float *rs,
                         double s, SolverState *st, const int *gate, int epi) {
  DslashExtra e1;
  e1.gate = gate;
  launch_dslash_q16(Lq, 1 - parity, tmp, p, EPI_PLAIN, e1);
  launch_dslash_q16(Lq, parity, Ap, tmp, epi, pkp_extra(s, p, rs, st, gate));
}

static void sloppy_apply(const LinksT<Q16> &Lq, int parity, HalfSite *p, HalfSite *tmp, float *Ap, float *rs,
                         double s, SolverState *st, const int *gate, int epi) {
  DslashExtra e1;
  e1.gate = gate;
  launch_dslash_q16_half(Lq, 1 - parity, tmp, nullptr, p, EPI_PLAIN, e1);
  launch_dslash_q16_half(Lq, parity, nullptr, Ap, tmp, epi, pkp_extra(s, p, rs, st, gate));
}

// P: storage of the sloppy vectors (float or HalfSite); LK: of the sloppy
// links (float, or Q16 16-bit fixed point).
template <typename P, typename LK>
static int invert_mixed(QPHIX_info_t *info, LinksT<double> *L, QPHIX_invert_arg_t *inv,
                        QPHIX_resid_arg_t *res, double mass, VectorT<double> *out,
                        VectorT<double> *in) {
  const double t0 = wall_now();
  const int vh = ctx().g.vh;
  const size_t nv = 6 * (size_t)vh;
  const LinksT<LK> *Lfp = [&]() -> const LinksT<LK> * {
    if constexpr (std::is_same<LK, Q16>::value) return sloppy_links16(L);
    else return sloppy_links_solver(L);
  }();
  const size_t psz = sizeof(P) == sizeof(HalfSite) ? sizeof(HalfSite) * (size_t)vh : nv * sizeof(float);
  char *ws = Lfp ? (char *)ws_get(2 * nv * sizeof(double) + 2 * psz + 3 * nv * sizeof(float) +
                                   sizeof(SolverState) + 1024)
                 : nullptr;
  if (!ws) {  // out of device memory, error recorded
    info->status = QPHIX_MEM_ERROR;
    return 0;
  }
  const LinksT<LK> &Lf = *Lfp;
  double *r64 = (double *)ws, *tmp64 = r64 + nv;
  P *p = (P *)(tmp64 + nv);
  P *tmps = (P *)((char *)p + psz);
  float *rs = (float *)((char *)tmps + psz), *xs = rs + nv, *Ap = xs + nv;
  SolverState *dst = (SolverState *)(((uintptr_t)(Ap + nv) + 255) & ~(uintptr_t)255);
  int *gate = &dst->active;
  const double msq_x4 = 4.0 * mass * mass;
  const int nb = (vh + BLAS_BLOCK - 1) / BLAS_BLOCK;
  const int nbc = cons_nb(vh);
  const int nb_dslash = (vh + DSLASH_BLOCK - 1) / DSLASH_BLOCK;
  const char *de = std::getenv("QPHIX_MIXED_DELTA");
  const int batch = 16;
  const bool fs = (fewsums_mode() & 2) != 0;  // single-reduction sloppy steps

  SolverState hs{};
  hs.niter = inv->max;
  hs.max_cg = inv->max * inv->nrestart;
  hs.mx_delta = de ? std::atof(de) : 0.1;
  // Stopping rule of the reference's CG (ks_d_congrad_fn.cpp:424-427): the
  // squared relative true residual below resid and the relative residue
  // (ks_blas_utils_c.cpp:110-149) below relresid, a test with a target <= 0
  // counting as met.  Both are evaluated on the FP64 true residual at every
  // reliable update; the sloppy steps only decide when to take one.
  hs.rsqmin = res->resid;
  hs.relrsqmin = res->relresid;
  hs.vol = (double)ctx().g.nx * ctx().g.ny * ctx().g.nz * ctx().g.nt;
  auto converged = [&](const SolverState &h) {
    return (res->resid <= 0 || h.final_rsq < res->resid) &&
           (res->relresid <= 0 || h.final_rel < res->relresid);
  };
  bool all_met = true;
  int todo[2], ntodo;
  if (inv->parity == QPHIX_EVENODD) { todo[0] = 0; todo[1] = 1; ntodo = 2; }
  else { todo[0] = inv->parity == QPHIX_EVEN ? 0 : 1; ntodo = 1; }
  int total = 0, updates = 0, iteration = 0;
  info->status = QPHIX_SUCCESS;
  for (int pass = 0; pass < ntodo; ++pass) {
    const int parity = todo[pass], other = 1 - parity;
    double *x = out->half[parity];
    const double *b = in->half[parity];
    BatchGraph steps;  // this parity's batch of sloppy CG steps
    hs.iteration = 0;
    hs.active = 0;
    write_state(dst, hs);
    blas_norm2<double>(b, dst, SOP_SRC_NORM);
    read_state(dst, hs);
    if (hs.src_norm == 0.0) {
      HIP_CHECK(hipMemsetAsync(x, 0, nv * sizeof(double), ctx().stream));
      hs.final_rsq = 0;
      hs.final_rel = 0;
      iteration = 0;
      continue;
    }
    hs.rsqstop = res->resid > 0 ? res->resid * hs.src_norm : 0.0;
    write_state(dst, hs);
    bool first = true;
    while (true) {
      // reliable update: x += x_s, r = b - M x in FP64
      if (!first) hipLaunchKernelGGL(k_mx_accum, dim3(nb), dim3(BLAS_BLOCK), 0, ctx().stream, x, xs, vh);
      DslashExtra e1;
      launch_dslash<double>(*L, other, tmp64, x, EPI_PLAIN, e1);
      DslashExtra e2;
      e2.s = -msq_x4;
      e2.aux = x;
      e2.b = b;
      e2.r = r64;
      e2.p = r64;
      e2.state = dst;
      e2.scalar_op = SOP_MX_TRUE;
      launch_dslash<double>(*L, parity, r64, tmp64, EPI_RESTART, e2);
      read_state(dst, hs);
      if (converged(hs) || hs.iteration >= hs.max_cg || !std::isfinite(hs.rsq)) {
        if (!converged(hs)) all_met = false;
        break;
      }
      if (first) {
        hipLaunchKernelGGL(k_mx_init<P>, dim3(nb), dim3(BLAS_BLOCK), 0, ctx().stream, rs, r64, p, xs, vh);
        hipLaunchKernelGGL(k_mx_start, dim3(1), dim3(1), 0, ctx().stream, dst);
        first = false;
      } else {
        const ReduceCtl rc = reduce_ctl(dst, SOP_MX_RP);
        hipLaunchKernelGGL(k_mx_rp<P>, dim3(nb), dim3(BLAS_BLOCK), 0, ctx().stream, rs, r64, p, vh, rc);
        reduce_finish(1, nb, rc, nullptr);
        hipLaunchKernelGGL(k_mx_newp<P>, dim3(nb), dim3(BLAS_BLOCK), 0, ctx().stream, rs, p, vh, dst);
        ++updates;
      }
      // sloppy CG steps until the device flags the next reliable update
      while (true) {
        steps.run([&]() {
          for (int k = 0; k < batch; ++k) {
            if (fs) {
              sloppy_apply(Lf, parity, p, tmps, Ap, rs, -msq_x4, dst, gate, EPI_PKP_FS);
              hipLaunchKernelGGL(k_mx_step_fs<P>, dim3(nbc), dim3(cons_block()), 0, ctx().stream, xs, p, rs,
                                 Ap, vh, dst, ctx().partials, nb_dslash, gate);
            } else {
              sloppy_apply(Lf, parity, p, tmps, Ap, rs, -msq_x4, dst, gate, EPI_PKP);
              hipLaunchKernelGGL(k_mx_step_c<P>, dim3(nbc), dim3(cons_block()), 0, ctx().stream, xs, p, rs,
                                 Ap, vh, dst, ctx().partials, nb_dslash, partials_b(), gate);
              hipLaunchKernelGGL(k_mx_axpy_c<P>, dim3(nbc), dim3(cons_block()), 0, ctx().stream, rs, p, vh,
                                 dst, partials_b(), nbc, gate);
            }
          }
        });
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(ctx().pinned, dst, HEAD, hipMemcpyDeviceToHost, ctx().stream));
        HIP_CHECK(hipStreamSynchronize(ctx().stream));
        std::memcpy(&hs, ctx().pinned, HEAD);
        if (hs.active == 0) break;
      }
    }
    iteration = hs.iteration;
    total += iteration;
    if (!std::isfinite(hs.rsq)) info->status = QPHIX_FAIL;
  }
  HIP_CHECK(hipStreamSynchronize(ctx().stream));
  res->final_rsq = hs.final_rsq;
  res->final_rel = hs.final_rel;
  res->size_r = hs.final_rsq;
  res->size_relr = hs.final_rel;
  res->final_iter = total;
  res->final_restart = updates;
  // no counterpart in the reference (its CG leaves status alone): a solve cut
  // off by max * nrestart before meeting its target reports failure
  if (!all_met) info->status = QPHIX_FAIL;
  info->final_sec = wall_now() - t0;
  info->final_flop = 1187.0 * total * 2.0 * vh;
  return iteration;
}

// Sloppy vectors of sloppy_bits (32: FP32, 16: FP16-stored); sloppy links of
// sloppy_link_bits() (32: FP32, 16: 16-bit fixed point).
static int mixed_dispatch(QPHIX_info_t *info, LinksT<double> *L, QPHIX_invert_arg_t *inv,
                          QPHIX_resid_arg_t *res, double mass, VectorT<double> *out, VectorT<double> *in,
                          int sloppy_bits) {
  const bool l16 = sloppy_link_bits() == 16;
  if (sloppy_bits == 16)
    return l16 ? invert_mixed<HalfSite, Q16>(info, L, inv, res, mass, out, in)
               : invert_mixed<HalfSite, float>(info, L, inv, res, mass, out, in);
  return l16 ? invert_mixed<float, Q16>(info, L, inv, res, mass, out, in)
             : invert_mixed<float, float>(info, L, inv, res, mass, out, in);
}

int mixed_solve(QPHIX_info_t *info, LinksT<double> *L, QPHIX_invert_arg_t *inv,
                QPHIX_resid_arg_t *res, double mass, VectorT<double> *out, VectorT<double> *in,
                int sloppy_bits) {
  return mixed_dispatch(info, L, inv, res, mass, out, in, sloppy_bits);
}

} // namespace qphix

using namespace qphix;

extern "C" {

double QPHIX_D3_asqtad_dslash_sloppy(QPHIX_D3_FermionLinksAsqtad *L, QPHIX_F3_ColorVector *out,
                                     QPHIX_F3_ColorVector *in, QPHIX_evenodd_t parity, int link_bits,
                                     int reps) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called: asqtad_dslash_sloppy", -1.0);
  QPHIX_REQUIRE(parity == QPHIX_EVEN || parity == QPHIX_ODD, "asqtad_dslash_sloppy: parity must be EVEN or ODD",
                -1.0);
  QPHIX_REQUIRE(link_bits == 32 || link_bits == 16, "asqtad_dslash_sloppy: link_bits must be 32 or 16", -1.0);
  const int cb = parity == QPHIX_EVEN ? 0 : 1;
  QPHIX_REQUIRE(L && out && in && out->half[cb] && in->half[1 - cb] && L->fat[cb],
                "asqtad_dslash_sloppy: bad arguments", -1.0);
  if (reps < 1) reps = 1;
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  DslashExtra ex;
  if (link_bits == 16) {
    const LinksT<Q16> *Lqp = sloppy_links16(L);
    if (!Lqp) return -1.0;  // error recorded
    const LinksT<Q16> &Lq = *Lqp;
    launch_dslash_q16(Lq, cb, out->half[cb], in->half[1 - cb], EPI_PLAIN, ex);
    HIP_CHECK(hipEventRecord(e0, ctx().stream));
    for (int r = 0; r < reps; ++r) launch_dslash_q16(Lq, cb, out->half[cb], in->half[1 - cb], EPI_PLAIN, ex);
  } else {
    const LinksT<float> *Lfp = sloppy_links_solver(L);
    if (!Lfp) return -1.0;  // error recorded
    const LinksT<float> &Lf = *Lfp;
    launch_dslash<float>(Lf, cb, out->half[cb], in->half[1 - cb], EPI_PLAIN, ex);
    HIP_CHECK(hipEventRecord(e0, ctx().stream));
    for (int r = 0; r < reps; ++r) launch_dslash<float>(Lf, cb, out->half[cb], in->half[1 - cb], EPI_PLAIN, ex);
  }
  HIP_CHECK(hipEventRecord(e1, ctx().stream));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  return (double)ms / reps;
}

int QPHIX_D3_asqtad_invert_mixed(QPHIX_info_t *info, QPHIX_D3_FermionLinksAsqtad *L,
                                 QPHIX_invert_arg_t *inv, QPHIX_resid_arg_t *res, double mass,
                                 QPHIX_D3_ColorVector *out, QPHIX_D3_ColorVector *in,
                                 int sloppy_bits) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called: asqtad_invert_mixed", 0);
  QPHIX_REQUIRE(info && L && inv && res && out && in, "asqtad_invert_mixed: null argument", 0);
  info->status = QPHIX_FAIL;
  QPHIX_REQUIRE(sloppy_bits == 32 || sloppy_bits == 16, "asqtad_invert_mixed: sloppy precision must be 32 or 16", 0);
  QPHIX_REQUIRE(in->parity == out->parity && in->parity == (int)inv->parity,
                "asqtad_invert_mixed: in, out and inv_arg parities must agree", 0);
  QPHIX_REQUIRE(inv->max > 0 && inv->nrestart > 0, "asqtad_invert_mixed: bad iteration limits", 0);
  QPHIX_REQUIRE(L->fat[0] && L->fat[1], "asqtad_invert_mixed: both link parities are required", 0);
  for (int cb = 0; cb < 2; ++cb)
    if (inv->parity & (cb ? QPHIX_ODD : QPHIX_EVEN))
      QPHIX_REQUIRE(out->half[cb] && in->half[cb], "asqtad_invert_mixed: vector lacks a solved parity", 0);
  return mixed_dispatch(info, L, inv, res, mass, out, in, sloppy_bits);
}

} // extern "C"
