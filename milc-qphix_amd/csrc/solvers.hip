// solvers.hip -- the Dslash entry points, CG and multi-shift CG drivers.
//
// The drivers keep the reference's control flow (restart rule, iteration
// counting C9, initial guess C10, squared relative residual C8, 4m^2 mass
// convention C7) but run every vector operation and every scalar recurrence on
// the GPU.  The host queues batches of iterations and only reads the solver
// state between batches; a device flag turns the remaining kernels of a batch
// into no-ops once the reference would leave its inner loop, so iteration
// counts are exact.
#include "blas.h"
#include "dslash_kernel.h"
#include "handles.h"
#include "qphix_mi355x.h"

#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

namespace qphix {

static double wall_now() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static const size_t STATE_HEAD = offsetof(SolverState, shifts);

static void state_upload(SolverState *d, const SolverState &h, size_t bytes = sizeof(SolverState)) {
  std::memcpy(ctx().pinned, &h, bytes);
  HIP_CHECK(hipMemcpyAsync(d, ctx().pinned, bytes, hipMemcpyHostToDevice, ctx().stream));
  HIP_CHECK(hipStreamSynchronize(ctx().stream));
}

static void state_read(const SolverState *d, SolverState &h, size_t bytes = STATE_HEAD) {
  HIP_CHECK(hipMemcpyAsync(ctx().pinned, d, bytes, hipMemcpyDeviceToHost, ctx().stream));
  HIP_CHECK(hipStreamSynchronize(ctx().stream));
  std::memcpy(&h, ctx().pinned, bytes);
}

static int steps_per_batch() {
  const char *e = std::getenv("QPHIX_BATCH");
  const int k = e ? std::atoi(e) : 16;
  return k < 1 ? 1 : k;
}

template <typename T> static T rnd_t(double v) { return (T)v; }

// ------------------------------------------------------------------ Dslash
template <typename T> static void dslash_api(LinksT<T> *L, VectorT<T> *out, VectorT<T> *in, int parity) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called:asqtad_dslash");
  QPHIX_REQUIRE(L && out && in, "asqtad_dslash: null argument");
  // qphix_d3.cpp:1252-1253 asserts parity != EVENODD and out->parity ==
  // QPHIX_OPP_PAR(in->parity); that macro maps EVEN to 9 and ODD to 10, so the
  // reference only accepts EVENODD vectors.  Here any vectors holding the
  // needed halves are accepted (a superset of the reference's contract).
  QPHIX_REQUIRE(parity == QPHIX_EVEN || parity == QPHIX_ODD, "asqtad_dslash: parity must be EVEN or ODD");
  const int cb = parity == QPHIX_EVEN ? 0 : 1;
  QPHIX_REQUIRE(out->half[cb] && in->half[1 - cb],
                "asqtad_dslash: out must hold the output parity and in the opposite one");
  DslashExtra ex;
  launch_dslash<T>(*L, cb, out->half[cb], in->half[1 - cb], EPI_PLAIN, ex);
  sync_if_needed();
}

template <typename T>
static double dslash_time(LinksT<T> *L, VectorT<T> *out, VectorT<T> *in, int parity, int reps) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called:dslash_time", -1.0);
  const int cb = parity == QPHIX_EVEN ? 0 : 1;
  QPHIX_REQUIRE(L && out && in && out->half[cb] && in->half[1 - cb] && L->fat[cb],
                "dslash_time: bad arguments", -1.0);
  if (reps < 1) reps = 1;
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  DslashExtra ex;
  HIP_CHECK(hipEventRecord(e0, ctx().stream));
  for (int r = 0; r < reps; ++r) launch_dslash<T>(*L, cb, out->half[cb], in->half[1 - cb], EPI_PLAIN, ex);
  HIP_CHECK(hipEventRecord(e1, ctx().stream));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  return (double)ms / reps;
}

// ------------------------------------------------------------------ CG
// QPHIX_*3_asqtad_invert, ks_d_congrad_fn.cpp:221-619.
// ws_ext: a workspace of invert_ws_bytes() the caller provides (the
// no-allocation entry point); then nothing is allocated and the iteration
// batches are launched directly instead of through a captured graph.
template <typename T> static size_t cg_ws_bytes() {
  return 4 * (size_t)6 * ctx().g.vh * sizeof(T) + sizeof(SolverState) + 256;
}

template <typename T>
static int invert(QPHIX_info_t *info, LinksT<T> *L, QPHIX_invert_arg_t *inv, QPHIX_resid_arg_t *res,
                  T mass, VectorT<T> *out, VectorT<T> *in, char *ws_ext = nullptr) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called:asqtad_invert", 0);
  QPHIX_REQUIRE(info && L && inv && res && out && in, "asqtad_invert: null argument", 0);
  info->status = QPHIX_FAIL;
  QPHIX_REQUIRE(in->parity == out->parity && in->parity == (int)inv->parity,
                "asqtad_invert: in, out and inv_arg parities must agree", 0);
  QPHIX_REQUIRE(inv->max > 0, "asqtad_invert: inv_arg->max must be positive", 0);
  for (int cb = 0; cb < 2; ++cb)
    if (inv->parity & (cb ? QPHIX_ODD : QPHIX_EVEN))
      QPHIX_REQUIRE(out->half[cb] && in->half[cb] && L->fat[0] && L->fat[1],
                    "asqtad_invert: vectors and both link parities are required", 0);
  const double t0 = wall_now();
  const int vh = ctx().g.vh;
  const size_t nv = 6 * (size_t)vh;
  // workspace: ttt, tmp, p, r, then the device state
  char *ws = ws_ext ? ws_ext : (char *)ws_get(cg_ws_bytes<T>());
  if (!ws) {  // out of device memory, error recorded
    info->status = QPHIX_MEM_ERROR;
    return 0;
  }
  T *ttt = (T *)ws, *tmp = ttt + nv, *p = tmp + nv, *r = p + nv;
  SolverState *dst = (SolverState *)(((uintptr_t)(r + nv) + 255) & ~(uintptr_t)255);
  int *gate = &dst->active;

  SolverState hs{};
  hs.niter = inv->max;
  hs.max_cg = inv->nrestart * inv->max;
  hs.single_coef = sizeof(T) == 4;
  hs.rsqmin = (double)(T)res->resid;
  hs.relrsqmin = (double)(T)res->relresid;
  hs.msq_x4 = (double)(T)(4.0 * (double)mass * (double)mass);
  hs.vol = (double)ctx().g.nx * ctx().g.ny * ctx().g.nz * ctx().g.nt;
  const int max_restarts = inv->nrestart;

  int todo[2], ntodo;
  if (inv->parity == QPHIX_EVENODD) { todo[0] = 0; todo[1] = 1; ntodo = 2; }
  else { todo[0] = inv->parity == QPHIX_EVEN ? 0 : 1; ntodo = 1; }

  const int batch = steps_per_batch();
  const int nb_dslash = (vh + DSLASH_BLOCK - 1) / DSLASH_BLOCK;
  // single-reduction steps (FEWSUMS) when asked for and the relative
  // residue, which needs its own reduction of the updated vectors, is off
  const bool fs = (fewsums_mode() & 1) && hs.relrsqmin <= 0;
  int total_iters = 0, iteration = 0, nrestart = 0;
  info->final_flop = 0;
  info->status = QPHIX_SUCCESS;
  for (int pass = 0; pass < ntodo; ++pass) {
    const int parity = todo[pass], other = 1 - parity;
    T *xd = out->half[parity];
    const T *bs = in->half[parity];
    BatchGraph steps;  // this parity's batch of CG steps
    total_iters += iteration;
    iteration = 0;
    nrestart = 0;
    hs.iteration = 0;
    hs.active = 0;
    hs.size_r = 0;
    hs.size_relr = 0;
    state_upload(dst, hs, STATE_HEAD);
    blas_norm2<T>(bs, dst, SOP_SRC_NORM);
    state_read(dst, hs);
    if (hs.src_norm == 0.0) {  // trivial source, ks_d_congrad_fn.cpp:327-350
      HIP_CHECK(hipMemsetAsync(xd, 0, nv * sizeof(T), ctx().stream));
      hs.final_rsq = 0;
      hs.final_rel = 0;
      continue;
    }
    while (true) {
      // (re)start: true residual r = b - M x, p = r (ks_d_congrad_fn.cpp:372-419)
      DslashExtra e1;
      launch_dslash<T>(*L, other, tmp, xd, EPI_PLAIN, e1);
      DslashExtra e2;
      e2.s = -hs.msq_x4;
      e2.aux = xd;
      e2.b = bs;
      e2.r = r;
      e2.p = p;
      e2.state = dst;
      e2.scalar_op = SOP_CG_RESTART;
      launch_dslash<T>(*L, parity, ttt, tmp, EPI_RESTART, e2);
      state_read(dst, hs);
      iteration = hs.iteration;
      const bool conv = (hs.rsqmin <= 0 || hs.rsqmin > hs.final_rsq) &&
                        (hs.relrsqmin <= 0 || hs.relrsqmin > hs.final_rel);
      if (iteration >= hs.max_cg || nrestart >= max_restarts || conv) break;
      if (iteration != 0) nrestart++;
      set_int(gate, 1);
      // CG steps until the device clears the gate (restart rule met)
      while (true) {
        steps.run([&]() {
          for (int s = 0; s < batch; ++s) {
            DslashExtra a1;
            a1.gate = gate;
            launch_dslash<T>(*L, other, tmp, p, EPI_PLAIN, a1);
            DslashExtra a2;
            a2.gate = gate;
            a2.s = -hs.msq_x4;
            a2.aux = p;
            a2.state = dst;
            a2.scalar_op = SOP_CG_PKP;
            a2.consumer_finalize = true;
            if (fs) {
              a2.r = r;
              launch_dslash<T>(*L, parity, ttt, tmp, EPI_PKP_FS, a2);
              blas_cg_step_fs<T>(xd, p, r, ttt, dst, gate, nb_dslash);
            } else {
              launch_dslash<T>(*L, parity, ttt, tmp, EPI_PKP, a2);
              blas_cg_step_c<T>(xd, p, r, ttt, dst, gate, nb_dslash);
            }
          }
        }, ws_ext == nullptr);
        state_read(dst, hs);
        if (hs.active == 0) break;
      }
      iteration = hs.iteration;
    }
    iteration = hs.iteration;
  }
  HIP_CHECK(hipStreamSynchronize(ctx().stream));
  total_iters += iteration;
  res->final_rsq = hs.final_rsq;
  res->final_rel = hs.final_rel;
  res->size_r = hs.size_r;
  res->size_relr = hs.size_relr;
  res->final_iter = total_iters;
  res->final_restart = nrestart;
  info->final_sec = wall_now() - t0;
  info->final_flop = 1187.0 * total_iters * 2.0 * vh;  // ks_d_congrad_fn.cpp:610
  if (!std::isfinite(hs.final_rsq)) info->status = QPHIX_FAIL;
  return iteration;
}

// ------------------------------------------------------------------ multi-shift
// QPHIX_*3_asqtad_invert_multi, ks_multicg_offset.cpp:205-689.
template <typename T>
static int invert_multi(QPHIX_info_t *info, LinksT<T> *L, QPHIX_invert_arg_t *inv,
                        QPHIX_resid_arg_t **res, T *mass, int nshift,
                        typename Handles<T>::V **out, VectorT<T> *in) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called:asqtad_invert_multi", 0);
  if (nshift == 0) return 0;  // ks_multicg_offset.cpp:218-225
  QPHIX_REQUIRE(info && L && inv && res && mass && out && in && out[0],
                "asqtad_invert_multi: null argument", 0);
  info->status = QPHIX_FAIL;
  QPHIX_REQUIRE(nshift > 0 && nshift <= MAXNDMASS, "asqtad_invert_multi: bad number of offsets", 0);
  QPHIX_REQUIRE(in->parity == out[0]->parity && in->parity == (int)inv->parity,
                "asqtad_invert_multi: in, out[0] and inv_arg parities must agree", 0);
  QPHIX_REQUIRE(L->fat[0] && L->fat[1], "asqtad_invert_multi: both link parities are required", 0);
  for (int cb = 0; cb < 2; ++cb) {
    if (!(inv->parity & (cb ? QPHIX_ODD : QPHIX_EVEN))) continue;
    QPHIX_REQUIRE(in->half[cb], "asqtad_invert_multi: source lacks a solved parity", 0);
    for (int j = 0; j < nshift; ++j)
      QPHIX_REQUIRE(out[j] && out[j]->half[cb], "asqtad_invert_multi: out[j] lacks a solved parity", 0);
  }
  const double t0 = wall_now();
  const int vh = ctx().g.vh;
  const size_t nv = 6 * (size_t)vh;
  // shifted systems updated up to `fuse` steps at a time (fuse - 1 saved residuals)
  const int fuse = msx_fuse();
  char *ws = (char *)ws_get((3 + (size_t)nshift + (size_t)(fuse - 1)) * nv * sizeof(T) + sizeof(SolverState) +
                            2 * MAXNDMASS * sizeof(T *) + 512);
  if (!ws) {  // out of device memory, error recorded
    info->status = QPHIX_MEM_ERROR;
    return 0;
  }
  T *ttt = (T *)ws, *tmp = ttt + nv, *r = tmp + nv;
  T *pm0 = r + nv;
  T *rprev = pm0 + nshift * nv;
  SolverState *dst = (SolverState *)(((uintptr_t)(rprev + (size_t)(fuse - 1) * nv) + 255) & ~(uintptr_t)255);
  T **dptr = (T **)(dst + 1);  // [0..nshift) x_j, [nshift..2nshift) p_j
  int *gate = &dst->active;

  SolverState hs{};
  hs.single_coef = sizeof(T) == 4;
  hs.nshift = nshift;
  hs.fuse = fuse;
  double offset_low = 1.0e+20;
  int j_low = -1;
  for (int j = 0; j < nshift; ++j) {
    hs.shifts[j] = (double)(T)((double)mass[j] * (double)mass[j] * 4.0);
    if (hs.shifts[j] < offset_low) { offset_low = hs.shifts[j]; j_low = j; }
  }
  for (int j = 0; j < nshift; ++j)
    if (j != j_low) hs.shifts[j] = (double)(T)((T)hs.shifts[j] - (T)hs.shifts[j_low]);
  hs.j_low = j_low;
  hs.offset_low = offset_low;
  const double rsqmin = (double)(T)res[0]->resid;
  const int max_cg = inv->nrestart * inv->max;
  hs.max_cg = max_cg;

  int todo[2], ntodo;
  if (inv->parity == QPHIX_EVENODD) { todo[0] = 0; todo[1] = 1; ntodo = 2; }
  else { todo[0] = inv->parity == QPHIX_EVEN ? 0 : 1; ntodo = 1; }
  const int batch = steps_per_batch();
  int iteration = 0, total_iters = 0;
  info->status = QPHIX_SUCCESS;
  for (int pass = 0; pass < ntodo; ++pass) {
    const int parity = todo[pass], other = 1 - parity;
    const T *bs = in->half[parity];
    std::vector<T *> hp(2 * nshift);
    for (int j = 0; j < nshift; ++j) {
      hp[j] = out[j]->half[parity];
      hp[nshift + j] = pm0 + j * nv;
      HIP_CHECK(hipMemsetAsync(hp[j], 0, nv * sizeof(T), ctx().stream));
      HIP_CHECK(hipMemcpyAsync(hp[nshift + j], bs, nv * sizeof(T), hipMemcpyDeviceToDevice, ctx().stream));
    }
    HIP_CHECK(hipMemcpyAsync(r, bs, nv * sizeof(T), hipMemcpyDeviceToDevice, ctx().stream));
    std::memcpy(ctx().pinned, hp.data(), 2 * nshift * sizeof(T *));
    HIP_CHECK(hipMemcpyAsync(dptr, ctx().pinned, 2 * nshift * sizeof(T *), hipMemcpyHostToDevice,
                             ctx().stream));
    HIP_CHECK(hipStreamSynchronize(ctx().stream));
    iteration = 0;
    hs.iteration = 0;
    hs.active = 0;
    state_upload(dst, hs, STATE_HEAD);
    blas_norm2<T>(bs, dst, SOP_SRC_NORM);
    state_read(dst, hs);
    if (hs.src_norm == 0.0) { hs.rsq = 0; continue; }
    hs.rsq = hs.src_norm;
    hs.rsqstop = rsqmin * hs.src_norm;
    hs.num_now = nshift;
    hs.converged = 0;
    hs.active = 1;
    hs.pend = 0;
    hs.defer = 0;
    for (int j = 0; j < nshift; ++j) {
      hs.zeta_im1[j] = hs.zeta_i[j] = 1.0;
      hs.beta_im1[j] = -1.0;
      hs.alpha[j] = 0.0;
      hs.finished[j] = 0;
    }
    state_upload(dst, hs);
    T *plow = hp[nshift + j_low];
    int launched = 0;
    auto enqueue = [&](int nstep) {
      for (int s = 0; s < nstep; ++s) {
        DslashExtra a1;
        a1.gate = gate;
        launch_dslash<T>(*L, other, tmp, plow, EPI_PLAIN, a1);
        DslashExtra a2;
        a2.gate = gate;
        a2.s = -offset_low;
        a2.aux = plow;
        a2.state = dst;
        a2.scalar_op = SOP_MS_PKP;
        launch_dslash<T>(*L, parity, ttt, tmp, EPI_PKP, a2);
        blas_ms_update_r<T>(r, ttt, dst, gate);
        if (fuse > 1) blas_ms_update_pm_def<T>(dptr, dptr + nshift, r, rprev, nshift, dst, gate, false);
        else blas_ms_update_pm<T>(dptr, dptr + nshift, r, nshift, dst, gate);
      }
    };
    BatchGraph steps;  // full batches; the last one, cut at max_cg, is enqueued directly
    while (launched < max_cg) {
      const int nstep = std::min(batch, max_cg - launched);
      if (nstep == batch) steps.run([&]() { enqueue(batch); });
      else enqueue(nstep);
      launched += nstep;
      state_read(dst, hs);
      if (hs.active == 0) break;
    }
    iteration = hs.iteration;
    total_iters += iteration;
    // the shifted systems' pending steps, before the closing update
    if (fuse > 1) blas_ms_update_pm_def<T>(dptr, dptr + nshift, r, rprev, nshift, dst, gate, true);
    if (hs.converged) blas_ms_update_dest<T>(dptr, dptr + nshift, nshift, dst);
    HIP_CHECK(hipStreamSynchronize(ctx().stream));
  }
  for (int j = 0; j < nshift; ++j) {
    res[j]->final_rsq = hs.src_norm != 0 ? (double)(T)hs.rsq / hs.src_norm : 0.0;
    res[j]->size_r = res[j]->final_rsq;
    res[j]->size_relr = res[j]->final_rel;
    res[j]->final_iter = iteration;
    res[j]->final_restart = 0;
  }
  info->final_sec = wall_now() - t0;
  info->final_flop = (1205.0 + 15.0 * nshift) * total_iters * 2.0 * vh;  // ks_multicg_offset.cpp:554
  if (!std::isfinite(hs.rsq)) info->status = QPHIX_FAIL;
  return iteration;
}

// ------------------------------------------------------------------ copy timing
__global__ void k_copy16(float4 *dst, const float4 *src, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

static double copy_time(size_t bytes, int reps) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called:copy_time", -1.0);
  const size_t n = bytes / 16;
  float4 *a = (float4 *)dev_alloc(n * 16), *b = (float4 *)dev_alloc(n * 16);
  if (!a || !b) { dev_free(a); dev_free(b); return -1; }
  HIP_CHECK(hipMemsetAsync(a, 0, n * 16, ctx().stream));
  const int grid = ctx().n_cu * 16;
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, ctx().stream, b, a, n);
  HIP_CHECK(hipEventRecord(e0, ctx().stream));
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, ctx().stream, b, a, n);
  HIP_CHECK(hipEventRecord(e1, ctx().stream));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  dev_free(a);
  dev_free(b);
  return (double)ms / reps;
}

} // namespace qphix

using namespace qphix;

extern "C" {

#define QPHIX_SOLVER_API(P, T)                                                                   \
  void QPHIX_##P##3_asqtad_dslash(QPHIX_##P##3_FermionLinksAsqtad *L,                            \
                                  QPHIX_##P##3_ColorVector *out, QPHIX_##P##3_ColorVector *in,   \
                                  QPHIX_evenodd_t parity) {                                      \
    dslash_api<T>(L, out, in, parity);                                                           \
  }                                                                                              \
  double QPHIX_##P##3_asqtad_dslash_time(QPHIX_##P##3_FermionLinksAsqtad *L,                     \
                                         QPHIX_##P##3_ColorVector *out,                          \
                                         QPHIX_##P##3_ColorVector *in, QPHIX_evenodd_t parity,   \
                                         int reps) {                                             \
    return dslash_time<T>(L, out, in, parity, reps);                                             \
  }                                                                                              \
  int QPHIX_##P##3_asqtad_invert(QPHIX_info_t *info, QPHIX_##P##3_FermionLinksAsqtad *L,         \
                                 QPHIX_invert_arg_t *inv, QPHIX_resid_arg_t *res, T mass,        \
                                 QPHIX_##P##3_ColorVector *out, QPHIX_##P##3_ColorVector *in) {  \
    return invert<T>(info, L, inv, res, mass, out, in);                                          \
  }                                                                                              \
  int QPHIX_##P##3_asqtad_invert_multi(QPHIX_info_t *info, QPHIX_##P##3_FermionLinksAsqtad *L,   \
                                       QPHIX_invert_arg_t *inv, QPHIX_resid_arg_t *res[],        \
                                       T *mass, int n, QPHIX_##P##3_ColorVector *out[],          \
                                       QPHIX_##P##3_ColorVector *in) {                           \
    return invert_multi<T>(info, L, inv, res, mass, n, out, in);                                 \
  }

QPHIX_SOLVER_API(D, double)
QPHIX_SOLVER_API(F, float)

double QPHIX_MI355X_copy_time(size_t bytes, int reps) { return copy_time(bytes, reps); }

} // extern "C"
