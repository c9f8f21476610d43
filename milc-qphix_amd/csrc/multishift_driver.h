// multishift_driver.h -- host driver of the batched multi-shift solver.
// Textually part of multishift_batch.hip (uses its kernels); not a public header.
#pragma once

struct BatchStats {
  int ms_iters[8];
  int updates[8];
  double final_rsq[8];                  // base system's last true |r|^2 / |b|^2
  int shift_iters[8][MAXNDMASS];        // step at which each shift converged
};

// One batch of multi-shift steps for NR sources in precision T: kernels only
// (no memsets or copies), every argument fixed for the chunk and phase, every
// scalar read from the device states -- so it can be captured as a hipGraph
// and replayed; a step whose source is paused or finished is a no-op through
// its gate.
template <int NR, typename T>
static void msx_batch(const LinksT<T> &Lk, int parity, int batch, T *p, T *xs, T *rs, T *Ap, T *tmp,
                      T *rprev, int j_low, double offset_low, SolverState *st, int K, int nbd, int nbr,
                      int nlog, int nshift) {
  const int vh = ctx().g.vh;
  const size_t nvr = 6 * (size_t)vh * NR;
  const T *plow = p + (size_t)j_low * nvr;
  for (int k = 0; k < batch; ++k) {
    mrhs_dslash<NR, T>(Lk, 1 - parity, tmp, plow, EPI_PLAIN, (T)0, nullptr, st, K);
    mrhs_dslash<NR, T>(Lk, parity, Ap, tmp, EPI_PKP, (T)-offset_low, plow, st, K);
    hipLaunchKernelGGL(k_finalize_multi<NR>, dim3(1), dim3(FINALIZE_BLOCK), 0, ctx().stream,
                       ctx().partials, nbd, st, (int)SOP_MS_PKP, k);
    hipLaunchKernelGGL((k_msx_update_r<T, NR>), dim3(nbr), dim3(BLAS_BLOCK), 0, ctx().stream, rs, Ap, vh, st,
                       ctx().partials, nlog);
    hipLaunchKernelGGL(k_finalize_multi<NR>, dim3(1), dim3(FINALIZE_BLOCK), 0, ctx().stream,
                       ctx().partials, nbr, st, (int)SOP_MSX_RSQ2, k);
    if (msx_staged<NR>(nshift)) {
      const size_t lds = (size_t)NR * nshift * MSX_NC * sizeof(double);
      const dim3 g(msx_pm_grid(k_msx_update_pm<T, NR, true>, vh * NR, lds));
      hipLaunchKernelGGL((k_msx_update_pm<T, NR, true>), g, dim3(BLAS_BLOCK), (unsigned)lds, ctx().stream, xs, p, rs,
                         rprev, vh, nshift, st);
    } else {
      const dim3 g(msx_pm_grid(k_msx_update_pm<T, NR, false>, vh * NR, 0));
      hipLaunchKernelGGL((k_msx_update_pm<T, NR, false>), g, dim3(BLAS_BLOCK), 0, ctx().stream, xs, p, rs, rprev,
                         vh, nshift, st);
    }
    hipLaunchKernelGGL(k_msx_commit, dim3(1), dim3(NR), 0, ctx().stream, st, NR);
  }
}

// One chunk of up to NR sources on one parity.  x64[r*nshift+j] are the FP64
// solution halves (zeroed here), b64[r] the source halves, resid_src[r] each
// source's squared relative target.
//
// switch_frac > 0: the chunk first iterates in FP64 (FP64 multi-RHS Dslash and
// vectors) until every source's |r| has fallen below switch_frac * |b|, then
// in FP32 with FP64 reliable updates.  The shifted systems have no residual of
// their own to correct, so their solutions keep the rounding of every step;
// steps taken while |r| ~ |b| dominate it, and running those in FP64 leaves
// the shifted solutions at ~eps32 * switch_frac instead of ~eps32.
// switch_frac = 0: FP32 from the start.
template <int NR>
static bool run_chunk(LinksT<double> *L, int parity, int nrhs, int nshift, const double *mass,
                      const double *resid_src, int max_cg, double *const *x64,
                      const double *const *b64, int K, double switch_frac, BatchStats &bs) {
  const int vh = ctx().g.vh;
  const size_t nv = 6 * (size_t)vh, nvr = nv * NR;
  const LinksT<float> &Lf = *L->sloppy;
  const int other = 1 - parity;
  // FP32 workspace: p, xs [nshift][vh][NR][6]; rs, Ap, tmp [vh][NR][6]; rprev
  // [MSX_MAXDEFER - 1][vh][NR][6]; r64, tmp64; states; pointers
  constexpr size_t NRP = MSX_MAXDEFER - 1;
  char *w = (char *)ws2_get((2 * (size_t)nshift + 3 + NRP) * nvr * sizeof(float) + 2 * nv * sizeof(double) +
                            (NR + 2) * sizeof(SolverState) + NR * nshift * sizeof(double *) + 4096);
  if (!w) return false;
  float *p = (float *)w, *xs = p + nshift * nvr, *rs = xs + nshift * nvr, *Ap = rs + nvr, *tmp = Ap + nvr;
  float *rprev = tmp + nvr;
  double *r64 = (double *)(rprev + NRP * nvr), *tmp64 = r64 + nv;
  SolverState *st = (SolverState *)(((uintptr_t)(tmp64 + nv) + 255) & ~(uintptr_t)255);
  SolverState *tst = st + NR;
  double **dx = (double **)(tst + 1);
  // FP64-phase workspace, same shapes
  bool hi = switch_frac > 0;
  double *w64 = nullptr;
  if (hi) {
    w64 = (double *)dev_alloc((2 * (size_t)nshift + 3 + NRP) * nvr * sizeof(double));
    if (!w64) return false;
  }
  // Optional workspace for the FP64 true residuals of several paused sources
  // at once (their x_low packed, D x and D D x by the multi-RHS Dslash, then
  // unpacked): packed in, packed out, gate states, result states.  Without it
  // each source takes its own pair of Dslashes.
  struct TrBuf {
    void *p = nullptr;
    ~TrBuf() { dev_free(p); }
  } trbuf;
  const size_t trb = nrhs <= 2 ? 2 : (nrhs <= 4 ? 4 : 8);  // packed block of the checked sources
  trbuf.p = nrhs > 1 ? dev_try_alloc(2 * trb * nv * sizeof(double) + 16 * sizeof(SolverState) + 512) : nullptr;
  double *tr_in = (double *)trbuf.p, *tr_out = tr_in ? tr_in + trb * nv : nullptr;
  SolverState *tr_gate =
      tr_in ? (SolverState *)(((uintptr_t)(tr_out + trb * nv) + 255) & ~(uintptr_t)255) : nullptr;
  SolverState *tr_st = tr_gate ? tr_gate + 8 : nullptr;
  double *p_d = w64, *xs_d = hi ? p_d + nshift * nvr : nullptr;
  double *rs_d = hi ? xs_d + nshift * nvr : nullptr, *Ap_d = hi ? rs_d + nvr : nullptr;
  double *tmp_d = hi ? Ap_d + nvr : nullptr, *rprev_d = hi ? tmp_d + nvr : nullptr;
  const int nb = (vh + BLAS_BLOCK - 1) / BLAS_BLOCK;
  const int nlog = (int)(((size_t)vh * NR + BLAS_BLOCK - 1) / BLAS_BLOCK);
  const int K64 = std::min(K, 2);  // FP64: twice the registers per source
  const int nbd = mrhs_blocks<NR>(vh, K, Lf.recon, EPI_PKP);
  const int nbr = resident_grid(k_msx_update_r<float, NR>, nlog);
  const int nbd64 = mrhs_blocks<NR, double>(vh, K64, L->recon, EPI_PKP);
  const int nbr64 = resident_grid(k_msx_update_r<double, NR>, nlog);
  if ((size_t)NR * std::max(std::max(nbd, nbr), std::max(nbd64, nbr64)) > ctx().partials_len) {
    dev_free(w64);
    fail(__FILE__, __LINE__, "multi-RHS solver: reduction scratch too small for this block size");
    return false;
  }

  std::vector<SolverState> hs(NR);
  int j_low = 0;
  double offset_low = 1e300;
  std::vector<double> sh(nshift);
  for (int j = 0; j < nshift; ++j) {
    sh[j] = 4.0 * mass[j] * mass[j];
    if (sh[j] < offset_low) { offset_low = sh[j]; j_low = j; }
  }
  const char *de = std::getenv("QPHIX_MIXED_DELTA");
  const double delta = de ? std::atof(de) : 0.1;
  std::vector<char> done(NR, 1);
  for (int r = 0; r < NR; ++r) {
    bs.ms_iters[r] = 0;
    bs.updates[r] = 0;
    bs.final_rsq[r] = 0;
    for (int j = 0; j < nshift; ++j) bs.shift_iters[r][j] = 0;
  }
  // source norms (FP64)
  for (int r = 0; r < nrhs; ++r) {
    SolverState t{};
    std::memcpy(ctx().pinned, &t, sizeof(t));
    HIP_CHECK(hipMemcpyAsync(tst, ctx().pinned, sizeof(t), hipMemcpyHostToDevice, ctx().stream));
    blas_norm2<double>(b64[r], tst, SOP_SRC_NORM);
    HIP_CHECK(hipMemcpyAsync(ctx().pinned, tst, sizeof(SolverState), hipMemcpyDeviceToHost, ctx().stream));
    HIP_CHECK(hipStreamSynchronize(ctx().stream));
    std::memcpy(&t, ctx().pinned, sizeof(t));
    SolverState &s = hs[r];
    s = SolverState{};
    s.src_norm = t.src_norm;
    for (int j = 0; j < nshift; ++j)
      HIP_CHECK(hipMemsetAsync(x64[r * nshift + j], 0, nv * sizeof(double), ctx().stream));
    if (t.src_norm == 0.0) continue;
    done[r] = 0;
    s.single_coef = 0;
    s.nshift = nshift;
    s.num_now = nshift;
    s.j_low = j_low;
    s.offset_low = offset_low;
    for (int j = 0; j < nshift; ++j) {
      s.shifts[j] = j == j_low ? sh[j] : sh[j] - sh[j_low];
      s.zeta_im1[j] = s.zeta_i[j] = 1.0;
      s.beta_im1[j] = -1.0;
      s.alpha[j] = 0.0;
    }
    s.rsq = t.src_norm;
    s.rsqstop = resid_src[r] * t.src_norm;
    s.mx_maxr = std::sqrt(t.src_norm);
    s.mx_delta = hi ? switch_frac : delta;
    s.max_cg = max_cg;
    s.active = 1;
    s.fuse = nshift > 1 ? g_msx_fuse : 1;  // one shift: nothing to defer
    if (hi)
      hipLaunchKernelGGL(k_msx_init<double>, dim3(nb), dim3(BLAS_BLOCK), 0, ctx().stream, rs_d, p_d, xs_d,
                         b64[r], vh, NR, r, nshift);
    else
      hipLaunchKernelGGL(k_msx_init<float>, dim3(nb), dim3(BLAS_BLOCK), 0, ctx().stream, rs, p, xs, b64[r],
                         vh, NR, r, nshift);
  }
  // states and pointer table staged at distinct offsets of the pinned buffer
  // (both copies are in flight together)
  char *pin = (char *)ctx().pinned;
  const size_t ptr_off = PINNED_BYTES / 2;
  std::memcpy(pin, hs.data(), NR * sizeof(SolverState));
  std::memcpy(pin + ptr_off, x64, NR * nshift * sizeof(double *));
  HIP_CHECK(hipMemcpyAsync(st, pin, NR * sizeof(SolverState), hipMemcpyHostToDevice, ctx().stream));
  HIP_CHECK(hipMemcpyAsync(dx, pin + ptr_off, (size_t)NR * nshift * sizeof(double *), hipMemcpyHostToDevice, ctx().stream));
  HIP_CHECK(hipStreamSynchronize(ctx().stream));

  const int batch = 16;
  auto enqueue_batch = [&]() {
    if (hi)
      msx_batch<NR, double>(*L, parity, batch, p_d, xs_d, rs_d, Ap_d, tmp_d, rprev_d, j_low, offset_low, st,
                            K64, nbd64, nbr64, nlog, nshift);
    else
      msx_batch<NR, float>(Lf, parity, batch, p, xs, rs, Ap, tmp, rprev, j_low, offset_low, st, K, nbd, nbr,
                           nlog, nshift);
  };
  // grid queries (occupancy calculator) happen before any capture
  (void)mrhs_blocks<NR>(vh, K, Lf.recon, EPI_PLAIN);
  (void)mrhs_blocks<NR, double>(vh, K64, L->recon, EPI_PLAIN);
  BatchGraph steps;  // re-captured when the chunk drops to FP32

  // FP64 true residual b - (M + sigma_low) x_low of source r into r64; returns
  // its SolverState (rsq, final_rsq = rsq / |b|^2).
  auto true_residual = [&](int r) {
    SolverState t{};
    t.src_norm = hs[r].src_norm;
    std::memcpy(ctx().pinned, &t, sizeof(t));
    HIP_CHECK(hipMemcpyAsync(tst, ctx().pinned, sizeof(t), hipMemcpyHostToDevice, ctx().stream));
    double *xl = x64[r * nshift + j_low];
    DslashExtra e1;
    launch_dslash<double>(*L, other, tmp64, xl, EPI_PLAIN, e1);
    DslashExtra e2;
    e2.s = -offset_low;
    e2.aux = xl;
    e2.b = b64[r];
    e2.r = r64;
    e2.p = r64;
    e2.state = tst;
    e2.scalar_op = SOP_MX_TRUE;
    launch_dslash<double>(*L, parity, r64, tmp64, EPI_RESTART, e2);
    HIP_CHECK(hipMemcpyAsync(ctx().pinned, tst, sizeof(SolverState), hipMemcpyDeviceToHost, ctx().stream));
    HIP_CHECK(hipStreamSynchronize(ctx().stream));
    std::memcpy(&t, ctx().pinned, sizeof(t));
    return t;
  };
  // Source r after a true-residual check (state t, and on the device its
  // residual rr and state td): finished, or resumed in FP32.
  auto settle = [&](int r, const SolverState &t, double new_delta, const double *rr, const SolverState *td) {
    bs.ms_iters[r] = hs[r].iteration;
    bs.final_rsq[r] = t.final_rsq;
    if (t.final_rsq < resid_src[r] || hs[r].iteration >= max_cg || !std::isfinite(t.rsq)) {
      done[r] = 1;
      for (int j = 0; j < nshift; ++j)
        bs.shift_iters[r][j] = hs[r].sdone[j] ? hs[r].sdone_at[j] : hs[r].iteration;
      return;
    }
    ++bs.updates[r];
    hipLaunchKernelGGL(k_msx_set_r, dim3(nb), dim3(BLAS_BLOCK), 0, ctx().stream, rs, rr, vh, NR, r);
    hipLaunchKernelGGL(k_msx_resume, dim3(1), dim3(1), 0, ctx().stream, st + r, td, new_delta);
  };
  // True-residual checks of the paused sources R (their x already flushed),
  // then settle.  Two or more share the multi-RHS Dslash: x_low packed, D x,
  // D D x, unpacked, and r = b + D D x - sigma_low x per source, one host read
  // for all.
  auto check = [&](const std::vector<int> &R, double new_delta) {
    const int n = (int)R.size();
    if (n < 2 || !tr_in) {
      for (int r : R) settle(r, true_residual(r), new_delta, r64, tst);
      return;
    }
    const double *src[8];
    char *pin = (char *)ctx().pinned;
    for (int k = 0; k < n; ++k) {
      src[k] = x64[R[k] * nshift + j_low];
      SolverState t{};
      t.src_norm = hs[R[k]].src_norm;
      std::memcpy(pin + k * sizeof(SolverState), &t, sizeof(t));
    }
    HIP_CHECK(hipMemcpyAsync(tr_st, pin, n * sizeof(SolverState), hipMemcpyHostToDevice, ctx().stream));
    const int nr = mrhs_dslash_packed<double>(*L, other, src, n, tr_in, tr_out, tr_gate);  // D x
    if (nr == 2) mrhs_dslash<2, double>(*L, parity, tr_in, tr_out, EPI_PLAIN, 0.0, nullptr, tr_gate, 1);
    else if (nr == 4) mrhs_dslash<4, double>(*L, parity, tr_in, tr_out, EPI_PLAIN, 0.0, nullptr, tr_gate, 1);
    else mrhs_dslash<8, double>(*L, parity, tr_in, tr_out, EPI_PLAIN, 0.0, nullptr, tr_gate, 1);
    mrhs_unpack<double>(tr_in, tr_out, n, nr);  // D D x of source k at tr_out + k nv
    for (int k = 0; k < n; ++k) {
      const ReduceCtl rc = reduce_ctl(tr_st + k, SOP_MX_TRUE);
      hipLaunchKernelGGL(k_msx_true_res, dim3(nb), dim3(BLAS_BLOCK), 0, ctx().stream, tr_out + k * nv, src[k],
                         b64[R[k]], -offset_low, vh, rc);
      reduce_finish(2, nb, rc, nullptr);
    }
    HIP_CHECK(hipMemcpyAsync(pin, tr_st, n * sizeof(SolverState), hipMemcpyDeviceToHost, ctx().stream));
    HIP_CHECK(hipStreamSynchronize(ctx().stream));
    std::vector<SolverState> ts(n);
    for (int k = 0; k < n; ++k) std::memcpy(&ts[k], pin + k * sizeof(SolverState), sizeof(SolverState));
    for (int k = 0; k < n; ++k) settle(R[k], ts[k], new_delta, tr_out + k * nv, tr_st + k);
  };

  while (true) {
    bool any = false;
    for (int r = 0; r < nrhs; ++r) any |= !done[r];
    if (!any) break;
    steps.run(enqueue_batch);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(ctx().pinned, st, NR * sizeof(SolverState), hipMemcpyDeviceToHost, ctx().stream));
    HIP_CHECK(hipStreamSynchronize(ctx().stream));
    std::memcpy(hs.data(), ctx().pinned, NR * sizeof(SolverState));
    // a paused or finished source never has a deferred update pending
    for (int r = 0; r < nrhs; ++r)
      if (!done[r] && !hs[r].active && hs[r].pend) {
        dev_free(w64);
        fail(__FILE__, __LINE__, "multi-RHS solver: a paused source has a pending shifted update");
        return false;
      }
    if (hi) {
      // FP64 phase: a source pauses once |r| < switch_frac |b| (or on
      // convergence or budget); the chunk drops to FP32 when all have.
      bool all = true;
      for (int r = 0; r < nrhs; ++r)
        if (!done[r] && hs[r].active) all = false;
      if (!all) continue;
      std::vector<int> R;
      for (int r = 0; r < nrhs; ++r) {
        if (done[r]) continue;
        hipLaunchKernelGGL(k_msx_flush<double>, dim3(nb), dim3(BLAS_BLOCK), 0, ctx().stream, dx + r * nshift,
                           xs_d, vh, NR, r, nshift);
        R.push_back(r);
      }
      check(R, delta);
      const size_t np = (size_t)nshift * nvr;  // reals
      const size_t nslot = np / 6;             // spinor slots (site x source, every shift)
      hipLaunchKernelGGL(k_msx_demote, dim3((unsigned)((nslot + BLAS_BLOCK - 1) / BLAS_BLOCK)), dim3(BLAS_BLOCK), 0,
                         ctx().stream, p, p_d, nslot);
      HIP_CHECK(hipMemsetAsync(xs, 0, np * sizeof(float), ctx().stream));
      HIP_CHECK(hipStreamSynchronize(ctx().stream));
      dev_free(w64);
      w64 = nullptr;
      hi = false;
      steps.reset();
      continue;
    }
    std::vector<int> R;
    for (int r = 0; r < nrhs; ++r) {
      if (done[r] || hs[r].active) continue;
      // reliable update of source r: flush x, FP64 true residual of the base system
      hipLaunchKernelGGL(k_msx_flush<float>, dim3(nb), dim3(BLAS_BLOCK), 0, ctx().stream, dx + r * nshift, xs,
                         vh, NR, r, nshift);
      R.push_back(r);
    }
    check(R, 0.0);
  }
  HIP_CHECK(hipStreamSynchronize(ctx().stream));
  dev_free(w64);
  return true;
}

// Autotuner: the multi-RHS Dslash is timed for every RHS block size NR in
// {1, 2, 4, 8} and sources-per-work-item K in {1, 2, 4} (K <= NR), once per
// (lattice, Naik storage), and cached; the block size with the lowest time per
// source and its best K drive the batched solver and the multi-vector Dslash.
// How much of the gathered neighbour data stays in L2 favours NR; the L1
// traffic of the links (once per K sources) against registers picks K.
// QPHIX_MRHS_K pins K.
struct MrhsTune {
  int nr = 0;            // best block size; 0 = tuning failed
  int k[9] = {};         // best K per block size
  double us[9][3] = {};  // microseconds per source [nr][K = 1, 2, 4]
};

static const MrhsTune &autotune_mrhs(const LinksT<float> &Lf) {
  static std::map<std::pair<long, int>, MrhsTune> cache;
  static MrhsTune failed_tune;
  const int vh = ctx().g.vh;
  const auto key = std::make_pair((long)vh, Lf.recon);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  const size_t nvr = 6 * (size_t)vh * 8;
  char *w = (char *)ws2_get(2 * nvr * sizeof(float) + 8 * sizeof(SolverState) + 1024);
  if (!w) return failed_tune;
  float *in = (float *)w, *out = in + nvr;
  SolverState *st = (SolverState *)(((uintptr_t)(out + nvr) + 255) & ~(uintptr_t)255);
  HIP_CHECK(hipMemsetAsync(in, 0, nvr * sizeof(float), ctx().stream));
  std::vector<SolverState> hs(8);
  for (auto &x : hs) x.active = 1;
  std::memcpy(ctx().pinned, hs.data(), 8 * sizeof(SolverState));
  HIP_CHECK(hipMemcpyAsync(st, ctx().pinned, 8 * sizeof(SolverState), hipMemcpyHostToDevice, ctx().stream));
  const char *pin_k = std::getenv("QPHIX_MRHS_K");
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  MrhsTune t;
  double best_t = 1e300;
  const int cand[4] = {1, 2, 4, 8};
  const int kc[3] = {1, 2, 4};
  for (int c = 0; c < 4; ++c) {
    const int nr = cand[c];
    double best_m = 1e300;
    for (int q = 0; q < 3; ++q) {
      const int K = kc[q];
      if (K > nr || (pin_k && std::atoi(pin_k) != K)) continue;
      auto launch = [&]() {
        if (nr == 1) mrhs_dslash<1, float>(Lf, 0, out, in, EPI_PLAIN, 0.f, nullptr, st, K);
        else if (nr == 2) mrhs_dslash<2, float>(Lf, 0, out, in, EPI_PLAIN, 0.f, nullptr, st, K);
        else if (nr == 4) mrhs_dslash<4, float>(Lf, 0, out, in, EPI_PLAIN, 0.f, nullptr, st, K);
        else mrhs_dslash<8, float>(Lf, 0, out, in, EPI_PLAIN, 0.f, nullptr, st, K);
      };
      launch();
      HIP_CHECK(hipEventRecord(e0, ctx().stream));
      for (int r = 0; r < 3; ++r) launch();
      HIP_CHECK(hipEventRecord(e1, ctx().stream));
      HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0;
      HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      const double per = 1e3 * ms / 3.0 / nr;
      t.us[nr][q] = per;
      if (per < best_m) { best_m = per; t.k[nr] = K; }
    }
    if (best_m < best_t) { best_t = best_m; t.nr = nr; }
  }
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  if (t.nr == 0) {  // QPHIX_MRHS_K excluded every candidate
    t.nr = 1;
    t.k[1] = 1;
  }
  return cache[key] = t;
}

__global__ void __launch_bounds__(BLAS_BLOCK) k_axpy64(double *x, const double *e, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += e[i];
}

template <int NR>
static int invert_multi_rhs(QPHIX_info_t *info, LinksT<double> *L, QPHIX_invert_arg_t *inv,
                            QPHIX_resid_arg_t **res, const double *mass, int nshift, int nrhs,
                            QPHIX_D3_ColorVector **out, QPHIX_D3_ColorVector **in, int K) {
  const double t0 = wall_now();
  const int vh = ctx().g.vh;
  const size_t nv = 6 * (size_t)vh;
  const int max_cg = inv->max * inv->nrestart;
  int todo[2], ntodo;
  if (inv->parity == QPHIX_EVENODD) { todo[0] = 0; todo[1] = 1; ntodo = 2; }
  else { todo[0] = inv->parity == QPHIX_EVEN ? 0 : 1; ntodo = 1; }
  auto target = [&](int r, int j) {
    const double t = res[(size_t)r * nshift + j]->resid;
    return t > 0 ? t : res[0]->resid;
  };
  int ret = 0;
  std::vector<int> iters((size_t)nrhs * nshift, 0), upd((size_t)nrhs * nshift, 0);
  std::vector<double> frsq((size_t)nrhs * nshift, 0);
  double flops = 0;
  info->status = QPHIX_SUCCESS;
  // refinement scratch: NR residual vectors, NR corrections, one tmp, one state
  double *rbuf = (double *)dev_alloc((2 * NR + 1) * nv * sizeof(double) + sizeof(SolverState) + 256);
  if (!rbuf) { info->status = QPHIX_MEM_ERROR; return 0; }
  double *ebuf = rbuf + NR * nv, *tmp64 = ebuf + NR * nv;
  SolverState *tst = (SolverState *)(((uintptr_t)(tmp64 + nv) + 255) & ~(uintptr_t)255);
  const int nb = (int)((nv + BLAS_BLOCK - 1) / BLAS_BLOCK);
  const char *vb = std::getenv("QPHIX_VERBOSE");
  const bool verbose = vb && std::atoi(vb) > 0;
  // |r| / |b| at which the multi-shift iteration drops from FP64 to FP32
  // (QPHIX_MSX_SWITCH; 0 = FP32 throughout).  Measured at 32^3x64, 8 sources
  // x 12 shifts: 3.41 / 2.98 / 2.91 / 2.92 / 2.89 s for 3e-3 / 1e-3 / 3e-4 /
  // 1e-4 / 1e-5 (a later switch leaves the shifted solutions less accurate and
  // the refinement longer); 48^3x96, 4 x 16 shifts: 7.97 s at 1e-3, 7.74 s at 1e-4.
  const char *sw = std::getenv("QPHIX_MSX_SWITCH");
  const double switch_frac = sw ? std::atof(sw) : 3e-4;
  for (int pass = 0; pass < ntodo; ++pass) {
    const int parity = todo[pass], other = 1 - parity;
    // 1. multi-shift on blocks of NR sources
    for (int r0 = 0; r0 < nrhs; r0 += NR) {
      const int nr = std::min(NR, nrhs - r0);
      std::vector<double *> x(NR * nshift, nullptr);
      std::vector<const double *> b(NR, nullptr);
      std::vector<double> rs(NR, res[0]->resid);
      for (int r = 0; r < nr; ++r) {
        b[r] = in[r0 + r]->half[parity];
        for (int j = 0; j < nshift; ++j) x[r * nshift + j] = out[(r0 + r) * nshift + j]->half[parity];
      }
      BatchStats bs;
      if (!run_chunk<NR>(L, parity, nr, nshift, mass, rs.data(), max_cg, x.data(), b.data(), K, switch_frac, bs)) {
        dev_free(rbuf);
        info->status = QPHIX_MEM_ERROR;
        return 0;
      }
      for (int r = 0; r < nr; ++r) {
        ret = std::max(ret, bs.ms_iters[r]);
        flops += (1205.0 + 15.0 * nshift) * bs.ms_iters[r] * 2.0 * vh;
        for (int j = 0; j < nshift; ++j) {
          iters[(size_t)(r0 + r) * nshift + j] += bs.shift_iters[r][j];
          upd[(size_t)(r0 + r) * nshift + j] += bs.updates[r];
        }
      }
    }
    const double t_ms = wall_now();
    if (verbose)
      std::fprintf(stderr, "[qphix] multi_rhs parity %d: multi-shift phase %.3f s (block %d, %d sources per work-item)\n",
                   parity, t_ms - t0, NR, K);
    // 2. per-shift FP64 check; sources that miss their target are refined
    //    together, NR at a time, by the same solver on the correction equation
    //    (M + sigma_j) e = b - (M + sigma_j) x_j, then x_j += e.
    for (int j = 0; j < nshift; ++j) {
      const double sig = 4.0 * mass[j] * mass[j];
      const double t_j = wall_now();
      int refine_steps = 0;
      std::vector<int> todo_r;
      size_t processed = 0;
      std::vector<double> bnorm(nrhs, 0), rnorm(nrhs, 0);
      for (int r = 0; r < nrhs; ++r) {
        double *xj = out[(size_t)r * nshift + j]->half[parity];
        const double *br = in[r]->half[parity];
        SolverState t{};
        std::memcpy(ctx().pinned, &t, sizeof(t));
        HIP_CHECK(hipMemcpyAsync(tst, ctx().pinned, sizeof(t), hipMemcpyHostToDevice, ctx().stream));
        blas_norm2<double>(br, tst, SOP_SRC_NORM);
        DslashExtra e1;
        launch_dslash<double>(*L, other, tmp64, xj, EPI_PLAIN, e1);
        DslashExtra e2;
        e2.s = -sig;
        e2.aux = xj;
        e2.b = br;
        e2.r = rbuf + (todo_r.size() - processed) * nv;  // next free slot of the block
        e2.p = e2.r;
        e2.state = tst;
        e2.scalar_op = SOP_MX_TRUE;
        launch_dslash<double>(*L, parity, (double *)e2.r, tmp64, EPI_RESTART, e2);
        HIP_CHECK(hipMemcpyAsync(ctx().pinned, tst, sizeof(SolverState), hipMemcpyDeviceToHost, ctx().stream));
        HIP_CHECK(hipStreamSynchronize(ctx().stream));
        std::memcpy(&t, ctx().pinned, sizeof(t));
        bnorm[r] = t.src_norm;
        rnorm[r] = t.rsq;
        frsq[(size_t)r * nshift + j] = t.src_norm > 0 ? t.rsq / t.src_norm : 0;
        if (t.src_norm > 0 && t.rsq >= target(r, j) * t.src_norm && std::isfinite(t.rsq))
          todo_r.push_back(r);
        // refine a full block (or the last partial one) as soon as it is ready
        const bool last = (r == nrhs - 1);
        const size_t pending = todo_r.size() - processed;
        if (pending == (size_t)NR || (last && pending > 0)) {
          const int k0 = (int)processed;
          const int nr = (int)pending;
          processed = todo_r.size();
          std::vector<double *> e(NR, nullptr);
          std::vector<const double *> src(NR, nullptr);
          std::vector<double> rt(NR, 1.0);
          for (int q = 0; q < nr; ++q) {
            const int rr = todo_r[k0 + q];
            src[q] = rbuf + q * nv;
            e[q] = ebuf + q * nv;
            rt[q] = target(rr, j) * bnorm[rr] / rnorm[rr];
          }
          BatchStats bs;
          if (!run_chunk<NR>(L, parity, nr, 1, &mass[j], rt.data(), max_cg, e.data(), src.data(), K, 0.0, bs)) {
            dev_free(rbuf);
            info->status = QPHIX_MEM_ERROR;
            return 0;
          }
          int chunk_steps = 0;
          for (int q = 0; q < nr; ++q) chunk_steps = std::max(chunk_steps, bs.ms_iters[q]);
          refine_steps += chunk_steps;
          for (int q = 0; q < nr; ++q) {
            const int rr = todo_r[k0 + q];
            double *xj = out[(size_t)rr * nshift + j]->half[parity];
            hipLaunchKernelGGL(k_axpy64, dim3(nb), dim3(BLAS_BLOCK), 0, ctx().stream, xj, e[q], (int)nv);
            iters[(size_t)rr * nshift + j] += bs.ms_iters[q];
            upd[(size_t)rr * nshift + j] += bs.updates[q];
            frsq[(size_t)rr * nshift + j] = bs.final_rsq[q] * rnorm[rr] / bnorm[rr];
            flops += 1187.0 * bs.ms_iters[q] * 2.0 * vh;
          }
          HIP_CHECK(hipStreamSynchronize(ctx().stream));
        }
      }
      if (verbose) {
        double worst = 0;
        for (int r = 0; r < nrhs; ++r) worst = std::max(worst, frsq[(size_t)r * nshift + j]);
        std::fprintf(stderr, "[qphix]   shift %2d sigma %.3e: %zu sources refined, %d steps, %.3f s, worst rsq %.2e\n",
                     j, sig, todo_r.size(), refine_steps, wall_now() - t_j, worst);
      }
    }
  }
  dev_free(rbuf);
  for (int r = 0; r < nrhs; ++r)
    for (int j = 0; j < nshift; ++j) {
      QPHIX_resid_arg_t *rj = res[(size_t)r * nshift + j];
      rj->final_rsq = frsq[(size_t)r * nshift + j];
      rj->size_r = rj->final_rsq;
      rj->final_rel = 0;
      rj->size_relr = 0;
      rj->final_iter = iters[(size_t)r * nshift + j];
      rj->final_restart = upd[(size_t)r * nshift + j];
      if (!(rj->final_rsq < target(r, j))) info->status = QPHIX_FAIL;
    }
  info->final_sec = wall_now() - t0;
  info->final_flop = flops;
  return ret;
}

