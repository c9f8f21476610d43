// gauge_update.hip -- the gauge-field half of a molecular-dynamics step, and
// pure-gauge MD trajectories built from it and the Symanzik force.
//
// The reference's update_h_u_gforce_imp (gauge_force_imp.cpp:221-369) updates
// the momentum and then the gauge field ("Update Momentum / Then update Gauge
// field", :215-217), calling update_mg_body with epsilonU = eps
// (gauge_force_imp.cpp:185, :334); that body is in the absent codegen
// repository.  Its semantics are MILC's update_u: every link
//     U_mu(x) <- exp(eps A_mu(x)) U_mu(x),
// A the traceless anti-Hermitian momentum held by the Force field (H = -iA, 8
// reals per link), which keeps U in SU(3) and with the force step
// A <- A - eps TA(W) (gauge_force.hip) conserves H = -1/2 sum Tr A^2 + S(U).
//
// The exponential: Q = -i eps A is Hermitian and traceless, so by the
// Cayley-Hamilton theorem Q^3 = c1 Q + c0 (c0 = det Q, c1 = Tr Q^2 / 2) and
// exp(iQ) = f0 + f1 Q + f2 Q^2 (Morningstar & Peardon, Phys. Rev. D 69 (2004)
// 054501, sec. III).  f0..f2 are summed here as the Taylor series of exp(iQ')
// for Q' = Q / 2^s with its eigenvalues below 0.5 in magnitude (21 terms,
// truncation < 1e-25), and squared s times in that basis (scaling and
// squaring, Moler & Van Loan, SIAM Rev. 45 (2003) 3) -- scalar arithmetic
// on three complex coefficients; only Q^2 and the final products are 3x3.
//
// One work-item per link, both parities in one launch: reads U (all rows) and
// the 8-real momentum, writes U and, when the field keeps one, its two-row
// copy (rows 0-1 of the new U).  HBM-bound: 144 + 64 B read, 144 (+ 96) B
// written per FP64 link.
#include "blas.h"
#include "handles.h"
#include "qphix_mi355x.h"

#include <chrono>
#include <cmath>

namespace qphix {

struct ZC {
  double re, im;
};
__device__ __forceinline__ ZC zmul(ZC a, ZC b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ ZC zadd(ZC a, ZC b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ ZC zscale(ZC a, double s) { return {a.re * s, a.im * s}; }

// exp(iQ) = f[0] + f[1] Q + f[2] Q^2 for Hermitian traceless Q with
// c0 = det Q, c1 = Tr Q^2 / 2.
__device__ inline void expiq_coeffs(double c0, double c1, ZC f[3]) {
  // eigenvalues |lambda| <= sqrt(4 c1 / 3); scale until that is <= 0.5
  int s = 0;
  while (c1 > 0.1875 && s < 60) {
    c1 *= 0.25;
    c0 *= 0.125;
    ++s;
  }
  // sum_n (i^n / n!) Q^n with Q^n = a I + b Q + d Q^2 (a, b, d real):
  // Q^{n+1} = d c0 I + (a + d c1) Q + b Q^2
  double a = 1, b = 0, d = 0, coef = 1;  // coef = 1 / n!
  f[0] = {0, 0};
  f[1] = {0, 0};
  f[2] = {0, 0};
  for (int n = 0; n <= 20; ++n) {
    // i^n: 1, i, -1, -i
    const int q = n & 3;
    const double re = q == 0 ? coef : (q == 2 ? -coef : 0.0);
    const double im = q == 1 ? coef : (q == 3 ? -coef : 0.0);
    f[0].re += re * a; f[0].im += im * a;
    f[1].re += re * b; f[1].im += im * b;
    f[2].re += re * d; f[2].im += im * d;
    const double na = d * c0, nb = a + d * c1, nd = b;
    a = na;
    b = nb;
    d = nd;
    coef /= (double)(n + 1);
  }
  // square s times: (f0 + f1 Q + f2 Q^2)^2 reduced with Q^3 = c0 + c1 Q,
  // Q^4 = c0 Q + c1 Q^2; the basis is Q' = Q / 2^k at step k, so c0, c1 grow
  for (int k = 0; k < s; ++k) {
    const ZC f0 = f[0], f1 = f[1], f2 = f[2];
    const ZC f1f2 = zmul(f1, f2), f2f2 = zmul(f2, f2);
    f[0] = zadd(zmul(f0, f0), zscale(f1f2, 2 * c0));
    f[1] = zadd(zadd(zscale(zmul(f0, f1), 2.0), zscale(f1f2, 2 * c1)), zscale(f2f2, c0));
    f[2] = zadd(zadd(zscale(zmul(f0, f2), 2.0), zmul(f1, f1)), zscale(f2f2, c1));
    // the squared matrix in the basis of 2 Q': f1 -> f1 / 2, f2 -> f2 / 4
    f[1] = zscale(f[1], 0.5);
    f[2] = zscale(f[2], 0.25);
    c0 *= 8.0;
    c1 *= 4.0;
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_gauge_update(T *u0, T *u1, T *w0, T *w1, const T *h0, const T *h1, int vh, double eps) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;  // link index within a parity: 4 i + mu
  const int cb = blockIdx.y;
  if (t >= 4 * vh) return;
  const int i = t >> 2, mu = t & 3;
  T *u = cb ? u1 : u0;
  T *w = cb ? w1 : w0;
  const T *h = (cb ? h1 : h0) + 8 * (size_t)t;
  // Q = eps H (H = -iA Hermitian traceless, MILC's 8-real order)
  ZC Q[9];
  Q[1] = {eps * (double)h[0], eps * (double)h[1]};
  Q[2] = {eps * (double)h[2], eps * (double)h[3]};
  Q[5] = {eps * (double)h[4], eps * (double)h[5]};
  Q[3] = {Q[1].re, -Q[1].im};
  Q[6] = {Q[2].re, -Q[2].im};
  Q[7] = {Q[5].re, -Q[5].im};
  Q[0] = {eps * (double)h[6], 0};
  Q[4] = {eps * (double)h[7], 0};
  Q[8] = {-Q[0].re - Q[4].re, 0};
  ZC Q2[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ZC s = {0, 0};
#pragma unroll
      for (int k = 0; k < 3; ++k) s = zadd(s, zmul(Q[3 * r + k], Q[3 * k + c]));
      Q2[3 * r + c] = s;
    }
  const double c1 = 0.5 * (Q2[0].re + Q2[4].re + Q2[8].re);
  // det Q = Tr Q^3 / 3 (traceless)
  double tr3 = 0;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k) tr3 += zmul(Q2[3 * r + k], Q[3 * k + r]).re;
  const double c0 = tr3 / 3.0;
  ZC f[3];
  expiq_coeffs(c0, c1, f);
  ZC E[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    ZC v = zadd(zmul(f[1], Q[e]), zmul(f[2], Q2[e]));
    if (e % 4 == 0) v = zadd(v, f[0]);
    E[e] = v;
  }
  // U <- E U
  ZC U[9];
  const size_t sv = (size_t)vh;
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    const size_t o = 2 * ((size_t)(mu * 9 + e) * sv + i);
    U[e] = {(double)u[o], (double)u[o + 1]};
  }
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ZC s = {0, 0};
#pragma unroll
      for (int k = 0; k < 3; ++k) s = zadd(s, zmul(E[3 * r + k], U[3 * k + c]));
      const size_t o = 2 * ((size_t)(mu * 9 + 3 * r + c) * sv + i);
      u[o] = (T)s.re;
      u[o + 1] = (T)s.im;
      if (w && r < 2) {
        const size_t ow = 2 * ((size_t)(mu * 6 + 3 * r + c) * sv + i);
        w[ow] = (T)s.re;
        w[ow + 1] = (T)s.im;
      }
    }
}

template <typename T> static void launch_update(GaugeT<T> *G, ForceT<T> *F, double eps) {
  const int vh = ctx().g.vh;
  const bool rows2 = G->u12[0] && G->u12[1];
  const dim3 grid((unsigned)((4 * vh + 255) / 256), 2);
  hipLaunchKernelGGL(k_gauge_update<T>, grid, dim3(256), 0, ctx().stream, G->u[0], G->u[1],
                     rows2 ? G->u12[0] : nullptr, rows2 ? G->u12[1] : nullptr, F->h[0], F->h[1], vh, eps);
  HIP_CHECK(hipGetLastError());
}

static double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

template <typename T> static bool gu_check(QPHIX_info_t *info, GaugeT<T> *G, ForceT<T> *F, const char *name) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called: gauge update", false);
  QPHIX_REQUIRE(info && G && F, name, false);
  info->status = QPHIX_FAIL;
  QPHIX_REQUIRE(G->u[0] && G->u[1] && F->h[0] && F->h[1], name, false);
  return true;
}

template <typename T> static void update_gauge(QPHIX_info_t *info, GaugeT<T> *G, ForceT<T> *F, double eps) {
  if (!gu_check<T>(info, G, F, "update_gauge: null or incomplete field")) return;
  const double t0 = now_s();
  launch_update<T>(G, F, eps);
  HIP_CHECK(hipStreamSynchronize(ctx().stream));
  info->final_sec = now_s() - t0;
  // per link: Q^2 and Tr Q^3 (~250), the coefficient series (~300), E and E U (~430)
  info->final_flop = 980.0 * 4 * 2 * ctx().g.vh;
  info->status = QPHIX_SUCCESS;
}

// Pure-gauge molecular dynamics: nsteps steps of size eps with the Symanzik
// force of coeffs.  scheme 0: leapfrog (half force step, then alternately a
// gauge step eps and a force step eps, ending with a half force step);
// scheme 1: Omelyan's second-order minimum-norm scheme (Omelyan, Mryglod &
// Folk, Comput. Phys. Commun. 151 (2003) 272, lambda = 0.1931833275037836):
// F(l e) U(e/2) F((1 - 2l) e) U(e/2) F(l e) per step.  Consecutive force
// steps of adjacent MD steps are merged.
template <typename T>
static void gauge_md(QPHIX_info_t *info, GaugeT<T> *G, ForceT<T> *F, QPHIX_gauge_coeffs_t *c, double eps,
                     int nsteps, int scheme) {
  if (!gu_check<T>(info, G, F, "gauge_md: null or incomplete field")) return;
  QPHIX_REQUIRE(c, "gauge_md: null coefficients");
  QPHIX_REQUIRE(nsteps >= 0 && (scheme == 0 || scheme == 1), "gauge_md: bad step count or scheme");
  const double t0 = now_s();
  QPHIX_info_t fi{};
  bool ok = true;
  auto force = [&](double e) {
    if (!ok) return;
    gauge_force_apply<T>(&fi, G, F, c, e);
    ok = fi.status == QPHIX_SUCCESS;
  };
  if (scheme == 0) {
    for (int k = 0; k < nsteps; ++k) {
      force(k == 0 ? 0.5 * eps : eps);
      launch_update<T>(G, F, eps);
    }
    if (nsteps > 0) force(0.5 * eps);
  } else {
    const double lam = 0.1931833275037836;
    for (int k = 0; k < nsteps; ++k) {
      force(k == 0 ? lam * eps : 2 * lam * eps);
      launch_update<T>(G, F, 0.5 * eps);
      force((1 - 2 * lam) * eps);
      launch_update<T>(G, F, 0.5 * eps);
    }
    if (nsteps > 0) force(lam * eps);
  }
  HIP_CHECK(hipStreamSynchronize(ctx().stream));
  if (!ok) return;  // the force recorded its error
  info->final_sec = now_s() - t0;
  info->final_flop = 0;
  info->status = QPHIX_SUCCESS;
}

} // namespace qphix

using namespace qphix;

extern "C" {

#define QPHIX_GU_API(P, T)                                                                             \
  void QPHIX_##P##3_update_gauge(QPHIX_info_t *info, QPHIX_##P##3_GaugeField *G, QPHIX_##P##3_Force *F,  \
                                 T eps) {                                                              \
    update_gauge<T>(info, G, F, (double)eps);                                                          \
  }                                                                                                    \
  void QPHIX_##P##3_gauge_md(QPHIX_info_t *info, QPHIX_##P##3_GaugeField *G, QPHIX_##P##3_Force *F,      \
                             QPHIX_gauge_coeffs_t *coeffs, T eps, int nsteps, int scheme) {            \
    gauge_md<T>(info, G, F, coeffs, (double)eps, nsteps, scheme);                                      \
  }

QPHIX_GU_API(D, double)
QPHIX_GU_API(F, float)

} // extern "C"
