// gauge_force.hip -- Symanzik one-loop gauge force on the momentum.
//
// Replaces QPHIX_*3_symanzik_1loop_gauge_force (gauge_force_imp.cpp:136-199)
// and its two-phase scatter (update_h_u_gforce_imp, :221-369, whose codegen
// bodies gforce_imp_body / update_mg_body are absent).  Semantics:
//   S(U) = - sum over geometric loops P of c_P Re Tr U_P,
//   mom_mu(x) <- mom_mu(x) - eps * TA(W_mu(x)),
//   W_mu(x) = U_mu(x) * sum_{loops P through (x,mu)} c_P * R_P,
// where R_P is the rest of P after its first step +mu from x and
// TA(M) = (M - M^dag)/2 - Tr(M - M^dag)/6.  c = coeffs->{plaquette, rectangle,
// parallelogram}; adjoint_plaquette is ignored (as in the reference).  The
// gauge field is not updated (the reference's update_mg_body may also touch U;
// that cannot be determined without its absent body).
//
// MI355X form: gather, not scatter -- one work-item per link computes the 48
// staples through its link (6 plaquette, 18 rectangle, 24 parallelogram) by
// walking link paths from L2-resident neighbours, in FP64, and writes only its
// own momentum.  No atomics, no second pass.  The links are stored
// complex-major (gauge_elem) and a wavefront's lanes are consecutive sites of
// one direction, so every step of a walk is a coalesced load.
#include "blas.h"
#include "handles.h"

#include <chrono>
#include <type_traits>

namespace qphix {

struct GFGeo {
  int L[4];
  int nxh, vh;
};

template <typename T> struct GFArgs {
  const T *u[2];
  T *h[2];
  GFGeo g;
  double c[3];
  double eps;
};

struct DM3 {
  double re[9], im[9];
};

__device__ __forceinline__ void dm_mul(const DM3 &a, const DM3 &b, DM3 &r) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double sr = 0, si = 0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        sr = fma(a.re[3 * i + k], b.re[3 * k + j], sr);
        sr = fma(-a.im[3 * i + k], b.im[3 * k + j], sr);
        si = fma(a.re[3 * i + k], b.im[3 * k + j], si);
        si = fma(a.im[3 * i + k], b.re[3 * k + j], si);
      }
      r.re[3 * i + j] = sr;
      r.im[3 * i + j] = si;
    }
}

template <typename T>
__device__ __forceinline__ void load_link(const GFArgs<T> &a, const int *y, int d, bool dag, DM3 &m) {
  int c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int v = y[k];
    v = v < 0 ? v + a.g.L[k] : (v >= a.g.L[k] ? v - a.g.L[k] : v);
    c[k] = v;
  }
  const int p = (c[0] + c[1] + c[2] + c[3]) & 1;
  const int i = (c[0] >> 1) + a.g.nxh * (c[1] + a.g.L[1] * (c[2] + a.g.L[2] * c[3]));
  // complex-major storage: one 2 x T load per element, coalesced across lanes
  // that walk the same loop from consecutive sites
  using T2 = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
  const T2 *s = reinterpret_cast<const T2 *>(a.u[p]);
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      const int src = dag ? 3 * cc + r : 3 * r + cc;
      const T2 v = s[gauge_elem(d, src, i, a.g.vh)];
      m.re[3 * r + cc] = (double)v.x;
      m.im[3 * r + cc] = dag ? -(double)v.y : (double)v.y;
    }
}

// r <- r * m, one row at a time (one row of temporaries instead of a third
// matrix: the walk keeps two 3x3 FP64 matrices live, which sets occupancy).
__device__ __forceinline__ void dm_rmul(DM3 &r, const DM3 &m) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double tr[3], ti[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double sr = 0, si = 0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        sr = fma(r.re[3 * i + k], m.re[3 * k + j], sr);
        sr = fma(-r.im[3 * i + k], m.im[3 * k + j], sr);
        si = fma(r.re[3 * i + k], m.im[3 * k + j], si);
        si = fma(r.im[3 * i + k], m.re[3 * k + j], si);
      }
      tr[j] = sr;
      ti[j] = si;
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      r.re[3 * i + j] = tr[j];
      r.im[3 * i + j] = ti[j];
    }
  }
}

// Product of the links along `n` signed steps (+(d+1) forward, -(d+1)
// backward) from vertex y, accumulated into acc with weight c.
template <int N, typename T>
__device__ __forceinline__ void walk_acc(const GFArgs<T> &a, const int *y0, const int *steps, double c,
                                         DM3 &acc) {
  int y[4] = {y0[0], y0[1], y0[2], y0[3]};
  DM3 r, m;
#pragma unroll 1
  for (int k = 0; k < N; ++k) {
    const int s = steps[k], d = (s > 0 ? s : -s) - 1;
    if (s > 0) {
      load_link(a, y, d, false, m);
      y[d] += 1;
    } else {
      y[d] -= 1;
      load_link(a, y, d, true, m);
    }
    if (k == 0) r = m;
    else dm_rmul(r, m);
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    acc.re[e] = fma(c, r.re[e], acc.re[e]);
    acc.im[e] = fma(c, r.im[e], acc.im[e]);
  }
}

// Grid: x over sites of one parity (blocks renumbered so that XCD x sweeps
// the x-th contiguous slab, keeping the paths' link loads, up to 2 sites away,
// in that XCD's L2), y = direction mu, z = parity.  All lanes of a wavefront
// walk the same loops from consecutive sites, so every link load of a step is
// coalesced.
__device__ __forceinline__ int gf_block_slot() {
  const int nb = gridDim.x, b = blockIdx.x;
  return nb % 8 == 0 ? (b % 8) * (nb / 8) + b / 8 : b;
}

// Occupancy target (waves per SIMD) for the force kernel; unset, the compiler
// keeps every walk's registers (2 waves per SIMD).
#ifdef QPHIX_GF_WAVES_PER_EU
#define QPHIX_GF_OCC __attribute__((amdgpu_waves_per_eu(QPHIX_GF_WAVES_PER_EU)))
#else
#define QPHIX_GF_OCC
#endif

template <typename T> __global__ void __launch_bounds__(256) QPHIX_GF_OCC k_gauge_force(const GFArgs<T> a) {
  const int i = gf_block_slot() * blockDim.x + threadIdx.x;
  if (i >= a.g.vh) return;
  const int mu = blockIdx.y, cb = blockIdx.z;
  int x[4];
  {
    const int xh = i % a.g.nxh;
    int r = i / a.g.nxh;
    x[1] = r % a.g.L[1];
    r /= a.g.L[1];
    x[2] = r % a.g.L[2];
    x[3] = r / a.g.L[2];
    x[0] = 2 * xh + ((x[1] + x[2] + x[3] + cb) & 1);
  }
  int y[4] = {x[0], x[1], x[2], x[3]};
  y[mu] += 1;  // every rest path starts at x + mu
  DM3 S;
#pragma unroll
  for (int e = 0; e < 9; ++e) S.re[e] = S.im[e] = 0;
  const int m = mu + 1;
  // one loop walk live at a time: unrolling these would interleave walks and
  // overflow the register file
#pragma unroll 1
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == mu) continue;
#pragma unroll 1
    for (int sg = -1; sg <= 1; sg += 2) {
      const int n = sg * (nu + 1);
      if (a.c[0] != 0) {
        const int p[3] = {n, -m, -n};
        walk_acc<3>(a, y, p, a.c[0], S);
      }
      if (a.c[1] != 0) {
        const int ra[5] = {m, n, -m, -m, -n};
        const int rb[5] = {n, -m, -m, -n, m};
        const int rc[5] = {n, n, -m, -n, -n};
        walk_acc<5>(a, y, ra, a.c[1], S);
        walk_acc<5>(a, y, rb, a.c[1], S);
        walk_acc<5>(a, y, rc, a.c[1], S);
      }
    }
  }
  if (a.c[2] != 0) {
#pragma unroll 1
    for (int da = 0; da < 4; ++da) {
      if (da == mu) continue;
#pragma unroll 1
      for (int db = 0; db < 4; ++db) {
        if (db == mu || db == da) continue;
#pragma unroll 1
        for (int sa = -1; sa <= 1; sa += 2)
#pragma unroll 1
          for (int sb = -1; sb <= 1; sb += 2) {
            const int A = sa * (da + 1), B = sb * (db + 1);
            const int p[5] = {A, B, -m, -A, -B};
            walk_acc<5>(a, y, p, a.c[2], S);
          }
      }
    }
  }
  DM3 U, W;
  load_link(a, x, mu, false, U);
  dm_mul(U, S, W);
  // H <- H + i eps TA(W)  (momentum stored as H = -iA, 8 reals)
  const double imtr = (W.im[0] + W.im[4] + W.im[8]) / 3.0;
  T *h = a.h[cb] + 32 * (size_t)i + 8 * mu;
  const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int r = pr[q][0], c = pr[q][1];
    const double tre = 0.5 * (W.re[3 * r + c] - W.re[3 * c + r]);
    const double tim = 0.5 * (W.im[3 * r + c] + W.im[3 * c + r]);
    h[2 * q] = (T)((double)h[2 * q] - a.eps * tim);
    h[2 * q + 1] = (T)((double)h[2 * q + 1] + a.eps * tre);
  }
  h[6] = (T)((double)h[6] - a.eps * (W.im[0] - imtr));
  h[7] = (T)((double)h[7] - a.eps * (W.im[4] - imtr));
}

template <typename T>
static void gauge_force(QPHIX_info_t *info, GaugeT<T> *G, ForceT<T> *F, QPHIX_gauge_coeffs_t *c,
                        double eps) {
  using namespace std::chrono;
  const double t0 = duration<double>(steady_clock::now().time_since_epoch()).count();
  const Geometry &gg = ctx().g;
  GFArgs<T> a;
  a.u[0] = G->u[0];
  a.u[1] = G->u[1];
  a.h[0] = F->h[0];
  a.h[1] = F->h[1];
  a.g = GFGeo{{gg.nx, gg.ny, gg.nz, gg.nt}, gg.nxh, gg.vh};
  a.c[0] = c->plaquette;
  a.c[1] = c->rectangle;
  a.c[2] = c->parallelogram;
  a.eps = eps;
  hipLaunchKernelGGL(k_gauge_force<T>, dim3((gg.vh + 255) / 256, 4, 2), dim3(256), 0, ctx().stream, a);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(ctx().stream));
  info->status = QPHIX_SUCCESS;
  info->final_sec = duration<double>(steady_clock::now().time_since_epoch()).count() - t0;
  info->final_flop = 61092.0;  // the reference's constant, gauge_force_imp.cpp:197
}

} // namespace qphix

using namespace qphix;

extern "C" {

#define QPHIX_GF_API(P, T)                                                                       \
  void QPHIX_##P##3_symanzik_1loop_gauge_force(QPHIX_info_t *info, QPHIX_##P##3_GaugeField *G,   \
                                               QPHIX_##P##3_Force *F,                            \
                                               QPHIX_gauge_coeffs_t *coeffs, T eps) {            \
    QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called: gauge force");                \
    QPHIX_REQUIRE(info && G && F && coeffs, "symanzik_1loop_gauge_force: null argument");        \
    info->status = QPHIX_FAIL;                                                                   \
    QPHIX_REQUIRE(F->parity == G->parity,                                                        \
                  "symanzik_1loop_gauge_force: force and gauge parities differ");                \
    gauge_force<T>(info, G, F, coeffs, (double)eps);                                             \
  }

QPHIX_GF_API(D, double)
QPHIX_GF_API(F, float)

} // extern "C"
