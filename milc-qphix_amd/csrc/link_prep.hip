// link_prep.hip -- asqtad / HISQ fermion links from the thin gauge field, on
// the device.
//
// In the reference the smeared links arrive from MILC (asqtad_create_L_from_raw,
// qphix_d3.cpp:970-1085) with the staggered phases already folded in; its own
// rephase and long-link kernels are disabled (ks_dslash_complete_specialization
// .cpp:238-275, `#if 0`).  Here they are built from a GaugeField (SU(3) links,
// no phases) on the GPU:
//
//   fat_mu(x) = eta_mu(x) b_mu(x) [ c1 U_mu(x) + sum over staple paths ],
//   lng_mu(x) = eta_mu(x) b3_mu(x) cN U_mu(x) U_mu(x+mu) U_mu(x+2mu),
//
// with the paths of Orginos, Toussaint & Sugar (Phys. Rev. D 60 (1999) 054503)
// and Lepage (Phys. Rev. D 59 (1999) 074502): 6 three-link staples (weight c3
// each), 24 five-link and 48 seven-link staples (c5, c7 per ordered path) and
// 6 five-link Lepage paths (cL), written with the physical (unphased) signs:
// fat7 = (1/8, 1/16, 1/64, 1/384), asqtad adds cL = -1/16, cN = -1/24 and
// c1 = 5/8.  The staples are nested: with the staple map
//   S_{d,s}(M)(y) = U_d(y) M(y+d) U_d(y+mu)^dag          (s = +1)
//                 = U_d(y-d)^dag M(y-d) U_d(y-d+mu)       (s = -1),
//   fat = c1 U + sum_{d,s} S_{d,s}( c3 U + cL S_{d,s}(U)
//                  + sum_{e,s'} S_{e,s'}( c5 U + c7 sum_{f,s''} S_{f,s''}(U) ) ),
// d, e, f distinct directions other than mu: 168 SU(3) products per link.
// eta_mu(x) is MILC's staggered phase (C11: (-1)^t, (-1)^(t+x), (-1)^(t+x+y),
// 1); b_mu = -1 on t-links whose path crosses the last time slice when the
// time boundary is antiperiodic (the phases and the boundary commute with the
// smearing: every path from x to x+mu carries the same signs).  Back links are
// MILC's double store, slot 2mu at x = -fwd_mu(x - hop mu)^dag.
//
// HISQ (Follana et al., Phys. Rev. D 75 (2007) 054502): W = P_U(3)(fat7(U)),
// the unitary projection V (V^dag V)^(-1/2) (here the polar factor, by the
// scaled Newton iteration X <- (z X + (z X)^(-dag)) / 2, z = |det X|^(-1/3)
// (Higham, SIAM J. Sci. Stat. Comput. 7 (1986) 1160)); then the fat links are
// the asqtad-like smearing of W with c1 = 1, c3 = 1/16, c5 = 1/64, c7 = 1/384,
// cL = -1/8, and the Naik links -1/24 (1 + eps_N) W W W.
//
// The phased fat/long links and back links are written per parity in the raw
// MILC order into a staging buffer that the import path of
// asqtad_create_L_from_raw (fields.hip) consumes, so storage (12/18 Naik
// rows, tile-major pairs) is decided as for links handed over by MILC.
#include "blas.h"
#include "handles.h"
#include "qphix_mi355x.h"

#include <chrono>

namespace qphix {

struct LPGeo {
  int L[4];
  int vh;
};

struct CM3 {
  double re[9], im[9];
};

__device__ __forceinline__ void cm_zero(CM3 &a) {
#pragma unroll
  for (int e = 0; e < 9; ++e) a.re[e] = a.im[e] = 0;
}

// r = a * b^(dag if DB) with a^(dag if DA)
template <bool DA, bool DB> __device__ __forceinline__ void cm_mul(const CM3 &a, const CM3 &b, CM3 &r) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double sr = 0, si = 0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int ia = DA ? 3 * k + i : 3 * i + k, ib = DB ? 3 * j + k : 3 * k + j;
        const double ar = a.re[ia], ai = DA ? -a.im[ia] : a.im[ia];
        const double br = b.re[ib], bi = DB ? -b.im[ib] : b.im[ib];
        sr = fma(ar, br, sr);
        sr = fma(-ai, bi, sr);
        si = fma(ar, bi, si);
        si = fma(ai, br, si);
      }
      r.re[3 * i + j] = sr;
      r.im[3 * i + j] = si;
    }
}

__device__ __forceinline__ void cm_axpy(double c, const CM3 &a, CM3 &r) {
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    r.re[e] = fma(c, a.re[e], r.re[e]);
    r.im[e] = fma(c, a.im[e], r.im[e]);
  }
}

// Link U_d at coordinates c (wrapped) from a field in the gauge layout
// (gauge_elem: complex (d*9 + e)*vh + i of parity (c0+c1+c2+c3) & 1).
template <typename T>
__device__ __forceinline__ void lp_load(const T *u0, const T *u1, const LPGeo &g, const int c[4], int d, CM3 &m) {
  const int r = c[0] + g.L[0] * (c[1] + g.L[1] * (c[2] + g.L[2] * c[3]));
  const T *u = ((c[0] + c[1] + c[2] + c[3]) & 1) ? u1 : u0;
  const size_t i = (size_t)(r >> 1);
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    const size_t o = 2 * ((size_t)(d * 9 + e) * (size_t)g.vh + i);
    m.re[e] = (double)u[o];
    m.im[e] = (double)u[o + 1];
  }
}

__device__ __forceinline__ void lp_shift(const LPGeo &g, const int c[4], int d, int s, int out[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) out[k] = c[k];
  int v = out[d] + s;
  v = v < 0 ? v + g.L[d] : (v >= g.L[d] ? v - g.L[d] : v);
  out[d] = v;
}

// acc += c * S_{d,s}(M)(y), with M given at y + s d (Mdisp) for the link
// direction mu.
template <typename T>
__device__ __forceinline__ void staple_acc(const T *u0, const T *u1, const LPGeo &g, const int y[4], int mu, int d,
                                           int s, const CM3 &Mdisp, double c, CM3 &acc) {
  CM3 A, B, t, r;
  if (s > 0) {
    int ymu[4];
    lp_shift(g, y, mu, 1, ymu);
    lp_load(u0, u1, g, y, d, A);      // U_d(y)
    lp_load(u0, u1, g, ymu, d, B);    // U_d(y+mu), used daggered
    cm_mul<false, false>(A, Mdisp, t);
    cm_mul<false, true>(t, B, r);
  } else {
    int ym[4], ymmu[4];
    lp_shift(g, y, d, -1, ym);
    lp_shift(g, ym, mu, 1, ymmu);
    lp_load(u0, u1, g, ym, d, A);     // U_d(y-d), used daggered
    lp_load(u0, u1, g, ymmu, d, B);   // U_d(y-d+mu)
    cm_mul<true, false>(A, Mdisp, t);
    cm_mul<false, false>(t, B, r);
  }
  cm_axpy(c, r, acc);
}

struct SmearCoef {
  double c1, c3, c5, c7, cl, cn;
};
// HISQ's second level (Follana et al. 2007): one link 1/8 + 2 * 3/8 + 1/8,
// staples as fat7, Lepage twice the asqtad value, Naik -(1 + eps)/24.
static SmearCoef hisq_level2(double naik_eps) {
  return SmearCoef{1.0, 1.0 / 16, 1.0 / 64, 1.0 / 384, -1.0 / 8, -(1.0 + naik_eps) / 24.0};
}

// One work-item per link: out = smeared fat link (no phases) in the gauge
// layout; nout (optional) = cN U U U.
template <typename T>
__global__ void __launch_bounds__(128)
    k_smear(const T *u0, const T *u1, double *o0, double *o1, double *n0, double *n1, LPGeo g, SmearCoef k) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int cb = blockIdx.y;
  if (t >= 4 * g.vh) return;
  const int i = t >> 2, mu = t & 3;
  // coordinates of half-site i of parity cb (raw MILC index)
  const int nxh = g.L[0] / 2;
  int c[4];
  int r = i / nxh;
  const int xh = i - r * nxh;
  c[1] = r % g.L[1];
  r /= g.L[1];
  c[2] = r % g.L[2];
  c[3] = r / g.L[2];
  c[0] = 2 * xh + ((c[1] + c[2] + c[3] + cb) & 1);
  CM3 Umu, fat;
  lp_load(u0, u1, g, c, mu, Umu);
  cm_zero(fat);
  cm_axpy(k.c1, Umu, fat);
  for (int d = 0; d < 4; ++d) {
    if (d == mu) continue;
    for (int s = -1; s <= 1; s += 2) {
      int y[4];
      lp_shift(g, c, d, s, y);  // the displaced link's site
      CM3 in1, Uy;
      lp_load(u0, u1, g, y, mu, Uy);
      cm_zero(in1);
      cm_axpy(k.c3, Uy, in1);
      if (k.cl != 0) {  // Lepage: S_{d,s}(S_{d,s}(U)), inner staple at y
        int y2[4];
        lp_shift(g, y, d, s, y2);
        CM3 U2;
        lp_load(u0, u1, g, y2, mu, U2);
        staple_acc(u0, u1, g, y, mu, d, s, U2, k.cl, in1);
      }
      if (k.c5 != 0 || k.c7 != 0) {
        for (int e = 0; e < 4; ++e) {
          if (e == mu || e == d) continue;
          for (int s2 = -1; s2 <= 1; s2 += 2) {
            int z[4];
            lp_shift(g, y, e, s2, z);
            CM3 in2, Uz;
            lp_load(u0, u1, g, z, mu, Uz);
            cm_zero(in2);
            cm_axpy(k.c5, Uz, in2);
            if (k.c7 != 0) {
              for (int f = 0; f < 4; ++f) {
                if (f == mu || f == d || f == e) continue;
                for (int s3 = -1; s3 <= 1; s3 += 2) {
                  int w[4];
                  lp_shift(g, z, f, s3, w);
                  CM3 Uw;
                  lp_load(u0, u1, g, w, mu, Uw);
                  staple_acc(u0, u1, g, z, mu, f, s3, Uw, k.c7, in2);
                }
              }
            }
            staple_acc(u0, u1, g, y, mu, e, s2, in2, 1.0, in1);
          }
        }
      }
      staple_acc(u0, u1, g, c, mu, d, s, in1, 1.0, fat);
    }
  }
  double *o = cb ? o1 : o0;
  const size_t vh = (size_t)g.vh;
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    const size_t q = 2 * ((size_t)(mu * 9 + e) * vh + i);
    o[q] = fat.re[e];
    o[q + 1] = fat.im[e];
  }
  if (n0) {
    int c1[4], c2[4];
    lp_shift(g, c, mu, 1, c1);
    lp_shift(g, c1, mu, 1, c2);
    CM3 a, b, p, nk;
    lp_load(u0, u1, g, c1, mu, a);
    lp_load(u0, u1, g, c2, mu, b);
    cm_mul<false, false>(Umu, a, p);
    cm_mul<false, false>(p, b, nk);
    double *n = cb ? n1 : n0;
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      const size_t q = 2 * ((size_t)(mu * 9 + e) * vh + i);
      n[q] = k.cn * nk.re[e];
      n[q + 1] = k.cn * nk.im[e];
    }
  }
}

// 3x3 complex inverse by cofactors: r = a^-1.
__device__ __forceinline__ void cm_inv(const CM3 &a, CM3 &r) {
  // cofactor C_ij (complex); inverse = C^T / det
  double cr[9], ci[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      // a[i1][j1] a[i2][j2] - a[i1][j2] a[i2][j1]
      const double pr = a.re[3 * i1 + j1] * a.re[3 * i2 + j2] - a.im[3 * i1 + j1] * a.im[3 * i2 + j2];
      const double pi = a.re[3 * i1 + j1] * a.im[3 * i2 + j2] + a.im[3 * i1 + j1] * a.re[3 * i2 + j2];
      const double qr = a.re[3 * i1 + j2] * a.re[3 * i2 + j1] - a.im[3 * i1 + j2] * a.im[3 * i2 + j1];
      const double qi = a.re[3 * i1 + j2] * a.im[3 * i2 + j1] + a.im[3 * i1 + j2] * a.re[3 * i2 + j1];
      cr[3 * i + j] = pr - qr;
      ci[3 * i + j] = pi - qi;
    }
  double dr = 0, di = 0;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    dr += a.re[j] * cr[j] - a.im[j] * ci[j];
    di += a.re[j] * ci[j] + a.im[j] * cr[j];
  }
  const double n = dr * dr + di * di;
  const double ir = dr / n, ii = -di / n;  // 1 / det
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double xr = cr[3 * j + i], xi = ci[3 * j + i];
      r.re[3 * i + j] = xr * ir - xi * ii;
      r.im[3 * i + j] = xr * ii + xi * ir;
    }
}

__device__ __forceinline__ double cm_absdet(const CM3 &a) {
  double dr = 0, di = 0;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    const double pr = a.re[3 + j1] * a.re[6 + j2] - a.im[3 + j1] * a.im[6 + j2];
    const double pi = a.re[3 + j1] * a.im[6 + j2] + a.im[3 + j1] * a.re[6 + j2];
    const double qr = a.re[3 + j2] * a.re[6 + j1] - a.im[3 + j2] * a.im[6 + j1];
    const double qi = a.re[3 + j2] * a.im[6 + j1] + a.im[3 + j2] * a.re[6 + j1];
    const double cr = pr - qr, ci = pi - qi;
    dr += a.re[j] * cr - a.im[j] * ci;
    di += a.re[j] * ci + a.im[j] * cr;
  }
  return sqrt(dr * dr + di * di);
}

// Unitary polar factor of every link, in place (gauge layout, FP64): scaled
// Newton iterations until the update is below 1e-15 (at most 30).
__global__ void __launch_bounds__(256) k_u3_project(double *v0, double *v1, int vh) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int cb = blockIdx.y;
  if (t >= 4 * vh) return;
  const int i = t >> 2, mu = t & 3;
  double *v = cb ? v1 : v0;
  CM3 X;
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    const size_t q = 2 * ((size_t)(mu * 9 + e) * (size_t)vh + i);
    X.re[e] = v[q];
    X.im[e] = v[q + 1];
  }
  for (int it = 0; it < 30; ++it) {
    const double z = pow(cm_absdet(X), -1.0 / 3.0);
    CM3 Xi;
    cm_inv(X, Xi);
    double delta = 0;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        // (z X + (z X)^(-dag)) / 2; (X^-1)^dag element (r, c) = conj(Xi(c, r))
        const double nr = 0.5 * (z * X.re[3 * r + c] + Xi.re[3 * c + r] / z);
        const double ni = 0.5 * (z * X.im[3 * r + c] - Xi.im[3 * c + r] / z);
        delta = fmax(delta, fabs(nr - X.re[3 * r + c]) + fabs(ni - X.im[3 * r + c]));
        X.re[3 * r + c] = nr;
        X.im[3 * r + c] = ni;
      }
    if (delta < 1e-15) break;
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    const size_t q = 2 * ((size_t)(mu * 9 + e) * (size_t)vh + i);
    v[q] = X.re[e];
    v[q + 1] = X.im[e];
  }
}

// Raw MILC-order fat / long links and back links of output parity cb with the
// staggered phases and the time boundary folded in: stage + 0: fat, + n: fat
// back, + 2n: long, + 3n: long back (n = 72 vh reals, [i][mu][row][col][re,im]).
__device__ __forceinline__ double lp_sign(const LPGeo &g, const int c[4], int mu, int hop, int apbc) {
  double s = 1;
  if (mu == 0) s = (c[3] & 1) ? -1 : 1;
  else if (mu == 1) s = ((c[3] + c[0]) & 1) ? -1 : 1;
  else if (mu == 2) s = ((c[3] + c[0] + c[1]) & 1) ? -1 : 1;
  if (mu == 3 && apbc && c[3] + hop >= g.L[3]) s = -s;
  return s;
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_links_raw(const double *f0, const double *f1, const double *n0, const double *n1, T *stage, LPGeo g, int cb,
                int apbc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * g.vh) return;
  const int i = t >> 2, mu = t & 3;
  const int nxh = g.L[0] / 2;
  int c[4];
  int r = i / nxh;
  const int xh = i - r * nxh;
  c[1] = r % g.L[1];
  r /= g.L[1];
  c[2] = r % g.L[2];
  c[3] = r / g.L[2];
  c[0] = 2 * xh + ((c[1] + c[2] + c[3] + cb) & 1);
  const size_t n = 72 * (size_t)g.vh, vh = (size_t)g.vh;
  for (int w = 0; w < 2; ++w) {  // 0 fat (hop 1), 1 long (hop 3)
    const int hop = w ? 3 : 1;
    const double *s0 = w ? n0 : f0, *s1 = w ? n1 : f1;
    // forward link at x
    {
      const double *src = cb ? s1 : s0;
      const double sg = lp_sign(g, c, mu, hop, apbc);
      T *dst = stage + (size_t)(2 * w) * n + 72 * (size_t)i + 18 * mu;
#pragma unroll
      for (int e = 0; e < 9; ++e) {
        const size_t q = 2 * ((size_t)(mu * 9 + e) * vh + i);
        dst[2 * e] = (T)(sg * src[q]);
        dst[2 * e + 1] = (T)(sg * src[q + 1]);
      }
    }
    // back link: -fwd(x - hop mu)^dag
    {
      int b[4] = {c[0], c[1], c[2], c[3]};
      int v = b[mu] - hop;
      b[mu] = v < 0 ? v + g.L[mu] : v;
      const int pb = (b[0] + b[1] + b[2] + b[3]) & 1;
      const int rb = b[0] + g.L[0] * (b[1] + g.L[1] * (b[2] + g.L[2] * b[3]));
      const size_t ib = (size_t)(rb >> 1);
      const double *src = pb ? s1 : s0;
      const double sg = -lp_sign(g, b, mu, hop, apbc);
      T *dst = stage + (size_t)(2 * w + 1) * n + 72 * (size_t)i + 18 * mu;
#pragma unroll
      for (int rr = 0; rr < 3; ++rr)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
          const size_t q = 2 * ((size_t)(mu * 9 + 3 * cc + rr) * vh + ib);  // element (cc, rr) of fwd
          dst[2 * (3 * rr + cc)] = (T)(sg * src[q]);
          dst[2 * (3 * rr + cc) + 1] = (T)(-sg * src[q + 1]);
        }
    }
  }
}

static LPGeo lp_geo() {
  const Geometry &gg = ctx().g;
  return LPGeo{{gg.nx, gg.ny, gg.nz, gg.nt}, gg.vh};
}

// The smeared (unphased) fat and Naik fields of the thin links G, FP64, in
// the gauge layout: fat in buf[0..1], Naik in buf[2..3], each 72 vh doubles;
// HISQ also uses buf[4..5] for its intermediate W.  hisq = false: one level
// with coefficients c; true: HISQ with Naik factor (1 + naik_eps).
template <typename T>
static void smear_links(const GaugeT<T> &G, const SmearCoef &c, bool hisq, double naik_eps, double *const *buf) {
  const LPGeo g = lp_geo();
  const dim3 grid((unsigned)((4 * g.vh + 127) / 128), 2);
  if (!hisq) {
    hipLaunchKernelGGL(k_smear<T>, grid, dim3(128), 0, ctx().stream, G.u[0], G.u[1], buf[0], buf[1], buf[2], buf[3],
                       g, c);
    HIP_CHECK(hipGetLastError());
    return;
  }
  // level 1: V = fat7(U) into buf[4..5], then W = unitary polar factor of V
  const SmearCoef fat7{1.0 / 8, 1.0 / 16, 1.0 / 64, 1.0 / 384, 0.0, 0.0};
  hipLaunchKernelGGL(k_smear<T>, grid, dim3(128), 0, ctx().stream, G.u[0], G.u[1], buf[4], buf[5],
                     (double *)nullptr, (double *)nullptr, g, fat7);
  const dim3 pgrid((unsigned)((4 * g.vh + 255) / 256), 2);
  hipLaunchKernelGGL(k_u3_project, pgrid, dim3(256), 0, ctx().stream, buf[4], buf[5], g.vh);
  // level 2 from W: fat links and the Naik links -(1 + eps)/24 W W W
  hipLaunchKernelGGL(k_smear<double>, grid, dim3(128), 0, ctx().stream, (const double *)buf[4],
                     (const double *)buf[5], buf[0], buf[1], buf[2], buf[3], g, hisq_level2(naik_eps));
  HIP_CHECK(hipGetLastError());
}

} // namespace qphix

namespace qphix {

static double lp_now() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// flags bit 0: periodic time boundary (default antiperiodic).
template <typename T>
static typename Handles<T>::L *links_from_gauge(QPHIX_info_t *info, GaugeT<T> *G, const SmearCoef &c, bool hisq,
                                                double naik_eps, int flags) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called: links from gauge field", nullptr);
  QPHIX_REQUIRE(info && G, "links from gauge field: null argument", nullptr);
  info->status = QPHIX_FAIL;
  QPHIX_REQUIRE(G->u[0] && G->u[1], "links from gauge field: the gauge field holds no links", nullptr);
  const double t0 = lp_now();
  const LPGeo g = lp_geo();
  const size_t nh = 72 * (size_t)g.vh;  // doubles per parity of one field
  const int nbuf = hisq ? 6 : 4;
  double *buf = (double *)dev_alloc(sizeof(double) * nh * nbuf);
  if (!buf) {  // error recorded
    info->status = QPHIX_MEM_ERROR;
    return nullptr;
  }
  double *b[6];
  for (int k = 0; k < 6; ++k) b[k] = k < nbuf ? buf + k * nh : nullptr;
  smear_links<T>(*G, c, hisq, naik_eps, b);
  const int apbc = (flags & 1) ? 0 : 1;
  std::function<bool(int, T *)> fill = [&](int cb, T *stage) {
    hipLaunchKernelGGL(k_links_raw<T>, dim3((unsigned)((4 * g.vh + 255) / 256)), dim3(256), 0, ctx().stream, b[0],
                       b[1], b[2], b[3], stage, g, cb, apbc);
    HIP_CHECK(hipGetLastError());
    return true;
  };
  auto *L = create_L_device<T>(QPHIX_EVENODD, fill);
  HIP_CHECK(hipStreamSynchronize(ctx().stream));
  dev_free(buf);
  if (!L) {
    info->status = QPHIX_MEM_ERROR;
    return nullptr;
  }
  info->final_sec = lp_now() - t0;
  // per link: 168 staple products and 2 Naik products, 216 flops each (x 2 + the
  // projection for HISQ)
  info->final_flop = (hisq ? 2.0 : 1.0) * 170.0 * 216.0 * 8.0 * g.vh;
  info->status = QPHIX_SUCCESS;
  return L;
}

} // namespace qphix

using namespace qphix;

extern "C" {

#define QPHIX_LP_API(P, T)                                                                                   \
  QPHIX_##P##3_FermionLinksAsqtad *QPHIX_##P##3_asqtad_create_L_from_G(                                      \
      QPHIX_info_t *info, const QPHIX_MI355X_asqtad_coeffs_t *coeffs, QPHIX_##P##3_GaugeField *G, int flags) { \
    QPHIX_REQUIRE(coeffs, "asqtad_create_L_from_G: null coefficients", nullptr);                             \
    const SmearCoef c{coeffs->one_link,     coeffs->three_staple, coeffs->five_staple,                      \
                      coeffs->seven_staple, coeffs->lepage,       coeffs->naik};                            \
    return static_cast<QPHIX_##P##3_FermionLinksAsqtad *>(links_from_gauge<T>(info, G, c, false, 0.0, flags)); \
  }                                                                                                          \
  QPHIX_##P##3_FermionLinksAsqtad *QPHIX_##P##3_hisq_create_L_from_G(QPHIX_info_t *info,                      \
                                                                     QPHIX_##P##3_GaugeField *G,              \
                                                                     double naik_eps, int flags) {            \
    return static_cast<QPHIX_##P##3_FermionLinksAsqtad *>(                                                   \
        links_from_gauge<T>(info, G, hisq_level2(naik_eps), true, naik_eps, flags));                         \
  }

QPHIX_LP_API(D, double)
QPHIX_LP_API(F, float)

} // extern "C"
