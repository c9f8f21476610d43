// fields.hip -- create / extract / destroy for colour vectors, asqtad links,
// gauge fields and momenta; raw MILC order <-> device layouts.
//
// Reference: qphix_d3.cpp:79-174 (V), 177-343 (G), 347-488 (F), 490-559 (V out),
// 562-742 (G out), 745-949 (F out), 970-1206 (L), 1208-1245 (destroy).
// Raw arrays are even sites first, 6 (vectors) or 72 (links, momenta) reals per
// site (convention C1).  Link import keeps every slot exactly as the caller gave
// it (C4); long links are stored as 12 reals only when row 2 of every long link
// is +-conj(row0 x row1) to rounding, with the sign kept per link (1 bit),
// otherwise as 18 -- so the stored operator always equals the given one.
#include "handles.h"
#include "device.h"
#include "qphix_mi355x.h"

#include <cstring>
#include <vector>

namespace qphix {

static int g_recon_mode = 0;
int long_recon_mode() { return g_recon_mode; }

// ------------------------------------------------------------------ kernels
template <typename T>
__global__ void k_import_pairs(T *dst, const T *fwd, const T *back, int vh, int ncl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  constexpr int CPV = vec16<T>::ncplx;
  const int nc = 2 * ncl, nch = nc / CPV;
  for (int mu = 0; mu < 4; ++mu)
    for (int k = 0; k < nc; ++k) {
      const int kk = k < ncl ? k : k - ncl;
      const T *src = (k < ncl ? fwd : back) + 72 * (size_t)i + 18 * mu + 2 * kk;
      const size_t chunk = link_chunk(mu, k / CPV, i, nch);
      T *d = dst + chunk * (2 * CPV) + 2 * (k % CPV);
      d[0] = src[0];
      d[1] = src[1];
    }
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long *a, double v) {
  atomicMax(a, (unsigned long long)__double_as_longlong(v));
}

// Per long link: s = sign(Re<conj(r0 x r1), r2>), error |r2 - s conj(r0 x r1)|.
template <typename T>
__global__ void k_recon_check(const T *fwd, const T *back, int vh, uint8_t *sign,
                              unsigned long long *maxbits) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double worst = 0;
  if (i < vh) {
    unsigned bits = 0;
    for (int mu = 0; mu < 4; ++mu)
      for (int fb = 0; fb < 2; ++fb) {
        const T *u = (fb ? fwd : back) + 72 * (size_t)i + 18 * mu;
        double cr[3], ci[3], dot = 0;
        for (int c = 0; c < 3; ++c) {
          const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
          const double ar = u[2 * c1], ai = u[2 * c1 + 1], br = u[6 + 2 * c2], bi = u[6 + 2 * c2 + 1];
          const double er = u[2 * c2], ei = u[2 * c2 + 1], fr = u[6 + 2 * c1], fi = u[6 + 2 * c1 + 1];
          const double xr = ar * br - ai * bi - (er * fr - ei * fi);
          const double xi = ar * bi + ai * br - (er * fi + ei * fr);
          cr[c] = xr;
          ci[c] = -xi;
          dot += cr[c] * u[12 + 2 * c] + ci[c] * u[12 + 2 * c + 1];
        }
        const double s = dot < 0 ? -1.0 : 1.0;
        for (int c = 0; c < 3; ++c) {
          worst = fmax(worst, fabs(u[12 + 2 * c] - s * cr[c]));
          worst = fmax(worst, fabs(u[12 + 2 * c + 1] - s * ci[c]));
        }
        if (s < 0) bits |= 1u << (2 * mu + fb);
      }
    sign[i] = (uint8_t)bits;
  }
  for (int off = 32; off > 0; off >>= 1) worst = fmax(worst, __shfl_down(worst, off, 64));
  if ((threadIdx.x & 63) == 0) atomic_max_nonneg(maxbits, worst);
}

// Forward links back to raw order (extract_L); 12-stored long links get their
// third row rebuilt with the stored sign.
template <typename T>
__global__ void k_export_fwd(const T *fat, const T *lng, const uint8_t *sign, int vh, int ncl,
                             T *fat_raw, T *lng_raw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  constexpr int CPV = vec16<T>::ncplx;
  for (int mu = 0; mu < 4; ++mu) {
    for (int k = 0; k < 9; ++k) {
      const size_t chunk = link_chunk(mu, k / CPV, i, 18 / CPV);
      const T *s = fat + chunk * (2 * CPV) + 2 * (k % CPV);
      fat_raw[72 * (size_t)i + 18 * mu + 2 * k] = s[0];
      fat_raw[72 * (size_t)i + 18 * mu + 2 * k + 1] = s[1];
    }
    cplx<double> w[9];
    const int nch = 2 * ncl / CPV;
    for (int k = 0; k < ncl; ++k) {
      const size_t chunk = link_chunk(mu, k / CPV, i, nch);
      const T *s = lng + chunk * (2 * CPV) + 2 * (k % CPV);
      w[k].re = s[0];
      w[k].im = s[1];
    }
    if (ncl == 6) recon_row2(w, ((sign[i] >> (2 * mu + 1)) & 1u) ? -1.0 : 1.0);
    for (int k = 0; k < 9; ++k) {
      lng_raw[72 * (size_t)i + 18 * mu + 2 * k] = (T)w[k].re;
      lng_raw[72 * (size_t)i + 18 * mu + 2 * k + 1] = (T)w[k].im;
    }
  }
}

// Load complex k (row-major) of the forward (fb=1) or back (fb=0) link of
// direction mu at site i.
template <typename T>
__device__ __forceinline__ cplx<double> link_elem(const T *base, int vh, int ncl, int mu, int fb,
                                                  int k, int i) {
  constexpr int CPV = vec16<T>::ncplx;
  const int kk = fb ? k : ncl + k;
  const size_t chunk = link_chunk(mu, kk / CPV, i, 2 * ncl / CPV);
  const T *s = base + chunk * (2 * CPV) + 2 * (kk % CPV);
  return cplx<double>{(double)s[0], (double)s[1]};
}

template <typename T>
__device__ void load_link(const T *base, const uint8_t *sign, int vh, int ncl, int mu, int fb,
                          int i, cplx<double> *w) {
  for (int k = 0; k < ncl; ++k) w[k] = link_elem(base, vh, ncl, mu, fb, k, i);
  if (ncl == 6) recon_row2(w, ((sign[i] >> (2 * mu + fb)) & 1u) ? -1.0 : 1.0);
}

// max |back_mu(x) -/+ U_mu(x - hop)^dagger| for fat (hop 1) and long (hop 3).
template <typename T>
__global__ void k_backcheck(const T *fat_c, const T *lng_c, const uint8_t *sgn_c, const T *fat_o,
                            const T *lng_o, const uint8_t *sgn_o, int ncl, Geo g, int cb,
                            unsigned long long *out4) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double m[4] = {0, 0, 0, 0};
  if (i < g.vh) {
    SiteNbr n;
    site_neighbours(g, i, cb, n);
    for (int mu = 0; mu < 4; ++mu)
      for (int kind = 0; kind < 2; ++kind) {
        cplx<double> b[9], u[9];
        if (kind == 0) {
          load_link(fat_c, nullptr, g.vh, 9, mu, 0, i, b);
          load_link(fat_o, nullptr, g.vh, 9, mu, 1, n.b1[mu], u);
        } else {
          load_link(lng_c, sgn_c, g.vh, ncl, mu, 0, i, b);
          load_link(lng_o, sgn_o, g.vh, ncl, mu, 1, n.b3[mu], u);
        }
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            const cplx<double> a{u[3 * c + r].re, -u[3 * c + r].im};  // adj(u)[r][c]
            const cplx<double> x = b[3 * r + c];
            m[2 * kind] = fmax(m[2 * kind], fmax(fabs(x.re + a.re), fabs(x.im + a.im)));
            m[2 * kind + 1] = fmax(m[2 * kind + 1], fmax(fabs(x.re - a.re), fabs(x.im - a.im)));
          }
      }
  }
  for (int k = 0; k < 4; ++k) {
    double v = m[k];
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(out4 + k, v);
  }
}

// FP64 chunk layout -> FP32 chunk layout (same pairs, same recon).
__global__ void k_links_to_float(float *dst, const double *src, int vh, int nc_pair) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  for (int mu = 0; mu < 4; ++mu)
    for (int k = 0; k < nc_pair; ++k) {
      const double *s = src + link_chunk(mu, k, i, nc_pair) * 2;
      float *d = dst + link_chunk(mu, k / 2, i, nc_pair / 2) * 4 + 2 * (k % 2);
      d[0] = (float)s[0];
      d[1] = (float)s[1];
    }
}

// tan(arg z / 2) for the 8-real form (recon8), from the half-angle identities
// tan(t/2) = sin t / (1 + cos t) = (1 - cos t) / sin t, each used where it does
// not cancel; arg z = pi maps to 1e18 (phase -1 to FP32 rounding), z = 0 to 0.
__device__ double half_tan(double x, double y) {
  const double r = hypot(x, y);
  if (r == 0) return 0;
  const double h = x >= 0 ? y / (r + x) : (y == 0 ? 1e18 : (r - x) / y);
  return fmin(fmax(h, -1e18), 1e18);
}

// FP64 12-real long links -> FP32 8-real form: per link U01, U02, U10 and the
// half-angle tangents of U00 and U20 = s conj(U01 U12 - U02 U11) (s the stored
// row-2 sign).  The 8-real rebuild takes |U00| and |U20| from unitarity, so a
// link has no usable 8-real form when its rows 0 and 1 are not orthonormal to
// 1e-6 (the 12-real check only ties row 2 to them) or n = |U01|^2 + |U02|^2 <
// 1e-10; such links are counted in *bad and the caller keeps the 12-real form.
__global__ void k_links_to_float8(float *dst, const double *src, const uint8_t *sign, int vh,
                                  unsigned *bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  unsigned nbad = 0;
  for (int mu = 0; mu < 4; ++mu)
    for (int fb = 0; fb < 2; ++fb) {  // fb 0: forward link (chunks 0-5), 1: back link (6-11)
      double u[12];
      for (int k = 0; k < 6; ++k) {
        const double *c = src + link_chunk(mu, 6 * fb + k, i, 12) * 2;
        u[2 * k] = c[0];
        u[2 * k + 1] = c[1];
      }
      const double s = ((sign[i] >> (2 * mu + (fb ? 0 : 1))) & 1u) ? -1.0 : 1.0;
      // U01 U12 - U02 U11
      const double xr = u[2] * u[10] - u[3] * u[11] - (u[4] * u[8] - u[5] * u[9]);
      const double xi = u[2] * u[11] + u[3] * u[10] - (u[4] * u[9] + u[5] * u[8]);
      const double n = u[2] * u[2] + u[3] * u[3] + u[4] * u[4] + u[5] * u[5];
      const double n0 = u[0] * u[0] + u[1] * u[1] + n;
      const double n1 = u[6] * u[6] + u[7] * u[7] + u[8] * u[8] + u[9] * u[9] + u[10] * u[10] + u[11] * u[11];
      double dr = 0, di = 0;  // row 0 . conj(row 1)
      for (int c = 0; c < 3; ++c) {
        dr += u[2 * c] * u[6 + 2 * c] + u[2 * c + 1] * u[6 + 2 * c + 1];
        di += u[2 * c + 1] * u[6 + 2 * c] - u[2 * c] * u[6 + 2 * c + 1];
      }
      if (!(n >= 1e-10) || !(fabs(n0 - 1) <= 1e-6) || !(fabs(n1 - 1) <= 1e-6) || !(hypot(dr, di) <= 1e-6)) ++nbad;
      const float v[8] = {(float)u[2], (float)u[3], (float)u[4], (float)u[5], (float)u[6], (float)u[7],
                          (float)half_tan(u[0], u[1]), (float)half_tan(s * xr, -s * xi)};
      for (int k = 0; k < 4; ++k) {
        float *d = dst + link_chunk(mu, (4 * fb + k) / 2, i, 4) * 4 + 2 * (k % 2);
        d[0] = v[2 * k];
        d[1] = v[2 * k + 1];
      }
    }
  if (nbad) atomicAdd(bad, nbad);
}

// FP64 links (double2 chunks) -> 16-bit fixed point (Q16 chunks), per site
// and direction one pair of nc_pair complex elements (forward link, then back
// link), each link with its own scale.
__global__ void k_links_to_q16(uint4 *dst, const double *src, int vh, int nc_pair) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  const int nch = nc_pair == 12 ? link_nch<Q16>(12) : link_nch<Q16>(18);
  for (int mu = 0; mu < 4; ++mu) {
    double x[36];
    for (int k = 0; k < nc_pair; ++k) {
      const double *c = src + link_chunk(mu, k, i, nc_pair) * 2;
      x[2 * k] = c[0];
      x[2 * k + 1] = c[1];
    }
    float sc[2];
    for (int h = 0; h < 2; ++h) {
      double m = 0;
      for (int e = h * nc_pair; e < (h + 1) * nc_pair; ++e) m = fmax(m, fabs(x[e]));
      float f = (float)m;
      if ((double)f < m) f = nextafterf(f, INFINITY);  // |x| / f <= 1
      sc[h] = f;
    }
    unsigned w[20] = {};
    for (int e = 0; e < 2 * nc_pair; ++e) {
      const float f = sc[e < nc_pair ? 0 : 1];
      int q = f > 0.f ? (int)rint(x[e] / (double)f * 32767.0) : 0;
      q = q > 32767 ? 32767 : (q < -32767 ? -32767 : q);
      w[e / 2] |= ((unsigned)q & 0xffffu) << (16 * (e % 2));
    }
    w[4 * nch - 2] = __float_as_uint(sc[0]);
    w[4 * nch - 1] = __float_as_uint(sc[1]);
    for (int k = 0; k < nch; ++k)
      dst[link_chunk(mu, k, i, nch)] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
  }
}

// Momentum raw (anti-Hermitian A, 18 reals) -> 8 reals of H = -iA
// (qphix_d3.cpp:420-433): [ReH01, ImH01, ReH02, ImH02, ReH12, ImH12, H00, H11].
template <typename T> __global__ void k_mom_in(T *h, const T *raw, int nlinks) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlinks) return;
  const T *a = raw + 18 * (size_t)l;
  T *o = h + 8 * (size_t)l;
  // H_rc = -i A_rc : Re H = Im A, Im H = -Re A
  o[0] = a[2 * 1 + 1];   o[1] = -a[2 * 1];        // (0,1)
  o[2] = a[2 * 2 + 1];   o[3] = -a[2 * 2];        // (0,2)
  o[4] = a[6 + 2 * 2 + 1]; o[5] = -a[6 + 2 * 2];  // (1,2)
  o[6] = a[1];           o[7] = a[6 + 2 * 1 + 1]; // ImA00, ImA11
}

// Inverse (qphix_d3.cpp:809-842): A = iH, Re A_kk = 0, Im A22 = -Im A00 - Im A11.
template <typename T> __global__ void k_mom_out(T *raw, const T *h, int nlinks) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlinks) return;
  const T *o = h + 8 * (size_t)l;
  T *a = raw + 18 * (size_t)l;
  const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int q = 0; q < 3; ++q) {
    const int r = pr[q][0], c = pr[q][1];
    const T hr = o[2 * q], hi = o[2 * q + 1];
    a[6 * r + 2 * c] = -hi;  // A_rc = i H_rc
    a[6 * r + 2 * c + 1] = hr;
    a[6 * c + 2 * r] = hi;   // A_cr = -conj(A_rc)
    a[6 * c + 2 * r + 1] = hr;
  }
  a[0] = 0; a[1] = o[6];
  a[8] = 0; a[9] = o[7];
  a[16] = 0; a[17] = -o[6] - o[7];
}

// ------------------------------------------------------------------ helpers
static inline int nblk(int n, int b = 256) { return (n + b - 1) / b; }

static inline void h2d(void *d, const void *h, size_t bytes) {
  HIP_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx().stream));
}
static inline void d2h(void *h, const void *d, size_t bytes) {
  HIP_CHECK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, ctx().stream));
}
static inline void sync() { HIP_CHECK(hipStreamSynchronize(ctx().stream)); }

template <typename T> static double recon_tol() { return sizeof(T) == 8 ? 1e-12 : 4e-6; }

template <typename T> void links_free(LinksT<T> *L) {
  if (!L) return;
  for (int cb = 0; cb < 2; ++cb) {
    dev_free(L->fat[cb]);
    dev_free(L->lng[cb]);
    dev_free(L->lsign[cb]);
    L->fat[cb] = L->lng[cb] = nullptr;
    L->lsign[cb] = nullptr;
  }
  if (L->sloppy) {
    links_free(L->sloppy);
    delete L->sloppy;
    L->sloppy = nullptr;
  }
  if (L->sloppy16) {
    links_free(L->sloppy16);
    delete L->sloppy16;
    L->sloppy16 = nullptr;
  }
  if (L->sloppy8) {
    links_free(L->sloppy8);
    delete L->sloppy8;
    L->sloppy8 = nullptr;
  }
}
template void links_free<double>(LinksT<double> *);
template void links_free<float>(LinksT<float> *);
template void links_free<Q16>(LinksT<Q16> *);

bool links_to_float(const LinksT<double> &Ld, LinksT<float> &Lf) {
  const int vh = ctx().g.vh;
  Lf.recon = Ld.recon;
  Lf.parity = Ld.parity;
  for (int cb = 0; cb < 2; ++cb) {
    if (!Ld.fat[cb]) continue;
    Lf.fat[cb] = (float *)dev_alloc(sizeof(float) * 4 * 36 * link_sites(vh));
    Lf.lng[cb] = (float *)dev_alloc(sizeof(float) * 4 * 2 * Ld.recon * link_sites(vh));
    if (Ld.lsign[cb]) Lf.lsign[cb] = (uint8_t *)dev_alloc((size_t)vh);
    if (!Lf.fat[cb] || !Lf.lng[cb] || (Ld.lsign[cb] && !Lf.lsign[cb])) {  // error recorded
      HIP_CHECK(hipStreamSynchronize(ctx().stream));  // earlier parity's copies done
      links_free(&Lf);
      return false;
    }
    hipLaunchKernelGGL(k_links_to_float, dim3(nblk(vh)), dim3(256), 0, ctx().stream, Lf.fat[cb],
                       Ld.fat[cb], vh, 18);
    hipLaunchKernelGGL(k_links_to_float, dim3(nblk(vh)), dim3(256), 0, ctx().stream, Lf.lng[cb],
                       Ld.lng[cb], vh, Ld.recon == 12 ? 12 : 18);
    if (Ld.lsign[cb])
      HIP_CHECK(hipMemcpyAsync(Lf.lsign[cb], Ld.lsign[cb], (size_t)vh, hipMemcpyDeviceToDevice,
                               ctx().stream));
  }
  HIP_CHECK(hipGetLastError());
  return true;
}

// FP32 copy with 8-real long links (fat links as links_to_float).  Returns 0 on
// success, 1 when some long link has no 8-real form (nothing kept, no error),
// -1 when out of device memory (error recorded).
static int links_to_float8(const LinksT<double> &Ld, LinksT<float> &Lf) {
  const int vh = ctx().g.vh;
  Lf.recon = 8;
  Lf.parity = Ld.parity;
  unsigned *bad = (unsigned *)dev_alloc(sizeof(unsigned));
  if (!bad) return -1;
  HIP_CHECK(hipMemsetAsync(bad, 0, sizeof(unsigned), ctx().stream));
  for (int cb = 0; cb < 2; ++cb) {
    if (!Ld.fat[cb]) continue;
    Lf.fat[cb] = (float *)dev_alloc(sizeof(float) * 4 * 36 * link_sites(vh));
    Lf.lng[cb] = (float *)dev_alloc(sizeof(float) * 4 * 2 * 8 * link_sites(vh));
    Lf.lsign[cb] = (uint8_t *)dev_alloc((size_t)vh);
    if (!Lf.fat[cb] || !Lf.lng[cb] || !Lf.lsign[cb]) {  // error recorded
      HIP_CHECK(hipStreamSynchronize(ctx().stream));
      links_free(&Lf);
      dev_free(bad);
      return -1;
    }
    hipLaunchKernelGGL(k_links_to_float, dim3(nblk(vh)), dim3(256), 0, ctx().stream, Lf.fat[cb],
                       Ld.fat[cb], vh, 18);
    hipLaunchKernelGGL(k_links_to_float8, dim3(nblk(vh)), dim3(256), 0, ctx().stream, Lf.lng[cb],
                       Ld.lng[cb], Ld.lsign[cb], vh, bad);
    HIP_CHECK(hipMemcpyAsync(Lf.lsign[cb], Ld.lsign[cb], (size_t)vh, hipMemcpyDeviceToDevice, ctx().stream));
  }
  HIP_CHECK(hipGetLastError());
  unsigned nbad = 0;
  HIP_CHECK(hipMemcpyAsync(&nbad, bad, sizeof(unsigned), hipMemcpyDeviceToHost, ctx().stream));
  HIP_CHECK(hipStreamSynchronize(ctx().stream));
  dev_free(bad);
  if (nbad) {
    links_free(&Lf);
    return 1;
  }
  return 0;
}

bool links_to_q16(const LinksT<double> &Ld, LinksT<Q16> &Lq) {
  const int vh = ctx().g.vh;
  Lq.recon = Ld.recon;
  Lq.parity = Ld.parity;
  const int ncl = Ld.recon == 12 ? 12 : 18;
  for (int cb = 0; cb < 2; ++cb) {
    if (!Ld.fat[cb]) continue;
    Lq.fat[cb] = (Q16 *)dev_alloc(16 * 4 * (size_t)link_nch<Q16>(18) * link_sites(vh));
    Lq.lng[cb] = (Q16 *)dev_alloc(16 * 4 * (size_t)link_nch<Q16>(ncl) * link_sites(vh));
    if (Ld.lsign[cb]) Lq.lsign[cb] = (uint8_t *)dev_alloc((size_t)vh);
    if (!Lq.fat[cb] || !Lq.lng[cb] || (Ld.lsign[cb] && !Lq.lsign[cb])) {  // error recorded
      HIP_CHECK(hipStreamSynchronize(ctx().stream));  // earlier parity's copies done
      links_free(&Lq);
      return false;
    }
    hipLaunchKernelGGL(k_links_to_q16, dim3(nblk(vh)), dim3(256), 0, ctx().stream,
                       reinterpret_cast<uint4 *>(Lq.fat[cb]), Ld.fat[cb], vh, 18);
    hipLaunchKernelGGL(k_links_to_q16, dim3(nblk(vh)), dim3(256), 0, ctx().stream,
                       reinterpret_cast<uint4 *>(Lq.lng[cb]), Ld.lng[cb], vh, ncl);
    if (Ld.lsign[cb])
      HIP_CHECK(hipMemcpyAsync(Lq.lsign[cb], Ld.lsign[cb], (size_t)vh, hipMemcpyDeviceToDevice,
                               ctx().stream));
  }
  HIP_CHECK(hipGetLastError());
  return true;
}

const LinksT<float> *sloppy_links(LinksT<double> *L) {
  if (!L->sloppy) {
    auto *s = new LinksT<float>;
    if (!links_to_float(*L, *s)) {
      delete s;
      return nullptr;
    }
    L->sloppy = s;
  }
  return L->sloppy;
}

static int g_sloppy_long_recon = 8;

// The FP32 links of the mixed solver's sloppy operator: long links in their
// 8-real form when set (the default), the FP64 links keep 12 reals and every
// long link has that form; otherwise the FP32 copy of sloppy_links.
const LinksT<float> *sloppy_links_solver(LinksT<double> *L) {
  if (g_sloppy_long_recon == 8 && L->recon == 12 && !L->no8) {
    if (!L->sloppy8) {
      auto *s = new LinksT<float>;
      const int rc = links_to_float8(*L, *s);
      if (rc < 0) {
        delete s;
        return nullptr;
      }
      if (rc == 0) L->sloppy8 = s;
      else {
        delete s;
        L->no8 = true;
      }
    }
    if (L->sloppy8) return L->sloppy8;
  }
  return sloppy_links(L);
}

const LinksT<Q16> *sloppy_links16(LinksT<double> *L) {
  if (!L->sloppy16) {
    auto *s = new LinksT<Q16>;
    if (!links_to_q16(*L, *s)) {
      delete s;
      return nullptr;
    }
    L->sloppy16 = s;
  }
  return L->sloppy16;
}

static int g_sloppy_link_bits = 32;
int sloppy_link_bits() { return g_sloppy_link_bits; }

// ------------------------------------------------------------------ V
// Raw MILC order [site][colour][re, im] <-> the tile-major spinor storage
// (Spin<T>), one work-item per site of one parity.
template <typename T> __global__ void k_spin_from_raw(T *dst, const T *raw, int vh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  cplx<T> v[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) v[c] = cplx<T>{raw[6 * (size_t)i + 2 * c], raw[6 * (size_t)i + 2 * c + 1]};
  Spin<T>::store(dst, i, v);
}

template <typename T> __global__ void k_spin_to_raw(T *raw, const T *src, int vh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  cplx<T> v[3];
  Spin<T>::load(src, i, v);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    raw[6 * (size_t)i + 2 * c] = v[c].re;
    raw[6 * (size_t)i + 2 * c + 1] = v[c].im;
  }
}

// qphix_d3.cpp:79-174 (create_V): the raw half of each requested parity is
// copied to the device and reordered there.
template <typename T>
static typename Handles<T>::V *create_V(const T *src, int evenodd) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called:create_V", nullptr);
  const int vh = ctx().g.vh;
  const size_t nv = 6 * (size_t)vh;
  T *stage = nullptr;
  if (src) {
    stage = (T *)dev_alloc(sizeof(T) * nv);
    if (!stage) return nullptr;  // error recorded
  }
  auto *V = new typename Handles<T>::V;
  V->parity = evenodd;
  for (int cb = 0; cb < 2; ++cb) {
    const int bit = cb ? QPHIX_ODD : QPHIX_EVEN;
    if (!(evenodd & bit)) continue;
    V->half[cb] = (T *)dev_alloc(sizeof(T) * nv);
    if (!V->half[cb]) {  // dev_alloc has recorded the error
      dev_free(V->half[0]);
      dev_free(V->half[1]);
      dev_free(stage);
      delete V;
      return nullptr;
    }
    if (src) {
      h2d(stage, src + (cb ? nv : 0), sizeof(T) * nv);
      hipLaunchKernelGGL(k_spin_from_raw<T>, dim3((vh + 255) / 256), dim3(256), 0, ctx().stream,
                         V->half[cb], stage, vh);
      HIP_CHECK(hipGetLastError());
    } else {
      HIP_CHECK(hipMemsetAsync(V->half[cb], 0, sizeof(T) * nv, ctx().stream));
    }
  }
  sync();
  dev_free(stage);
  return V;
}

// The parities written are those the vector holds (src->parity), as in the
// reference (qphix_d3.cpp:495-553 ignores the evenodd argument).
template <typename T> static void extract_V(T *dest, const typename Handles<T>::V *V) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called:extract_V");
  QPHIX_REQUIRE(V && dest, "extract_V: null argument");
  const int vh = ctx().g.vh;
  const size_t nv = 6 * (size_t)vh;
  T *stage = (T *)dev_alloc(sizeof(T) * nv);
  if (!stage) return;  // error recorded
  for (int cb = 0; cb < 2; ++cb)
    if (V->half[cb]) {
      hipLaunchKernelGGL(k_spin_to_raw<T>, dim3((vh + 255) / 256), dim3(256), 0, ctx().stream, stage,
                         V->half[cb], vh);
      HIP_CHECK(hipGetLastError());
      d2h(dest + (cb ? nv : 0), stage, sizeof(T) * nv);
    }
  sync();
  dev_free(stage);
}

template <typename T> static void destroy_V(typename Handles<T>::V *V) {
  if (!V) return;
  if (ctx().initialised) HIP_CHECK(hipStreamSynchronize(ctx().stream));
  dev_free(V->half[0]);
  dev_free(V->half[1]);
  delete V;
}

// ------------------------------------------------------------------ L
// The import pipeline of create_L: fill(cb, stage, long_only) puts parity
// cb's raw links in MILC order into the device staging buffer, fat at stage,
// fat back at + n, long at + 2n, long back at + 3n (n = 72 vh reals); with
// long_only only the last two are needed.  Pass 1 decides the Naik storage
// (12 reals only if every supplied long link is SU(3) times a sign), pass 2
// reorders into the tile-major pairs.  Returns false when the fill failed.
template <typename T, typename Fill>
static typename Handles<T>::L *create_L_core(int evenodd, Fill &&fill) {
  const int vh = ctx().g.vh;
  const size_t nraw = 72 * (size_t)vh;
  auto *L = new typename Handles<T>::L;
  L->parity = evenodd;
  T *stage = (T *)dev_alloc(sizeof(T) * 4 * nraw);
  unsigned long long *dmax = (unsigned long long *)dev_alloc(sizeof(unsigned long long) * 4);
  if (!stage || !dmax) { dev_free(stage); dev_free(dmax); delete L; return nullptr; }
  uint8_t *signs[2] = {nullptr, nullptr};
  auto bail = [&]() -> typename Handles<T>::L * {
    sync();
    dev_free(signs[0]);
    dev_free(signs[1]);
    dev_free(stage);
    dev_free(dmax);
    links_free<T>(L);
    delete L;
    return nullptr;
  };
  // Pass 1: decide the long-link storage from every supplied long link.
  bool can12 = long_recon_mode() != 18;
  for (int cb = 0; cb < 2 && can12; ++cb) {
    if (!(evenodd & (cb ? QPHIX_ODD : QPHIX_EVEN))) continue;
    if (!fill(cb, stage, true)) return bail();
    HIP_CHECK(hipMemsetAsync(dmax, 0, sizeof(unsigned long long), ctx().stream));
    signs[cb] = (uint8_t *)dev_alloc((size_t)vh);
    if (!signs[cb]) return bail();  // error recorded
    hipLaunchKernelGGL(k_recon_check<T>, dim3(nblk(vh)), dim3(256), 0, ctx().stream, stage + 2 * nraw,
                       stage + 3 * nraw, vh, signs[cb], dmax);
    unsigned long long bits = 0;
    d2h(&bits, dmax, sizeof(bits));
    sync();
    double worst;
    std::memcpy(&worst, &bits, sizeof(worst));
    if (!(worst <= recon_tol<T>())) can12 = false;
  }
  L->recon = can12 ? 12 : 18;
  const int ncl = can12 ? 6 : 9;
  for (int cb = 0; cb < 2; ++cb) {
    if (!(evenodd & (cb ? QPHIX_ODD : QPHIX_EVEN))) continue;
    if (!fill(cb, stage, false)) return bail();
    L->fat[cb] = (T *)dev_alloc(sizeof(T) * 144 * link_sites(vh));
    L->lng[cb] = (T *)dev_alloc(sizeof(T) * 2 * 2 * ncl * 4 * link_sites(vh));
    if (!L->fat[cb] || !L->lng[cb]) return bail();  // error recorded
    hipLaunchKernelGGL(k_import_pairs<T>, dim3(nblk(vh)), dim3(256), 0, ctx().stream, L->fat[cb],
                       stage, stage + nraw, vh, 9);
    hipLaunchKernelGGL(k_import_pairs<T>, dim3(nblk(vh)), dim3(256), 0, ctx().stream, L->lng[cb],
                       stage + 2 * nraw, stage + 3 * nraw, vh, ncl);
    HIP_CHECK(hipGetLastError());
    sync();
    if (can12) { L->lsign[cb] = signs[cb]; signs[cb] = nullptr; }
  }
  dev_free(signs[0]);
  dev_free(signs[1]);
  dev_free(stage);
  dev_free(dmax);
  return L;
}

template <typename T>
static typename Handles<T>::L *create_L(const T *fat, const T *lng, const T *fatback,
                                        const T *lngback, int evenodd) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called:create_L", nullptr);
  QPHIX_REQUIRE(fat && lng && fatback && lngback, "create_L: null link array", nullptr);
  const size_t nraw = 72 * (size_t)ctx().g.vh;
  return create_L_core<T>(evenodd, [&](int cb, T *stage, bool long_only) {
    const size_t off = cb ? nraw : 0;
    if (!long_only) {
      h2d(stage, fat + off, sizeof(T) * nraw);
      h2d(stage + nraw, fatback + off, sizeof(T) * nraw);
    }
    h2d(stage + 2 * nraw, lng + off, sizeof(T) * nraw);
    h2d(stage + 3 * nraw, lngback + off, sizeof(T) * nraw);
    return true;
  });
}

// Links whose raw MILC-order blocks a device kernel writes (link_prep.hip).
template <typename T>
typename Handles<T>::L *create_L_device(int evenodd, const std::function<bool(int, T *)> &fill) {
  return create_L_core<T>(evenodd, [&](int cb, T *stage, bool) { return fill(cb, stage); });
}
template Handles<double>::L *create_L_device<double>(int, const std::function<bool(int, double *)> &);
template Handles<float>::L *create_L_device<float>(int, const std::function<bool(int, float *)> &);

template <typename T>
static void extract_L(T *fat, T *lng, const typename Handles<T>::L *L, int evenodd) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called:extract_L");
  QPHIX_REQUIRE(L && fat && lng, "extract_L: null argument");
  const int vh = ctx().g.vh;
  const size_t nraw = 72 * (size_t)vh;
  T *stage = (T *)dev_alloc(sizeof(T) * 2 * nraw);
  if (!stage) return;  // error recorded
  for (int cb = 0; cb < 2; ++cb) {
    if (!(evenodd & (cb ? QPHIX_ODD : QPHIX_EVEN))) continue;
    if (!L->fat[cb]) continue;
    hipLaunchKernelGGL(k_export_fwd<T>, dim3(nblk(vh)), dim3(256), 0, ctx().stream, L->fat[cb],
                       L->lng[cb], L->lsign[cb], vh, L->recon == 12 ? 6 : 9, stage, stage + nraw);
    d2h(fat + (cb ? nraw : 0), stage, sizeof(T) * nraw);
    d2h(lng + (cb ? nraw : 0), stage + nraw, sizeof(T) * nraw);
    sync();
  }
  dev_free(stage);
}

template <typename T> static void destroy_L(typename Handles<T>::L *L) {
  if (!L) return;
  if (ctx().initialised) HIP_CHECK(hipStreamSynchronize(ctx().stream));
  links_free<T>(L);
  delete L;
}

template <typename T> static void backcheck(const LinksT<T> *L, double out[4]) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called:backcheck");
  QPHIX_REQUIRE(L && L->fat[0] && L->fat[1], "backcheck needs both parities");
  const Geometry &gg = ctx().g;
  Geo g{gg.nxh, gg.ny, gg.nz, gg.nt, gg.vh, gg.nxy, gg.nxyz};
  unsigned long long *d = (unsigned long long *)dev_alloc(sizeof(unsigned long long) * 4);
  if (!d) return;  // error recorded
  HIP_CHECK(hipMemsetAsync(d, 0, sizeof(unsigned long long) * 4, ctx().stream));
  const int ncl = L->recon == 12 ? 6 : 9;
  for (int cb = 0; cb < 2; ++cb)
    hipLaunchKernelGGL(k_backcheck<T>, dim3(nblk(gg.vh)), dim3(256), 0, ctx().stream, L->fat[cb],
                       L->lng[cb], L->lsign[cb], L->fat[1 - cb], L->lng[1 - cb], L->lsign[1 - cb],
                       ncl, g, cb, d);
  unsigned long long h[4];
  d2h(h, d, sizeof(h));
  sync();
  for (int k = 0; k < 4; ++k) std::memcpy(&out[k], &h[k], sizeof(double));
  dev_free(d);
}

// Raw gauge links of one parity ([i][mu][3][3][re,im]) <-> complex-major layout.
template <typename T> __global__ void k_gauge_in(T *dst, const T *raw, int vh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  for (int mu = 0; mu < 4; ++mu)
    for (int e = 0; e < 9; ++e) {
      const size_t d = 2 * gauge_elem(mu, e, i, vh);
      dst[d] = raw[72 * (size_t)i + 18 * mu + 2 * e];
      dst[d + 1] = raw[72 * (size_t)i + 18 * mu + 2 * e + 1];
    }
}

template <typename T> __global__ void k_gauge_out(T *raw, const T *src, int vh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  for (int mu = 0; mu < 4; ++mu)
    for (int e = 0; e < 9; ++e) {
      const size_t d = 2 * gauge_elem(mu, e, i, vh);
      raw[72 * (size_t)i + 18 * mu + 2 * e] = src[d];
      raw[72 * (size_t)i + 18 * mu + 2 * e + 1] = src[d + 1];
    }
}

// Per link of one parity (complex-major): |r2 - conj(r0 x r1)|, the largest
// over the links into *maxbits (as the bits of a non-negative double).
template <typename T> __global__ void k_gauge_su3_check(const T *u, int vh, unsigned long long *maxbits) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double worst = 0;
  if (i < vh) {
    for (int mu = 0; mu < 4; ++mu) {
      double re[9], im[9];
      for (int e = 0; e < 9; ++e) {
        re[e] = u[2 * gauge_elem(mu, e, i, vh)];
        im[e] = u[2 * gauge_elem(mu, e, i, vh) + 1];
      }
      for (int c = 0; c < 3; ++c) {
        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        const double xr = re[c1] * re[3 + c2] - im[c1] * im[3 + c2] - (re[c2] * re[3 + c1] - im[c2] * im[3 + c1]);
        const double xi = re[c1] * im[3 + c2] + im[c1] * re[3 + c2] - (re[c2] * im[3 + c1] + im[c2] * re[3 + c1]);
        worst = fmax(worst, fabs(re[6 + c] - xr));
        worst = fmax(worst, fabs(im[6 + c] + xi));
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) worst = fmax(worst, __shfl_down(worst, off, 64));
  if ((threadIdx.x & 63) == 0) atomic_max_nonneg(maxbits, worst);
}

// Rows 0-1 of every link, complex-major with 6 elements per link.
template <typename T> __global__ void k_gauge_rows(T *dst, const T *u, int vh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  for (int mu = 0; mu < 4; ++mu)
    for (int e = 0; e < 6; ++e) {
      dst[2 * gauge_elem_n(mu, e, i, vh, 6)] = u[2 * gauge_elem(mu, e, i, vh)];
      dst[2 * gauge_elem_n(mu, e, i, vh, 6) + 1] = u[2 * gauge_elem(mu, e, i, vh) + 1];
    }
}

static int g_gauge_recon_mode = 0;  // 18: never keep the two-row copy

// ------------------------------------------------------------------ G, F
// Gauge field: all links are always loaded (the reference reads both raw
// blocks whatever the parity, qphix_d3.cpp:198-200); EVENODD is recorded as
// EVEN (qphix_d3.cpp:195).
template <typename T> static typename Handles<T>::G *create_G(const T *raw, int evenodd) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called:create_G", nullptr);
  QPHIX_REQUIRE(raw, "create_G: null array", nullptr);
  QPHIX_REQUIRE(evenodd == QPHIX_EVENODD || evenodd == QPHIX_EVEN || evenodd == QPHIX_ODD,
                "create_G: bad parity", nullptr);
  const int vh = ctx().g.vh;
  const size_t n = 72 * (size_t)vh;
  T *stage = (T *)dev_alloc(sizeof(T) * n);
  if (!stage) return nullptr;
  auto *G = new typename Handles<T>::G;
  G->parity = evenodd == QPHIX_EVENODD ? QPHIX_EVEN : evenodd;
  for (int cb = 0; cb < 2; ++cb) {
    G->u[cb] = (T *)dev_alloc(sizeof(T) * n);
    if (!G->u[cb]) {  // error recorded
      sync();
      dev_free(G->u[0]);
      dev_free(stage);
      delete G;
      return nullptr;
    }
    h2d(stage, raw + (cb ? n : 0), sizeof(T) * n);
    hipLaunchKernelGGL(k_gauge_in<T>, dim3(nblk(vh)), dim3(256), 0, ctx().stream, G->u[cb], stage, vh);
    HIP_CHECK(hipGetLastError());
    sync();
  }
  dev_free(stage);
  // Two-row copy for the force kernel when every link is SU(3) to rounding
  // (the exact links stay in u for extract_G).
  bool su3 = g_gauge_recon_mode != 18;
  unsigned long long *dmax = su3 ? (unsigned long long *)dev_try_alloc(sizeof(unsigned long long)) : nullptr;
  if (!dmax) su3 = false;
  for (int cb = 0; cb < 2 && su3; ++cb) {
    HIP_CHECK(hipMemsetAsync(dmax, 0, sizeof(unsigned long long), ctx().stream));
    hipLaunchKernelGGL(k_gauge_su3_check<T>, dim3(nblk(vh)), dim3(256), 0, ctx().stream, G->u[cb], vh, dmax);
    HIP_CHECK(hipGetLastError());
    unsigned long long bits = 0;
    d2h(&bits, dmax, sizeof(bits));
    sync();
    double worst;
    std::memcpy(&worst, &bits, sizeof(worst));
    if (!(worst <= recon_tol<T>())) su3 = false;
  }
  dev_free(dmax);
  if (su3) {
    for (int cb = 0; cb < 2; ++cb) G->u12[cb] = (T *)dev_try_alloc(sizeof(T) * 48 * (size_t)vh);
    if (!G->u12[0] || !G->u12[1]) {  // no room: the force kernel reads the full links
      dev_free(G->u12[0]);
      dev_free(G->u12[1]);
      G->u12[0] = G->u12[1] = nullptr;
    } else {
      for (int cb = 0; cb < 2; ++cb) {
        hipLaunchKernelGGL(k_gauge_rows<T>, dim3(nblk(vh)), dim3(256), 0, ctx().stream, G->u12[cb],
                           G->u[cb], vh);
        HIP_CHECK(hipGetLastError());
      }
    }
  }
  sync();
  return G;
}

template <typename T> static void extract_G(T *raw, const typename Handles<T>::G *G, int evenodd) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called:extract_G");
  QPHIX_REQUIRE(raw && G, "extract_G: null argument");
  const int vh = ctx().g.vh;
  const size_t n = 72 * (size_t)vh;
  T *stage = (T *)dev_alloc(sizeof(T) * n);
  if (!stage) return;
  for (int cb = 0; cb < 2; ++cb) {
    if (!(evenodd & (cb ? QPHIX_ODD : QPHIX_EVEN))) continue;
    hipLaunchKernelGGL(k_gauge_out<T>, dim3(nblk(vh)), dim3(256), 0, ctx().stream, stage, G->u[cb], vh);
    HIP_CHECK(hipGetLastError());
    d2h(raw + (cb ? n : 0), stage, sizeof(T) * n);
    sync();
  }
  dev_free(stage);
}

template <typename T> static void destroy_G(typename Handles<T>::G *G) {
  if (!G) return;
  if (ctx().initialised) HIP_CHECK(hipStreamSynchronize(ctx().stream));
  dev_free(G->u[0]);
  dev_free(G->u[1]);
  dev_free(G->u12[0]);
  dev_free(G->u12[1]);
  delete G;
}

// Momentum: both parities are always held (zero unless loaded), the requested
// parities are loaded (qphix_d3.cpp:360-488).
template <typename T> static typename Handles<T>::F *create_F(const T *raw, int evenodd) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called:create_F", nullptr);
  QPHIX_REQUIRE(raw, "create_F: null array", nullptr);
  QPHIX_REQUIRE(evenodd == QPHIX_EVENODD || evenodd == QPHIX_EVEN || evenodd == QPHIX_ODD,
                "create_F: bad parity", nullptr);
  const int vh = ctx().g.vh;
  const size_t n = 72 * (size_t)vh;
  T *stage = (T *)dev_alloc(sizeof(T) * n);
  if (!stage) return nullptr;  // error recorded
  auto *F = new typename Handles<T>::F;
  F->parity = evenodd == QPHIX_EVENODD ? QPHIX_EVEN : evenodd;
  for (int cb = 0; cb < 2; ++cb) {
    F->h[cb] = (T *)dev_alloc(sizeof(T) * 32 * (size_t)vh);
    if (!F->h[cb]) {  // error recorded
      sync();
      dev_free(F->h[0]);
      dev_free(stage);
      delete F;
      return nullptr;
    }
    HIP_CHECK(hipMemsetAsync(F->h[cb], 0, sizeof(T) * 32 * (size_t)vh, ctx().stream));
    if (!(evenodd & (cb ? QPHIX_ODD : QPHIX_EVEN))) continue;
    h2d(stage, raw + (cb ? n : 0), sizeof(T) * n);
    hipLaunchKernelGGL(k_mom_in<T>, dim3(nblk(4 * vh)), dim3(256), 0, ctx().stream, F->h[cb],
                       stage, 4 * vh);
    HIP_CHECK(hipGetLastError());
    sync();
  }
  sync();
  dev_free(stage);
  return F;
}

template <typename T> static void extract_F(T *raw, const typename Handles<T>::F *F, int evenodd) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called:extract_F");
  QPHIX_REQUIRE(raw && F, "extract_F: null argument");
  const int vh = ctx().g.vh;
  const size_t n = 72 * (size_t)vh;
  T *stage = (T *)dev_alloc(sizeof(T) * n);
  if (!stage) return;  // error recorded
  for (int cb = 0; cb < 2; ++cb) {
    if (!(evenodd & (cb ? QPHIX_ODD : QPHIX_EVEN))) continue;
    hipLaunchKernelGGL(k_mom_out<T>, dim3(nblk(4 * vh)), dim3(256), 0, ctx().stream, stage,
                       F->h[cb], 4 * vh);
    d2h(raw + (cb ? n : 0), stage, sizeof(T) * n);
    sync();
  }
  dev_free(stage);
}

template <typename T> static void destroy_F(typename Handles<T>::F *F) {
  if (!F) return;
  if (ctx().initialised) HIP_CHECK(hipStreamSynchronize(ctx().stream));
  dev_free(F->h[0]);
  dev_free(F->h[1]);
  delete F;
}

} // namespace qphix

using namespace qphix;

extern "C" {

#define QPHIX_FIELD_API(P, T)                                                                     \
  QPHIX_##P##3_ColorVector *QPHIX_##P##3_create_V_from_raw(T *src, QPHIX_evenodd_t evenodd) {     \
    return create_V<T>(src, evenodd);                                                             \
  }                                                                                               \
  void QPHIX_##P##3_extract_V_to_raw(T *dest, QPHIX_##P##3_ColorVector *src, QPHIX_evenodd_t) {   \
    extract_V<T>(dest, src);                                                                      \
  }                                                                                               \
  void QPHIX_##P##3_destroy_V(QPHIX_##P##3_ColorVector *V) { destroy_V<T>(V); }                   \
  QPHIX_##P##3_FermionLinksAsqtad *QPHIX_##P##3_asqtad_create_L_from_raw(                         \
      T *fat, T *lng, T *fatback, T *lngback, QPHIX_evenodd_t evenodd) {                          \
    return create_L<T>(fat, lng, fatback, lngback, evenodd);                                      \
  }                                                                                               \
  void QPHIX_##P##3_asqtad_extract_L_to_raw(T *fat, T *lng, QPHIX_##P##3_FermionLinksAsqtad *src, \
                                            QPHIX_evenodd_t evenodd) {                            \
    extract_L<T>(fat, lng, src, evenodd);                                                         \
  }                                                                                               \
  void QPHIX_##P##3_asqtad_destroy_L(QPHIX_##P##3_FermionLinksAsqtad *L) { destroy_L<T>(L); }    \
  int QPHIX_##P##3_asqtad_links_recon(QPHIX_##P##3_FermionLinksAsqtad *L) {                       \
    return L ? L->recon : 0;                                                                      \
  }                                                                                               \
  QPHIX_##P##3_GaugeField *QPHIX_##P##3_create_G_from_raw(T *raw, QPHIX_evenodd_t evenodd) {      \
    return create_G<T>(raw, evenodd);                                                             \
  }                                                                                               \
  void QPHIX_##P##3_extract_G_to_raw(T *raw, QPHIX_##P##3_GaugeField *G, QPHIX_evenodd_t eo) {    \
    extract_G<T>(raw, G, eo);                                                                     \
  }                                                                                               \
  void QPHIX_##P##3_destroy_G(QPHIX_##P##3_GaugeField *G) { destroy_G<T>(G); }                    \
  int QPHIX_##P##3_gauge_links_recon(QPHIX_##P##3_GaugeField *G) {                                \
    return G && G->u12[0] ? 12 : 18;                                                              \
  }                                                                                               \
  QPHIX_##P##3_Force *QPHIX_##P##3_create_F_from_raw(T *raw, QPHIX_evenodd_t evenodd) {           \
    return create_F<T>(raw, evenodd);                                                             \
  }                                                                                               \
  void QPHIX_##P##3_extract_F_to_raw(T *raw, QPHIX_##P##3_Force *F, QPHIX_evenodd_t eo) {         \
    extract_F<T>(raw, F, eo);                                                                     \
  }                                                                                               \
  void QPHIX_##P##3_destroy_F(QPHIX_##P##3_Force *F) { destroy_F<T>(F); }

QPHIX_FIELD_API(D, double)
QPHIX_FIELD_API(F, float)

void QPHIX_MI355X_set_long_recon(int mode) { g_recon_mode = mode == 18 ? 18 : 0; }
void QPHIX_MI355X_set_gauge_recon(int mode) { g_gauge_recon_mode = mode == 18 ? 18 : 0; }
void QPHIX_MI355X_set_sloppy_link_bits(int bits) { g_sloppy_link_bits = bits == 16 ? 16 : 32; }
void QPHIX_MI355X_set_sloppy_long_recon(int recon) { g_sloppy_long_recon = recon == 12 ? 12 : 8; }

void QPHIX_D3_asqtad_links_backcheck(QPHIX_D3_FermionLinksAsqtad *L, double out[4]) {
  backcheck<double>(L, out);
}

} // extern "C"
