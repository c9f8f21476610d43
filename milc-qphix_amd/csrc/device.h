// device.h -- device-side helpers: complex arithmetic, spinor and link access,
// lattice neighbour arithmetic.
#pragma once
#include "common.h"

namespace qphix {

template <typename C> struct cplx {
  C re, im;
};

template <typename C> __device__ __forceinline__ void cmac(cplx<C> &acc, cplx<C> a, cplx<C> b) {
  acc.re = fma(a.re, b.re, acc.re);
  acc.re = fma(-a.im, b.im, acc.re);
  acc.im = fma(a.re, b.im, acc.im);
  acc.im = fma(a.im, b.re, acc.im);
}

// acc += U v for U given row-major as 9 complex numbers
template <typename C>
__device__ __forceinline__ void matvec_acc(const cplx<C> *U, const cplx<C> *v, cplx<C> *acc) {
#pragma unroll
  for (int r = 0; r < 3; ++r) {
#pragma unroll
    for (int c = 0; c < 3; ++c) cmac(acc[r], U[3 * r + c], v[c]);
  }
}

// Per-site term of the relative residue (ks_blas_utils_c.cpp:110-149): each
// colour's |v_c|^2 is formed in the vector type and widened to double, as the
// reference's calc_soalen_rel_residue does (:118-121).
template <typename C> __device__ __forceinline__ double site_norm2_rounded(const cplx<C> *v) {
  double s = 0;
#pragma unroll
  for (int c = 0; c < 3; ++c) s += (double)(C)(v[c].re * v[c].re + v[c].im * v[c].im);
  return s;
}
// num / den with the reference's special cases: r = 0 skipped, x = 0 counts 1.
__device__ __forceinline__ double rel_term(double num, double den) {
  return num == 0 ? 0.0 : (den == 0 ? 1.0 : num / den);
}

// Third row of a link from its first two: s * conj(r0 x r1).
template <typename C>
__device__ __forceinline__ void recon_row2(cplx<C> *U, C s) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
    // (r0 x r1)_c = r0[c1] r1[c2] - r0[c2] r1[c1], as explicit fmas: the
    // rounding is then fixed by this code, not by where the compiler fuses
    C re = U[c1].re * U[3 + c2].re;
    re = fma(-U[c1].im, U[3 + c2].im, re);
    re = fma(-U[c2].re, U[3 + c1].re, re);
    re = fma(U[c2].im, U[3 + c1].im, re);
    C im = U[c1].re * U[3 + c2].im;
    im = fma(U[c1].im, U[3 + c2].re, im);
    im = fma(-U[c2].re, U[3 + c1].im, im);
    im = fma(-U[c2].im, U[3 + c1].re, im);
    U[6 + c].re = s * re;
    U[6 + c].im = -s * im;
  }
}

// A link U = s V (V in SU(3), s = +-1, so det U = s) from 8 reals:
// q[0] = U01, q[1] = U02, q[2] = U10, q[3] = (tan(arg U00 / 2), tan(arg U20 / 2)).
// Unitarity gives n = |U01|^2 + |U02|^2 = 1 - |U00|^2 = |U10|^2 + |U20|^2, so
// |U00| = sqrt(1 - n) and |U20| = sqrt(n - |U10|^2); a half-angle tangent h
// gives the unit phase ((1 - h^2), 2h) / (1 + h^2).  The 2x2 block solves the
// orthogonality of column 0 to columns 1 and 2 together with the cofactor
// identities of det U = s, conj(U01) = s (U12 U20 - U10 U22) and
// conj(U02) = s (U10 U21 - U11 U20):
//   U11 = -(conj(U00) U01 U10 + s conj(U20) conj(U02)) / n
//   U21 =  (s conj(U10) conj(U02) - conj(U00) U01 U20) / n
//   U12 =  (s conj(U20) conj(U01) - conj(U00) U02 U10) / n
//   U22 = -(s conj(U10) conj(U01) + conj(U00) U02 U20) / n
// Every term of a numerator is a product of two of the small off-diagonal
// elements, so the quotient keeps FP32 relative precision as n shrinks; n = 0
// (U block-diagonal) has no 8-real form and is rejected when links are packed.
// Hardware reciprocal and square root (1 ulp): this form feeds sloppy solver
// operators only.
__device__ __forceinline__ void recon8(const cplx<float> *q, float s, cplx<float> *U) {
  const cplx<float> a1 = q[0], a2 = q[1], b0 = q[2];
  const float n = fmaf(a1.re, a1.re, fmaf(a1.im, a1.im, fmaf(a2.re, a2.re, a2.im * a2.im)));
  const float ma = __builtin_amdgcn_sqrtf(fmaxf(1.f - n, 0.f));
  const float mc = __builtin_amdgcn_sqrtf(fmaxf(n - fmaf(b0.re, b0.re, b0.im * b0.im), 0.f));
  const float ha = q[3].re, hc = q[3].im;
  const float ra = ma * __builtin_amdgcn_rcpf(fmaf(ha, ha, 1.f));
  const float rc = mc * __builtin_amdgcn_rcpf(fmaf(hc, hc, 1.f));
  const cplx<float> a0{fmaf(-ha, ha, 1.f) * ra, 2.f * ha * ra};
  const cplx<float> c0{fmaf(-hc, hc, 1.f) * rc, 2.f * hc * rc};
  const float inv = __builtin_amdgcn_rcpf(n);
  auto mul = [](cplx<float> x, cplx<float> y) {
    return cplx<float>{fmaf(x.re, y.re, -x.im * y.im), fmaf(x.re, y.im, x.im * y.re)};
  };
  auto cjmul = [](cplx<float> x, cplx<float> y) {  // conj(x) y
    return cplx<float>{fmaf(x.re, y.re, x.im * y.im), fmaf(x.re, y.im, -x.im * y.re)};
  };
  const cplx<float> t1 = cjmul(a0, a1), t2 = cjmul(a0, a2);
  const cplx<float> sa1{s * a1.re, -s * a1.im}, sa2{s * a2.re, -s * a2.im};
  const cplx<float> p11 = mul(t1, b0), q11 = cjmul(c0, sa2);
  const cplx<float> p21 = cjmul(b0, sa2), q21 = mul(t1, c0);
  const cplx<float> p12 = cjmul(c0, sa1), q12 = mul(t2, b0);
  const cplx<float> p22 = cjmul(b0, sa1), q22 = mul(t2, c0);
  U[0] = a0;
  U[1] = a1;
  U[2] = a2;
  U[3] = b0;
  U[4] = cplx<float>{-(p11.re + q11.re) * inv, -(p11.im + q11.im) * inv};
  U[5] = cplx<float>{(p12.re - q12.re) * inv, (p12.im - q12.im) * inv};
  U[6] = c0;
  U[7] = cplx<float>{(p21.re - q21.re) * inv, (p21.im - q21.im) * inv};
  U[8] = cplx<float>{-(p22.re + q22.re) * inv, -(p22.im + q22.im) * inv};
}

// ---------------------------------------------------------------------------
// Spinor (colour vector) storage traits.  Compute type is double for double
// storage and float otherwise.
//
// double and float vectors are stored tile-major over 64 sites: colour c of
// site i is complex element spin_cidx(i, c) = ((i / 64) * 3 + c) * 64 + i % 64.
// (Measured and dropped: FP32 tiles with colours 0-1 packed into one 16-byte
// word per site, 2 loads per gather instead of 3: the FP32-link Dslash at 32^4
// 89 -> 93 us, the 16-bit-link one 62 -> 57 us.)
// A wavefront's 64 consecutive sites then read or write one contiguous block
// per colour (1 KB FP64, 512 B FP32), and a neighbour gather touches a third of
// the cache lines the site-major [site][colour] order needs (FP64 Dslash at
// 32^4: 186.6 -> 172.5 us).  Every lattice extent is a multiple of 4, so the
// sites of one parity (and the NR-interleaved site * NR + r index of the
// multi-RHS vectors) are a multiple of 128: tiles never straddle the end of a
// vector, and vector offsets of 6 * vh * k reals start on a tile.
// ---------------------------------------------------------------------------
constexpr int SPIN_TILE = 64;
__host__ __device__ __forceinline__ size_t spin_cidx(size_t i, int c) {
  return ((i / SPIN_TILE) * 3 + (size_t)c) * SPIN_TILE + i % SPIN_TILE;
}

template <typename S> struct Spin;

template <> struct Spin<double> {
  using C = double;
  static __device__ __forceinline__ void load(const double *p, int i, cplx<double> *v) {
    const double2 *q = reinterpret_cast<const double2 *>(p) + spin_cidx((size_t)i, 0);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double2 d = q[SPIN_TILE * c];
      v[c].re = d.x;
      v[c].im = d.y;
    }
  }
  static __device__ __forceinline__ void store(double *p, int i, const cplx<double> *v) {
    double2 *q = reinterpret_cast<double2 *>(p) + spin_cidx((size_t)i, 0);
#pragma unroll
    for (int c = 0; c < 3; ++c) q[SPIN_TILE * c] = make_double2(v[c].re, v[c].im);
  }
  static __device__ __forceinline__ void store_nt(double *p, int i, const cplx<double> *v) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    d2 *q = reinterpret_cast<d2 *>(p) + spin_cidx((size_t)i, 0);
#pragma unroll
    for (int c = 0; c < 3; ++c) __builtin_nontemporal_store(d2{v[c].re, v[c].im}, q + SPIN_TILE * c);
  }
};

template <> struct Spin<float> {
  using C = float;
  static __device__ __forceinline__ void load(const float *p, int i, cplx<float> *v) {
    const float2 *q = reinterpret_cast<const float2 *>(p) + spin_cidx((size_t)i, 0);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float2 d = q[SPIN_TILE * c];
      v[c].re = d.x;
      v[c].im = d.y;
    }
  }
  static __device__ __forceinline__ void store(float *p, int i, const cplx<float> *v) {
    float2 *q = reinterpret_cast<float2 *>(p) + spin_cidx((size_t)i, 0);
#pragma unroll
    for (int c = 0; c < 3; ++c) q[SPIN_TILE * c] = make_float2(v[c].re, v[c].im);
  }
  static __device__ __forceinline__ void store_nt(float *p, int i, const cplx<float> *v) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    f2 *q = reinterpret_cast<f2 *>(p) + spin_cidx((size_t)i, 0);
#pragma unroll
    for (int c = 0; c < 3; ++c) __builtin_nontemporal_store(f2{v[c].re, v[c].im}, q + SPIN_TILE * c);
  }
};

template <> struct Spin<HalfSite> {
  using C = float;
  static __device__ __forceinline__ void load(const HalfSite *p, int i, cplx<float> *v) {
    const uint4 raw = *reinterpret_cast<const uint4 *>(p + i);
    const __half2 h0 = *reinterpret_cast<const __half2 *>(&raw.x);
    const __half2 h1 = *reinterpret_cast<const __half2 *>(&raw.y);
    const __half2 h2 = *reinterpret_cast<const __half2 *>(&raw.z);
    const float nrm = __uint_as_float(raw.w);
    const float2 f0 = __half22float2(h0), f1 = __half22float2(h1), f2 = __half22float2(h2);
    v[0].re = f0.x * nrm; v[0].im = f0.y * nrm;
    v[1].re = f1.x * nrm; v[1].im = f1.y * nrm;
    v[2].re = f2.x * nrm; v[2].im = f2.y * nrm;
  }
  static __device__ __forceinline__ void store(HalfSite *p, int i, const cplx<float> *v) {
    float m = 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) m = fmaxf(m, fmaxf(fabsf(v[c].re), fabsf(v[c].im)));
    const float inv = m > 0.f ? 1.f / m : 0.f;
    uint4 raw;
    __half2 h0 = __floats2half2_rn(v[0].re * inv, v[0].im * inv);
    __half2 h1 = __floats2half2_rn(v[1].re * inv, v[1].im * inv);
    __half2 h2 = __floats2half2_rn(v[2].re * inv, v[2].im * inv);
    raw.x = *reinterpret_cast<unsigned *>(&h0);
    raw.y = *reinterpret_cast<unsigned *>(&h1);
    raw.z = *reinterpret_cast<unsigned *>(&h2);
    raw.w = __float_as_uint(m);
    *reinterpret_cast<uint4 *>(p + i) = raw;
  }
  static __device__ __forceinline__ void store_nt(HalfSite *p, int i, const cplx<float> *v) {
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    float m = 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) m = fmaxf(m, fmaxf(fabsf(v[c].re), fabsf(v[c].im)));
    const float inv = m > 0.f ? 1.f / m : 0.f;
    __half2 h0 = __floats2half2_rn(v[0].re * inv, v[0].im * inv);
    __half2 h1 = __floats2half2_rn(v[1].re * inv, v[1].im * inv);
    __half2 h2 = __floats2half2_rn(v[2].re * inv, v[2].im * inv);
    const u4 raw = {*reinterpret_cast<unsigned *>(&h0), *reinterpret_cast<unsigned *>(&h1),
                    *reinterpret_cast<unsigned *>(&h2), __float_as_uint(m)};
    __builtin_nontemporal_store(raw, reinterpret_cast<u4 *>(p + i));
  }
};

// ---------------------------------------------------------------------------
// Link pair loads from the chunk-major layout (see common.h).
// NC complex numbers per pair; chunk = 16 bytes.
// ---------------------------------------------------------------------------
typedef double nd2 __attribute__((ext_vector_type(2)));
typedef float nf4 __attribute__((ext_vector_type(4)));

// NT = true: streaming (non-temporal) loads, so links read once per
// application do not evict the neighbour spinors from L2 / Infinity Cache.
template <typename T, int NC, bool NT = false> struct PairLoad;

template <int NC, bool NT> struct PairLoad<double, NC, NT> {
  static __device__ __forceinline__ void load(const double2 *base, int vh, int i, cplx<double> *u) {
    const nd2 *b = reinterpret_cast<const nd2 *>(base);
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const nd2 d = NT ? __builtin_nontemporal_load(b + (size_t)k * vh + i) : b[(size_t)k * vh + i];
      u[k].re = d.x;
      u[k].im = d.y;
    }
  }
};

template <int NC, bool NT> struct PairLoad<float, NC, NT> {
  static __device__ __forceinline__ void load(const float4 *base, int vh, int i, cplx<float> *u) {
    const nf4 *b = reinterpret_cast<const nf4 *>(base);
#pragma unroll
    for (int k = 0; k < NC / 2; ++k) {
      const nf4 d = NT ? __builtin_nontemporal_load(b + (size_t)k * vh + i) : b[(size_t)k * vh + i];
      u[2 * k].re = d.x;
      u[2 * k].im = d.y;
      u[2 * k + 1].re = d.z;
      u[2 * k + 1].im = d.w;
    }
  }
};

// 16-bit fixed-point pair (see Q16): NC complex elements decoded to FP32.
typedef unsigned nu4 __attribute__((ext_vector_type(4)));
template <int NC, bool NT> struct PairLoad<Q16, NC, NT> {
  static __device__ __forceinline__ void load(const uint4 *base, int vh, int i, cplx<float> *u) {
    constexpr int NCH = link_nch<Q16>(NC);
    const nu4 *b = reinterpret_cast<const nu4 *>(base);
    nu4 c[NCH];
#pragma unroll
    for (int k = 0; k < NCH; ++k) c[k] = NT ? __builtin_nontemporal_load(b + (size_t)k * vh + i) : b[(size_t)k * vh + i];
    const float sf = __uint_as_float(c[NCH - 1][2]) * (1.0f / 32767.0f);
    const float sb = __uint_as_float(c[NCH - 1][3]) * (1.0f / 32767.0f);
#pragma unroll
    for (int e = 0; e < 2 * NC; ++e) {
      const unsigned w = c[e / 8][(e % 8) / 2];
      const int q = (e % 2) ? ((int)w >> 16) : (int)(short)(w & 0xffffu);
      const float x = (float)q * (e < NC ? sf : sb);
      if (e % 2) u[e / 2].im = x;
      else u[e / 2].re = x;
    }
  }
};

// ---------------------------------------------------------------------------
// Lattice arithmetic.  cb = parity of the site i; the returned indices are
// raw indices of the opposite parity.
// ---------------------------------------------------------------------------
struct Geo {
  int nxh, ny, nz, nt, vh, nxy, nxyz;
};

struct SiteNbr {
  int f1[4], b1[4], f3[4], b3[4];
};

__device__ __forceinline__ void site_coords(const Geo &g, int i, int &xh, int &y, int &z, int &t) {
  xh = i % g.nxh;
  int r = i / g.nxh;
  y = r % g.ny;
  r /= g.ny;
  z = r % g.nz;
  t = r / g.nz;
}

__device__ __forceinline__ int wrapc(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

__device__ __forceinline__ void site_neighbours(const Geo &g, int i, int cb, SiteNbr &n) {
  int xh, y, z, t;
  site_coords(g, i, xh, y, z, t);
  const int xodd = (y + z + t + cb) & 1;
  const int row = i - xh;
  n.f1[0] = row + wrapc(xh + xodd, g.nxh);
  n.b1[0] = row + wrapc(xh + xodd - 1, g.nxh);
  n.f3[0] = row + wrapc(xh + xodd + 1, g.nxh);
  n.b3[0] = row + wrapc(xh + xodd - 2, g.nxh);
  // y, z, t: single wrap suffices because every extent is >= 4
  const int cs[3] = {y, z, t};
  const int len[3] = {g.ny, g.nz, g.nt};
  const int st[3] = {g.nxh, g.nxy, g.nxyz};
  const int per[3] = {g.nxy, g.nxyz, g.vh};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int c = cs[d], L = len[d], s = st[d], P = per[d];
    n.f1[d + 1] = i + s - (c + 1 >= L ? P : 0);
    n.b1[d + 1] = i - s + (c - 1 < 0 ? P : 0);
    n.f3[d + 1] = i + 3 * s - (c + 3 >= L ? P : 0);
    n.b3[d + 1] = i - 3 * s + (c - 3 < 0 ? P : 0);
  }
}

} // namespace qphix
