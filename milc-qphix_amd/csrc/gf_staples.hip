// gf_staples.hip -- Symanzik gauge force built from stored plaquette staples.
//
// Same quantity as the loop-walk kernel of gauge_force.hip (the semantics of
// QPHIX_*3_symanzik_1loop_gauge_force, gauge_force_imp.cpp:136-199): for every
// link, W_mu(x) = U_mu(x) * sum over the 48 loops through it of c_P R_P, R_P
// the path from x + mu back to x, and mom_mu(x) <- mom_mu(x) - eps TA(W).
// Used when every link is SU(3) to rounding (the field keeps its two-row
// copy): the rectangles and parallelograms are then products of plaquette
// staples, because a link followed by its inverse cancels.
//
// With L_s(z) the link of a unit step s = +-k from z (L_{+k}(z) = U_k(z),
// L_{-k}(z) = U_k(z-k)^dag) and, for mu and a signed direction n orthogonal
// to it, the staple
//   Q_{mu,n}(y) = L_n(y+mu) L_{-mu}(y+mu+n) L_{-n}(y+n)     (y+mu -> y),
// the staple sum of link (x, mu) is
//   c_p sum_n Q_{mu,n}(x)                                       plaquettes
// + c_r U_mu(x+mu) sum_n Q_{mu,n}(x+mu) Q_{mu,n}(x)             rectangles mu,n,-mu,-mu,-n
// + c_r [sum_n Q_{mu,n}(x) Q_{mu,n}(x-mu)] U_mu(x-mu)           rectangles n,-mu,-mu,-n,mu
// + c_r sum_n L_n(x+mu) Q_{mu,n}(x+n) L_{-n}(x+n)               rectangles n,n,-mu,-n,-n
// + c_g sum_A L_A(x+mu) sum_{d _|_ mu,A} Q_{mu,d}(x+A) Q_{A,d}(x)   parallelograms A,d,-mu,-A,-d
// with Q_{-a,d}(x) = Q_{a,d}(x-a)^dag.  A rectangle n, n, -mu, -n, -n and the
// parallelograms with A = n begin with the same link, so each such group is
// one bracket times L_n(x+mu): 51 SU(3) products per link (with W = U S)
// instead of the walk kernel's 140, plus 10.5 per link to form the staples.
//
// Kernel 1 (k_gq_staples): one work-item per (site, pair mu < nu), the four
// staples Q_{mu,+-nu}, Q_{nu,+-mu}; the staples are stored as their rows 0-1 (they are SU(3)), complex-
// major: Q[parity][q = 6 mu + 2 rank(nu) + (n < 0)][element 0..5][site].
// Kernel 2 (k_gq_force): one work-item per link, as the walk kernel, with the
// 8 (mu, parity) blocks of a site block next to each other and each XCD
// sweeping a contiguous slab.
#include "blas.h"
#include "handles.h"

#include <chrono>
#include <cstdlib>
#include <type_traits>

namespace qphix {

struct GqGeo {
  int L[4];
  int nxh, vh;
  int stride[4];  // lexicographic strides 1, L0, L0 L1, L0 L1 L2
  int zreg;       // 1: each XCD's slab of site blocks is a z-region (gq_site)
};

// Site of logical index Lg: itself, or with zreg the XCD x = Lg / (vh / 8)
// takes the z-planes [x nz / 8, (x + 1) nz / 8) t-slice by t-slice, so the 8
// XCDs advance through t together and the +-1 neighbourhood they read is a
// few planes of each region instead of whole t-slices.
__device__ __forceinline__ int gq_site(const GqGeo &g, int Lg) {
  if (!g.zreg) return Lg;
  const int nxy = g.nxh * g.L[1], nzl = g.L[2] / 8;
  const int per = nxy * nzl * g.L[3], slice = nxy * nzl;
  const int x = Lg / per;
  int r = Lg - x * per;
  const int t = r / slice;
  r -= t * slice;
  const int z = x * nzl + r / nxy;
  return (r % nxy) + nxy * (z + g.L[2] * t);
}

namespace {


struct M3 {
  double re[9], im[9];
};

__device__ __forceinline__ void m_zero(M3 &a) {
#pragma unroll
  for (int e = 0; e < 9; ++e) a.re[e] = a.im[e] = 0.0;
}

// Row 2 of an SU(3) matrix from rows 0-1: conj(r0 x r1).
__device__ __forceinline__ void m_row2(M3 &m) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
    const double ar = m.re[c1], ai = m.im[c1], br = m.re[3 + c2], bi = m.im[3 + c2];
    const double er = m.re[c2], ei = m.im[c2], fr = m.re[3 + c1], fi = m.im[3 + c1];
    m.re[6 + c] = (ar * br - ai * bi) - (er * fr - ei * fi);
    m.im[6 + c] = -((ar * bi + ai * br) - (er * fi + ei * fr));
  }
}

// Element (i, j) of op(a): a, or a^dag when D.
template <bool D> __device__ __forceinline__ double ere(const M3 &a, int i, int j) {
  return D ? a.re[3 * j + i] : a.re[3 * i + j];
}
template <bool D> __device__ __forceinline__ double eim(const M3 &a, int i, int j) {
  return D ? -a.im[3 * j + i] : a.im[3 * i + j];
}

// r += op(a) op(b)
template <bool DA, bool DB> __device__ __forceinline__ void m_muladd(const M3 &a, const M3 &b, M3 &r) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double sr = r.re[3 * i + j], si = r.im[3 * i + j];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double ar = ere<DA>(a, i, k), ai = eim<DA>(a, i, k);
        const double br = ere<DB>(b, k, j), bi = eim<DB>(b, k, j);
        sr = fma(ar, br, sr);
        sr = fma(-ai, bi, sr);
        si = fma(ar, bi, si);
        si = fma(ai, br, si);
      }
      r.re[3 * i + j] = sr;
      r.im[3 * i + j] = si;
    }
}

// a *= c
__device__ __forceinline__ void m_scale(M3 &a, double c) {
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    a.re[e] *= c;
    a.im[e] *= c;
  }
}

// r = op(a) op(b)
template <bool DA, bool DB> __device__ __forceinline__ void m_mul(const M3 &a, const M3 &b, M3 &r) {
  m_zero(r);
  m_muladd<DA, DB>(a, b, r);
}

// s += c op(a) op(b), one row of temporaries
template <bool DA, bool DB>
__device__ __forceinline__ void m_muladd_scaled(double c, const M3 &a, const M3 &b, M3 &s) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double tr[3], ti[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double sr = 0, si = 0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double ar = ere<DA>(a, i, k), ai = eim<DA>(a, i, k);
        const double br = ere<DB>(b, k, j), bi = eim<DB>(b, k, j);
        sr = fma(ar, br, sr);
        sr = fma(-ai, bi, sr);
        si = fma(ar, bi, si);
        si = fma(ai, br, si);
      }
      tr[j] = sr;
      ti[j] = si;
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      s.re[3 * i + j] = fma(c, tr[j], s.re[3 * i + j]);
      s.im[3 * i + j] = fma(c, ti[j], s.im[3 * i + j]);
    }
  }
}

// A two-row matrix (rows 0-1 complex-major: element e of matrix m at
// base[(m * 6 + e) * vh + i]), row 2 rebuilt.  `base` and `m` are uniform.
template <typename T>
__device__ __forceinline__ void load2(const T *base, int m, int vh, int i, M3 &a) {
  using T2 = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
  const T2 *b = reinterpret_cast<const T2 *>(base) + (size_t)(m * 6) * (size_t)vh;
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const T2 v = b[(size_t)e * (size_t)vh + i];
    a.re[e] = (double)v.x;
    a.im[e] = (double)v.y;
  }
  m_row2(a);
}

// NT: streaming (nontemporal) stores, which do not keep the lines in L2.
template <typename T, bool NT = false>
__device__ __forceinline__ void store2(T *base, int m, int vh, int i, const M3 &a) {
  typedef T V2 __attribute__((ext_vector_type(2)));
  V2 *b = reinterpret_cast<V2 *>(base) + (size_t)(m * 6) * (size_t)vh;
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const V2 v{(T)a.re[e], (T)a.im[e]};
    if constexpr (NT) __builtin_nontemporal_store(v, b + (size_t)e * (size_t)vh + i);
    else b[(size_t)e * (size_t)vh + i] = v;
  }
}

// Coordinates of half-site i of parity cb.
__device__ __forceinline__ void q_coords(const GqGeo &g, int i, int cb, int c[4]) {
  const int xh = i % g.nxh;
  int r = i / g.nxh;
  c[1] = r % g.L[1];
  r /= g.L[1];
  c[2] = r % g.L[2];
  c[3] = r / g.L[2];
  c[0] = 2 * xh + ((c[1] + c[2] + c[3] + cb) & 1);
}

// Coordinate of direction d (uniform) without indexing a per-lane array.
__device__ __forceinline__ int q_coord(const int c[4], int d) {
  return d == 0 ? c[0] : d == 1 ? c[1] : d == 2 ? c[2] : c[3];
}

// Lexicographic index r moved by s (+-1) along d (uniform), coordinate cd
// along d before the move; wraps periodically.
__device__ __forceinline__ int q_move(const GqGeo &g, int r, int cd, int d, int s) {
  const int L = g.L[d], st = g.stride[d];
  const int v = cd + s;
  r += s * st;
  if (v < 0) r += L * st;
  else if (v >= L) r -= L * st;
  return r;
}

// The three directions other than mu, in increasing order; rank of d among
// the directions other than m.
__device__ __forceinline__ int q_other(int mu, int k) { return k < mu ? k : k + 1; }
__device__ __forceinline__ int q_rank(int m, int d) { return d < m ? d : d - 1; }

// Staple slot q of (link direction m, signed direction: d, negative = neg).
__device__ __forceinline__ int q_slot(int m, int d, bool neg) { return 6 * m + 2 * q_rank(m, d) + (neg ? 1 : 0); }

} // namespace

// Block -> slot: each XCD (blocks are dispatched round-robin over 8 XCDs)
// sweeps one contiguous slab of slots.
__device__ __forceinline__ int gq_slot() {
  const int nb = gridDim.x, b = blockIdx.x;
  return nb % 8 == 0 ? (b % 8) * (nb / 8) + b / 8 : b;
}

// Kernel 1.  One block per (256 sites of one parity, pair mu < nu, parity),
// the 12 (pair, parity) blocks of a site block in adjacent slots, so that the
// links they share are read once into their XCD's L2.  A work-item forms the
// four staples of its pair at its site:
//   Q_{mu,+nu}(y) = U_nu(y+mu) U_mu(y+nu)^dag U_nu(y)^dag = t U_nu(y)^dag,  t = U_nu(y+mu) U_mu(y+nu)^dag
//   Q_{nu,+mu}(y) = U_mu(y+nu) U_nu(y+mu)^dag U_mu(y)^dag = t^dag U_mu(y)^dag
//   Q_{mu,-nu}(y) = U_nu(y+mu-nu)^dag U_mu(y-nu)^dag U_nu(y-nu)
//   Q_{nu,-mu}(y) = U_mu(y+nu-mu)^dag U_nu(y-mu)^dag U_mu(y-mu)
// (10 link loads and 7 products for 4 staples: 10.5 products per link).
// The staples are stored streaming (nontemporal): they are read back only by
// the next kernel, and kept out of L2 they leave it to the links this kernel
// re-reads.  Measured at 24^4 (FP64): 275 -> 257 us, 655 -> 338 MB fetched per
// launch; the combine kernel after it 1013 -> 989 us.
template <typename T>
__global__ void __launch_bounds__(256) k_gq_staples(const T *u0, const T *u1, T *q0, T *q1, GqGeo g) {
  const int slot = gq_slot();
  const int sb = slot / 12, pp = slot % 12;
  const int il = sb * 256 + threadIdx.x;
  if (il >= g.vh) return;
  const int i = gq_site(g, il);
  const int pair = pp >> 1, cb = pp & 1;
  const int mu = pair < 3 ? 0 : (pair < 5 ? 1 : 2);
  const int nu = pair < 3 ? pair + 1 : (pair < 5 ? pair - 1 : 3);
  int c[4];
  q_coords(g, i, cb, c);
  const int r = c[0] + g.L[0] * (c[1] + g.L[1] * (c[2] + g.L[2] * c[3]));
  const int cmu = q_coord(c, mu), cnu = q_coord(c, nu);
  const int r_p = q_move(g, r, cmu, mu, 1);    // y + mu
  const int r_n = q_move(g, r, cnu, nu, 1);    // y + nu
  const int r_b = q_move(g, r, cnu, nu, -1);   // y - nu
  const int r_m = q_move(g, r, cmu, mu, -1);   // y - mu
  const int r_pb = q_move(g, r_p, cnu, nu, -1);  // y + mu - nu
  const int r_nm = q_move(g, r_n, cmu, mu, -1);  // y + nu - mu
  const T *uo = cb ? u0 : u1;  // links of the other parity
  const T *us = cb ? u1 : u0;  // links of this parity
  T *qs = cb ? q1 : q0;
  M3 a, b, t, w;
  load2(uo, nu, g.vh, r_p >> 1, a);
  load2(uo, mu, g.vh, r_n >> 1, b);
  m_mul<false, true>(a, b, t);
  load2(us, nu, g.vh, i, a);
  m_mul<false, true>(t, a, w);
  store2<T, true>(qs, q_slot(mu, nu, false), g.vh, i, w);
  load2(us, mu, g.vh, i, a);
  m_mul<true, true>(t, a, w);
  store2<T, true>(qs, q_slot(nu, mu, false), g.vh, i, w);
  load2(us, nu, g.vh, r_pb >> 1, a);
  load2(uo, mu, g.vh, r_b >> 1, b);
  m_mul<true, true>(a, b, t);
  load2(uo, nu, g.vh, r_b >> 1, a);
  m_mul<false, false>(t, a, w);
  store2<T, true>(qs, q_slot(mu, nu, true), g.vh, i, w);
  load2(us, mu, g.vh, r_nm >> 1, a);
  load2(uo, nu, g.vh, r_m >> 1, b);
  m_mul<true, true>(a, b, t);
  load2(uo, mu, g.vh, r_m >> 1, a);
  m_mul<false, false>(t, a, w);
  store2<T, true>(qs, q_slot(nu, mu, true), g.vh, i, w);
}

struct GqArgs {
  const void *u12[2];   // two-row links, parity 0 / 1
  const void *q[2];     // staples
  void *h[2];           // momentum
  GqGeo g;
  double cp, cr, cg, eps;
  int nbx;              // site blocks per (mu, parity)
};

// Block -> (site block, mu, parity): the 8 (mu, parity) blocks of a site block
// in adjacent slots of an XCD's slab.
__device__ __forceinline__ void gq_block(int &sb, int &mu, int &cb) {
  const int slot = gq_slot();
  sb = slot >> 3;
  mu = slot & 3;
  cb = (slot >> 2) & 1;
}

// The staple sum S of each work-item lives in LDS, column-wise
// ([element][lane]): it is touched once per term, and with the loops rolled
// the kernel fits 2 wavefronts per SIMD without spilling (fully unrolled, the
// compiler kept every matrix in flight in registers: 256 VGPRs, one
// wavefront; held to 3 wavefronts it spilled).
constexpr int GQ_BLOCK = 256;

__device__ __forceinline__ void s_axpy(double *S, double c, const M3 &a) {
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    S[(2 * e) * GQ_BLOCK] = fma(c, a.re[e], S[(2 * e) * GQ_BLOCK]);
    S[(2 * e + 1) * GQ_BLOCK] = fma(c, a.im[e], S[(2 * e + 1) * GQ_BLOCK]);
  }
}

// S += c op(a) op(b)
template <bool DA, bool DB> __device__ __forceinline__ void s_muladd(double *S, double c, const M3 &a, const M3 &b) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double sr = 0, si = 0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double ar = ere<DA>(a, i, k), ai = eim<DA>(a, i, k);
        const double br = ere<DB>(b, k, j), bi = eim<DB>(b, k, j);
        sr = fma(ar, br, sr);
        sr = fma(-ai, bi, sr);
        si = fma(ar, bi, si);
        si = fma(ai, br, si);
      }
      const int e = 3 * i + j;
      S[(2 * e) * GQ_BLOCK] = fma(c, sr, S[(2 * e) * GQ_BLOCK]);
      S[(2 * e + 1) * GQ_BLOCK] = fma(c, si, S[(2 * e + 1) * GQ_BLOCK]);
    }
}

template <typename T>
__global__ void __launch_bounds__(GQ_BLOCK) k_gq_force(const GqArgs a) {
  int sb, mu, cb;
  gq_block(sb, mu, cb);
  const GqGeo &g = a.g;
  const int il = sb * 256 + threadIdx.x;
  if (il >= g.vh) return;
  const int i = gq_site(g, il);
  const T *Us = static_cast<const T *>(a.u12[cb]), *Uo = static_cast<const T *>(a.u12[1 - cb]);
  const T *Qs = static_cast<const T *>(a.q[cb]), *Qo = static_cast<const T *>(a.q[1 - cb]);
  const int vh = g.vh;
  int c[4];
  q_coords(g, i, cb, c);
  const int r = c[0] + g.L[0] * (c[1] + g.L[1] * (c[2] + g.L[2] * c[3]));
  const int cmu = q_coord(c, mu);
  const int r_pmu = q_move(g, r, cmu, mu, 1), r_mmu = q_move(g, r, cmu, mu, -1);
  const int i_pmu = r_pmu >> 1, i_mmu = r_mmu >> 1;  // parity 1 - cb
  __shared__ double sacc[18 * GQ_BLOCK];
  __shared__ double sgb[18 * GQ_BLOCK];
  double *S = sacc + threadIdx.x, *GB = sgb + threadIdx.x;
#pragma unroll
  for (int e = 0; e < 18; ++e) S[e * GQ_BLOCK] = GB[e * GQ_BLOCK] = 0.0;
  M3 G, A, B;
  // The loops stay rolled: one term's matrices in registers at a time (the
  // latency is hidden by the other wavefronts), and every index is computed
  // from uniform directions, so nothing is indexed per lane.
  // Plaquettes and the rectangles along mu in one pass over n, each Q_{mu,n}(x)
  // loaded once: G = sum_n Q_{mu,n}(x+mu) Q_{mu,n}(x) in registers, the
  // other rectangle sum sum_n Q_{mu,n}(x) Q_{mu,n}(x-mu) in LDS (GB).
  m_zero(G);
#pragma unroll 2
  for (int n = 0; n < 6; ++n) {
    const int qn = q_slot(mu, q_other(mu, n >> 1), n & 1);
    load2(Qs, qn, vh, i, A);          // Q_{mu,n}(x)
    s_axpy(S, a.cp, A);
    load2(Qo, qn, vh, i_pmu, B);      // Q_{mu,n}(x+mu)
    m_muladd<false, false>(B, A, G);
    load2(Qo, qn, vh, i_mmu, B);      // Q_{mu,n}(x-mu)
    s_muladd<false, false>(GB, 1.0, A, B);
  }
  load2(Uo, mu, vh, i_pmu, A);        // U_mu(x+mu)
  s_muladd<false, false>(S, a.cr, A, G);
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    G.re[e] = GB[(2 * e) * GQ_BLOCK];
    G.im[e] = GB[(2 * e + 1) * GQ_BLOCK];
  }
  load2(Uo, mu, vh, i_mmu, A);        // U_mu(x-mu)
  s_muladd<false, false>(S, a.cr, G, A);
  // Per direction nu _|_ mu and sign, the rectangle n, n, -mu, -n, -n and the
  // parallelograms that start with the same link L_n(x+mu) share it:
  //   +nu: U_nu(x+mu) [c_r Q_{mu,+nu}(x+nu) U_nu(x)^dag + c_g sum_d Q_{mu,d}(x+nu) Q_{nu,d}(x)]
  //   -nu: U_nu(x+mu-nu)^dag [c_r Q_{mu,-nu}(x-nu) U_nu(x-nu) + c_g sum_d Q_{mu,d}(x-nu) Q_{nu,d}(x-nu)^dag]
  // (d over the four signed directions orthogonal to mu and nu).
#pragma unroll 1
  for (int k = 0; k < 3; ++k) {
    const int nu = q_other(mu, k), cn = q_coord(c, nu);
    const int i_p = q_move(g, r, cn, nu, 1) >> 1, i_m = q_move(g, r, cn, nu, -1) >> 1;  // parity 1 - cb
    const int i_mo = q_move(g, r_pmu, cn, nu, -1) >> 1;                                 // x+mu-nu, parity cb
    // +nu
    m_zero(G);
#pragma unroll 2
    for (int j = 0; j < 4; ++j) {
      const int dd = q_other(mu, (k + 1 + (j >> 1)) % 3);
      load2(Qo, q_slot(mu, dd, j & 1), vh, i_p, A);
      load2(Qs, q_slot(nu, dd, j & 1), vh, i, B);
      m_muladd<false, false>(A, B, G);
    }
    m_scale(G, a.cg);
    load2(Qo, q_slot(mu, nu, false), vh, i_p, A);
    load2(Us, nu, vh, i, B);
    m_scale(A, a.cr);
    m_muladd<false, true>(A, B, G);
    load2(Uo, nu, vh, i_pmu, A);
    s_muladd<false, false>(S, 1.0, A, G);
    // keeps the -nu loads from being hoisted into the +nu half (without it the
    // kernel takes 256 VGPRs, one wavefront per SIMD; with it 210, two)
    __builtin_amdgcn_sched_barrier(0);
    // -nu
    m_zero(G);
#pragma unroll 2
    for (int j = 0; j < 4; ++j) {
      const int dd = q_other(mu, (k + 1 + (j >> 1)) % 3);
      load2(Qo, q_slot(mu, dd, j & 1), vh, i_m, A);
      load2(Qo, q_slot(nu, dd, j & 1), vh, i_m, B);
      m_muladd<false, true>(A, B, G);
    }
    m_scale(G, a.cg);
    load2(Qo, q_slot(mu, nu, true), vh, i_m, A);
    load2(Uo, nu, vh, i_m, B);
    m_scale(A, a.cr);
    m_muladd<false, false>(A, B, G);
    load2(Us, nu, vh, i_mo, A);
    s_muladd<true, false>(S, 1.0, A, G);
  }
  // W = U_mu(x) S; H <- H + i eps TA(W) (momentum stored as H = -iA, 8 reals)
  load2(Us, mu, vh, i, A);            // U_mu(x), row 2 rebuilt like every other link here
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    B.re[e] = S[(2 * e) * GQ_BLOCK];
    B.im[e] = S[(2 * e + 1) * GQ_BLOCK];
  }
  M3 &W = G;
  m_mul<false, false>(A, B, W);
  const double imtr = (W.im[0] + W.im[4] + W.im[8]) / 3.0;
  // the link's 8 momentum reals (32 or 64 bytes, 16-byte aligned) as 16-byte
  // loads and stores: lanes are 4 links apart, so each access touches one
  // cache line per lane
  T *h = static_cast<T *>(a.h[cb]) + 32 * (size_t)i + 8 * mu;
  using V16 = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
  constexpr int NV = 8 * sizeof(T) / 16;
  V16 *hv = reinterpret_cast<V16 *>(h);
  V16 hr[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) hr[k] = hv[k];
  T *hh = reinterpret_cast<T *>(hr);
  const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
#pragma unroll
  for (int qq = 0; qq < 3; ++qq) {
    const int rr = pr[qq][0], cc = pr[qq][1];
    const double tre = 0.5 * (W.re[3 * rr + cc] - W.re[3 * cc + rr]);
    const double tim = 0.5 * (W.im[3 * rr + cc] + W.im[3 * cc + rr]);
    hh[2 * qq] = (T)((double)hh[2 * qq] - a.eps * tim);
    hh[2 * qq + 1] = (T)((double)hh[2 * qq + 1] + a.eps * tre);
  }
  hh[6] = (T)((double)hh[6] - a.eps * (W.im[0] - imtr));
  hh[7] = (T)((double)hh[7] - a.eps * (W.im[4] - imtr));
#pragma unroll
  for (int k = 0; k < NV; ++k) hv[k] = hr[k];
}

// The staple-product force, queued on the library stream.  False (nothing
// queued, no error recorded) when the field keeps no two-row copy (links not
// SU(3) to rounding) or the staple buffer cannot be allocated: the caller then
// runs the loop-walk kernel.  The staples take 2 x 24 x 6 complex numbers per
// half-site (FP64 at 24^4: 765 MB), held in the second workspace.
template <typename T>
bool gauge_force_staples(GaugeT<T> *G, ForceT<T> *F, const QPHIX_gauge_coeffs_t *c, double eps) {
  if (!gauge_force_staples_mode() || !G->u12[0] || !G->u12[1]) return false;
  const Geometry &gg = ctx().g;
  const size_t per = (size_t)24 * 6 * 2 * (size_t)gg.vh;  // reals per parity
  T *q = (T *)ws2_try_get(sizeof(T) * 2 * per);
  if (!q) return false;
  GqGeo g{};
  g.L[0] = gg.nx; g.L[1] = gg.ny; g.L[2] = gg.nz; g.L[3] = gg.nt;
  g.nxh = gg.nxh;
  g.vh = gg.vh;
  g.stride[0] = 1;
  g.stride[1] = gg.nx;
  g.stride[2] = gg.nx * gg.ny;
  g.stride[3] = gg.nx * gg.ny * gg.nz;
  const int nbx = (gg.vh + 255) / 256;
  // z-region slabs when they tile the grid exactly (QPHIX_GF_ORDER=t: t-slabs).
  // Measured at 24^4 (FP64, both kernels): 1.54 ms with t-slabs, 1.46 ms with
  // z-regions; the combine kernel fetched 4.4 GB per launch with t-slabs
  // against the 765 MB of staples (each XCD's +-1 t-slice window of whole
  // slices, 160 MB per XCD, misses the 256 MB Infinity Cache).  Walking each
  // region's y-rows in chunks (a smaller slice to reuse) measured no better:
  // 1.46 (whole rows) / 1.49 / 1.50 / 1.69 ms for 24 / 12 / 6 / 3-row chunks,
  // 8.5 GB fetched with 6-row chunks against 8.8 GB.
  static const bool tenv = [] {
    const char *e = std::getenv("QPHIX_GF_ORDER");
    return e && e[0] == 't';
  }();
  g.zreg = !tenv && gg.nz % 8 == 0 && gg.vh % 256 == 0 && nbx % 8 == 0 ? 1 : 0;
  hipLaunchKernelGGL(k_gq_staples<T>, dim3((unsigned)(12 * nbx)), dim3(256), 0, ctx().stream, G->u12[0],
                     G->u12[1], q, q + per, g);
  GqArgs a{};
  a.u12[0] = G->u12[0];
  a.u12[1] = G->u12[1];
  a.q[0] = q;
  a.q[1] = q + per;
  a.h[0] = F->h[0];
  a.h[1] = F->h[1];
  a.g = g;
  a.cp = c->plaquette;
  a.cr = c->rectangle;
  a.cg = c->parallelogram;
  a.eps = eps;
  a.nbx = nbx;
  hipLaunchKernelGGL(k_gq_force<T>, dim3((unsigned)(8 * nbx)), dim3(256), 0, ctx().stream, a);
  HIP_CHECK(hipGetLastError());
  return true;
}
template bool gauge_force_staples<double>(GaugeT<double> *, ForceT<double> *, const QPHIX_gauge_coeffs_t *, double);
template bool gauge_force_staples<float>(GaugeT<float> *, ForceT<float> *, const QPHIX_gauge_coeffs_t *, double);

} // namespace qphix
