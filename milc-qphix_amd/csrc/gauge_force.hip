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
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <type_traits>

namespace qphix {

struct GFGeo {
  int L[4];
  int nxh, vh;
};

// Steps of the 48 loop walks through one link (6 plaquettes, 18 rectangles,
// 24 parallelograms), grouped by their last link (see gf_step_tables): 6
// groups of 7 walks with 2 + 4 + 4 + 4 x 4 steps and a closing step, 1 group
// of 6 walks with 6 x 4 steps and a closing step, 1 padding step (even count).
constexpr int GF_MAX_STEPS = 6 * (2 + 4 + 4 + 16 + 1) + (6 * 4 + 1) + 1;
// Padding steps = the largest prefetch distance D compiled (distance 2
// measured the same as 1 at 24^4, with 17 more registers).
constexpr int GF_PAD = 2;

template <typename T> struct GFArgs {
  const T *u[2];    // all 3 rows (the link of the force's own site)
  const T *w0;      // links the walks read, parity 0: u[0], or the two-row copy
  ptrdiff_t pdiff;  // parity 1 - parity 0 of those, in units of 2 T
  T *h[2];
  GFGeo g;
  double c[3];
  double eps;
  int nsteps;       // steps per direction (walks of zero coefficient left out)
  int order;        // block order, see gf_block
  // per direction mu, 16 bits per step (see step_link), two to a word: the
  // kernel reads them with uniform indices, as scalar dword loads
  // (GF_PAD padding steps after the last: the pipeline's final prefetches load them)
  uint32_t tab[4][(GF_MAX_STEPS + GF_PAD + 1) / 2];
};

template <typename T> __device__ __forceinline__ unsigned gf_step(const GFArgs<T> &a, int mu, int k) {
  return (a.tab[mu][k >> 1] >> (16 * (k & 1))) & 0xffffu;
}

struct DM3 {
  double re[9], im[9];
};

// Work-items per block of the force kernel.  Each keeps its staple sum
// (18 doubles) in LDS, column-wise ([element][lane]), instead of registers: it
// is touched once per loop walk, and the 36 registers it frees raise occupancy.
constexpr int GF_BLOCK = 256;

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

// A walk's current vertex: wrapped coordinates and lexicographic index
// r = x0 + L0 (x1 + L1 (x2 + L2 x3)) (the half-site index is r >> 1, L0 even).
struct Walker {
  int c[4];
  int r;
};

// Moves w one site along direction D, forward (sgn = 1) or backward (-1);
// only the moved coordinate is wrapped.
template <int D> __device__ __forceinline__ void walker_move(Walker &w, int sgn, const GFGeo &g) {
  const int L = g.L[D];
  const int stride = D == 0 ? 1 : D == 1 ? g.L[0] : D == 2 ? g.L[0] * g.L[1] : g.L[0] * g.L[1] * g.L[2];
  int v = w.c[D] + sgn, r = w.r + sgn * stride;
  if (sgn > 0) {  // uniform
    const bool wrap = v == L;
    v = wrap ? 0 : v;
    r = wrap ? r - L * stride : r;
  } else {
    const bool wrap = v < 0;
    v = wrap ? L - 1 : v;
    r = wrap ? r + L * stride : r;
  }
  w.c[D] = v;
  w.r = r;
}
__device__ __forceinline__ void walker_move(Walker &w, int d, int sgn, const GFGeo &g) {
  switch (d) {  // d is uniform: one branch, no indexed registers
    case 0: walker_move<0>(w, sgn, g); break;
    case 1: walker_move<1>(w, sgn, g); break;
    case 2: walker_move<2>(w, sgn, g); break;
    default: walker_move<3>(w, sgn, g); break;
  }
}

// Link U_d at the walker's vertex, of parity p (uniform: every lane of the
// wavefront is at a vertex of the same parity), elements in storage order
// (m.re/im[3 row + column]); NE = 6: rows 0-1 only (row 2 is rebuilt by
// recon_row2 when the link is used).  Each element is one load from a uniform
// (scalar) base plus the lane's site offset.  Nothing here depends on whether
// the step uses the link daggered: the loads are the same either way, the
// values are not touched until the step applies them (so the loads stay in
// flight), and the registers they land in are fixed (a dag-dependent placement
// was compiled into indexed private memory).
template <int NE, typename T>
__device__ __forceinline__ void load_link_t(const GFArgs<T> &a, const Walker &w, int d, int p, DM3 &m) {
  using T2 = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
  const T2 *base = reinterpret_cast<const T2 *>(a.w0) + (ptrdiff_t)p * a.pdiff +
                   (size_t)(d * NE) * (size_t)a.g.vh;
  const unsigned i = (unsigned)w.r >> 1;
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const T2 v = (base + (size_t)e * (size_t)a.g.vh)[i];
    m.re[e] = (double)v.x;
    m.im[e] = (double)v.y;
  }
}

// Row 2 of U from rows 0-1: U2 = conj(U0 x U1).
__device__ __forceinline__ void recon_row2(DM3 &m) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
    const double ar = m.re[c1], ai = m.im[c1], br = m.re[3 + c2], bi = m.im[3 + c2];
    const double er = m.re[c2], ei = m.im[c2], fr = m.re[3 + c1], fi = m.im[3 + c1];
    m.re[6 + c] = (ar * br - ai * bi) - (er * fr - ei * fi);
    m.im[6 + c] = -((ar * bi + ai * br) - (er * fi + ei * fr));
  }
}

// m <- m^dag in place.
__device__ __forceinline__ void dm_dagger(DM3 &m) {
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    m.im[4 * r] = -m.im[4 * r];
#pragma unroll
    for (int c = r + 1; c < 3; ++c) {
      const double tr = m.re[3 * r + c], ti = m.im[3 * r + c];
      m.re[3 * r + c] = m.re[3 * c + r];
      m.im[3 * r + c] = -m.im[3 * c + r];
      m.re[3 * c + r] = tr;
      m.im[3 * c + r] = -ti;
    }
  }
}

// r <- r * m', m' = m (CONJ false) or m with its imaginary parts negated
// (CONJ true), one row at a time (one row of temporaries instead of a third
// matrix).
template <bool CONJ> __device__ __forceinline__ void dm_rmul_t(DM3 &r, const DM3 &m) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double tr[3], ti[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double sr = 0, si = 0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double mi = CONJ ? -m.im[3 * k + j] : m.im[3 * k + j];  // folds into the fma signs
        sr = fma(r.re[3 * i + k], m.re[3 * k + j], sr);
        sr = fma(-r.im[3 * i + k], mi, sr);
        si = fma(r.re[3 * i + k], mi, si);
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

// One step, as 16 bits of the per-direction step table: bits 0-1 direction
// d; GF_BACK: the walker moves back along d before the load (else forward
// after it); GF_DAG: the link is used daggered; GF_FIRST: first step of a walk
// (the walker starts at x + mu; the product restarts at this link); GF_ATX: the
// walker starts at x; GF_LAST: the walk's product, times its coefficient, is
// added to the staple sum S; GF_GACC: ... added to the group sum G instead;
// GF_GFIN: closing step of a group: S += G * link, G = 0; bits 5-6 the walk's
// coefficient; GF_PODD: parity of the link's site relative to x (unit moves
// from x, mod 2).  A backward step of a walk is GF_BACK | GF_DAG.
enum : unsigned {
  GF_BACK = 4, GF_FIRST = 8, GF_LAST = 16, GF_PODD = 128,
  GF_GACC = 256, GF_GFIN = 512, GF_ATX = 1024, GF_DAG = 2048
};

// Link of one step from the walker's vertex (x's parity cb), as load_link_t
// leaves it; the walker moves to the step's end.
template <int NE, typename T>
__device__ __forceinline__ void step_link(const GFArgs<T> &a, Walker &w, unsigned e, int cb, DM3 &m) {
  const int d = e & 3, p = cb ^ ((e & GF_PODD) ? 1 : 0);
  if (e & GF_BACK) walker_move(w, d, -1, a.g);
  load_link_t<NE>(a, w, d, p, m);
  if (!(e & GF_BACK)) walker_move(w, d, 1, a.g);
}

// The force's own link U_mu(x), all 3 rows, x at index i of parity cb.
template <typename T>
__device__ __forceinline__ void own_link(const GFArgs<T> &a, int cb, int mu, int i, DM3 &m) {
  using T2 = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
  const T2 *s = reinterpret_cast<const T2 *>(cb ? a.u[1] : a.u[0]);  // cb is uniform
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    const T2 v = s[gauge_elem(mu, e, i, a.g.vh)];
    m.re[e] = (double)v.x;
    m.im[e] = (double)v.y;
  }
}

// Grid: one block per (256 sites of one parity, direction mu, parity).  All
// lanes of a wavefront walk the same loops from consecutive sites, so every
// link load of a step is coalesced.  Blocks are renumbered so that XCD x
// sweeps the x-th contiguous slab of the work, keeping the paths' link loads,
// up to 2 sites away, in that XCD's L2.  order 0: the slab runs over the sites
// of one (mu, parity) at a time; order 1: over site blocks, with the 8
// (mu, parity) blocks of each site block next to each other (they read the
// same neighbourhood links).
struct GFBlock {
  int sb, mu, cb;
};
__device__ __forceinline__ GFBlock gf_block(int order, int nbx) {
  const int nb = gridDim.x, b = blockIdx.x;
  if (order == 1) {
    const int slot = nb % 8 == 0 ? (b % 8) * (nb / 8) + b / 8 : b;
    return {slot >> 3, slot & 3, (slot >> 2) & 1};
  }
  const int bx = b % nbx, mu = (b / nbx) & 3, cb = b / (4 * nbx);
  return {nbx % 8 == 0 ? (bx % 8) * (nbx / 8) + bx / 8 : bx, mu, cb};
}

// Occupancy target (waves per SIMD) for the force kernel; unset, the compiler
// keeps every walk's registers (2 waves per SIMD).
#ifdef QPHIX_GF_WAVES_PER_EU
#define QPHIX_GF_OCC __attribute__((amdgpu_waves_per_eu(QPHIX_GF_WAVES_PER_EU)))
#else
#define QPHIX_GF_OCC
#endif

template <typename T, int NE, int D>
__global__ void __launch_bounds__(GF_BLOCK) QPHIX_GF_OCC k_gauge_force(const GFArgs<T> a) {
  const GFBlock bk = gf_block(a.order, (a.g.vh + GF_BLOCK - 1) / GF_BLOCK);
  const int i = bk.sb * GF_BLOCK + threadIdx.x;
  if (i >= a.g.vh) return;
  const int mu = bk.mu, cb = bk.cb;
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
  // the walks start at x + mu (group closing steps at x)
  const Walker xw{{x[0], x[1], x[2], x[3]}, x[0] + a.g.L[0] * (x[1] + a.g.L[1] * (x[2] + a.g.L[2] * x[3]))};
  Walker y = xw;
  walker_move(y, mu, 1, a.g);
  // staple sum S and group sum G, per lane in LDS ([element][lane])
  __shared__ double sacc[36 * GF_BLOCK];
  double *S = sacc + threadIdx.x, *G = sacc + 18 * GF_BLOCK + threadIdx.x;
#pragma unroll
  for (int e = 0; e < 36; ++e) S[e * GF_BLOCK] = 0;
  // All 48 loop walks through (x, mu) as one software-pipelined stream of
  // steps: the link of step k + D (of this walk or a later one) is requested
  // before the product with step k's link, so each load's latency overlaps D
  // 3x3 products.  D + 2 matrices are live (product and D + 1 link buffers).
  // The step bytes are uniform over the block, so the branches on them are
  // uniform.
  const int n = a.nsteps;
  // request the link of step k into buf (the walker moves with the requests);
  // always a load, also for the padding steps after the last: a conditional
  // one leaves the number of loads in flight unknown, and the compiler then
  // waits for all of them
  auto fetch = [&](int k, Walker &w, DM3 &buf) -> unsigned {
    const unsigned e = gf_step(a, mu, k);
    if (e & GF_FIRST) w = y;
    if (e & GF_ATX) w = xw;
    step_link<NE>(a, w, e, cb, buf);
    return e;
  };
  // apply a step's link (table byte e) to the walk's running product r
  auto apply = [&](unsigned e, DM3 &r, DM3 &cur) {
    if constexpr (NE == 6) recon_row2(cur);
    if (e & GF_DAG) dm_dagger(cur);  // uniform branch
    if (e & GF_FIRST) {
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        r.re[q] = cur.re[q];
        r.im[q] = cur.im[q];
      }
    } else {
      if (e & GF_GFIN) {  // the group's shared last link: r = G (then G = 0)
#pragma unroll
        for (int q = 0; q < 9; ++q) {
          r.re[q] = G[(2 * q) * GF_BLOCK];
          r.im[q] = G[(2 * q + 1) * GF_BLOCK];
          G[(2 * q) * GF_BLOCK] = 0;
          G[(2 * q + 1) * GF_BLOCK] = 0;
        }
      }
      dm_rmul_t<false>(r, cur);
    }
    if (e & (GF_LAST | GF_GACC | GF_GFIN)) {
      const double c = (e & GF_GFIN) ? 1.0 : a.c[(e >> 5) & 3];
      double *acc = (e & GF_GACC) ? G : S;
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        acc[(2 * q) * GF_BLOCK] += c * r.re[q];
        acc[(2 * q + 1) * GF_BLOCK] += c * r.im[q];
      }
    }
  };
  if (n > 0) {
    Walker w = y;
    DM3 r, buf[D + 1];
    unsigned es[D + 1];
#pragma unroll
    for (int j = 0; j < D; ++j) es[j] = fetch(j, w, buf[j]);
    // D + 1 steps per trip, unrolled, with the buffers rotating (compile-time
    // indices, no copies: a copy would wait for the link in flight); n is a
    // multiple of 6 (every walk type contributes one)
#pragma unroll 1
    for (int k = 0; k < n; k += D + 1) {
#pragma unroll
      for (int j = 0; j <= D; ++j) {
        const int jn = (j + D) % (D + 1);
        es[jn] = fetch(k + j + D, w, buf[jn]);
        apply(es[j], r, buf[j]);
      }
    }
  }
  DM3 U, W, Ssum;
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    Ssum.re[e] = S[(2 * e) * GF_BLOCK];
    Ssum.im[e] = S[(2 * e + 1) * GF_BLOCK];
  }
  own_link(a, cb, mu, i, U);
  dm_mul(U, Ssum, W);
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

// Step tables of the loop walks through (x, mu), for each mu.  A walk is a
// list of signed steps (+(d+1) forward along d, -(d+1) backward) from x + mu
// back to x; for the link's own direction m = mu + 1 and another direction
// n = +-(nu+1):
//   plaquette   n, -m, -n
//   rectangles  m, n, -m, -m, -n  /  n, -m, -m, -n, m  /  n, n, -m, -n, -n
// and for two other directions A, B (ordered, signed):
//   parallelogram  A, B, -m, -A, -B.
// Walks that end with the same link share its product: for each n, the 7
// walks ending with -n (the plaquette, the first and third rectangles, the 4
// parallelograms with B = n) add their products without that last step,
// times their coefficients, to G, and a closing step adds G times the shared
// link (U_n(x)^dag for n > 0, U_|n|(x - |n|) for n < 0) to S; likewise the 6
// second rectangles, which end with +m (link U_m(x - m)).  180 link products
// become 139.  Walks of a zero coefficient are left out, and so is the
// closing step of an empty group.
template <typename T> static bool gf_step_tables(GFArgs<T> &a) {
  std::memset(a.tab, 0, sizeof(a.tab));
  int n = -1;
  for (int mu = 0; mu < 4; ++mu) {
    int k = 0;
    auto put = [&](unsigned e) {
      a.tab[mu][k >> 1] |= e << (16 * (k & 1));
      ++k;
    };
    // a walk without its last step; the product of the kept steps goes to G
    auto walk = [&](std::initializer_list<int> steps, unsigned ci) {
      const int kept = (int)steps.size() - 1;
      int j = 0;
      for (int st : steps) {
        if (j == kept) break;
        unsigned e = (unsigned)((st > 0 ? st : -st) - 1) | (ci << 5);
        if (st < 0) e |= GF_BACK | GF_DAG;
        // unit moves from x to the link's site: 1 (to x + mu) + j, + 1 more
        // for a backward step (its link sits at the step's end)
        if ((1 + j + (st < 0 ? 1 : 0)) & 1) e |= GF_PODD;
        if (j == 0) e |= GF_FIRST;
        if (j == kept - 1) e |= GF_GACC;
        put(e);
        ++j;
      }
    };
    // closing step of a group whose walks end with step `last` into x
    auto close = [&](int last) {
      unsigned e = (unsigned)((last > 0 ? last : -last) - 1) | GF_ATX | GF_GFIN;
      if (last > 0) e |= GF_BACK | GF_PODD;  // link at x - d (one move from x), undaggered
      else e |= GF_DAG;                       // link at x, daggered
      put(e);
    };
    const int m = mu + 1;
    for (int nu = 0; nu < 4; ++nu) {
      if (nu == mu) continue;
      for (int sg = -1; sg <= 1; sg += 2) {
        const int d = sg * (nu + 1);
        const int k0 = k;
        if (a.c[0] != 0) walk({d, -m, -d}, 0);
        if (a.c[1] != 0) {
          walk({m, d, -m, -m, -d}, 1);
          walk({d, d, -m, -d, -d}, 1);
        }
        if (a.c[2] != 0)
          for (int da = 0; da < 4; ++da) {
            if (da == mu || da == nu) continue;
            for (int sa = -1; sa <= 1; sa += 2) {
              const int A = sa * (da + 1);
              walk({A, d, -m, -A, -d}, 2);
            }
          }
        if (k > k0) close(-d);
      }
    }
    if (a.c[1] != 0) {
      for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        for (int sg = -1; sg <= 1; sg += 2) {
          const int d = sg * (nu + 1);
          walk({d, -m, -m, -d, m}, 1);
        }
      }
      close(m);
    }
    // an even count (two steps per pipeline trip): a one-step walk that adds
    // nothing; then the padding steps the last prefetches load, never applied
    if (k & 1) put(GF_FIRST | GF_PODD);
    QPHIX_REQUIRE(n < 0 || n == k, "gauge force: step tables differ between directions", false);
    n = k;
    for (int j = 0; j < GF_PAD; ++j) put(GF_FIRST | GF_PODD);
  }
  QPHIX_REQUIRE(n % 2 == 0 && n <= GF_MAX_STEPS, "gauge force: bad step table", false);
  a.nsteps = n;
  return true;
}

// Queues the force kernel (no synchronisation); info->status is set to
// QPHIX_SUCCESS once it is queued.
template <typename T>
void gauge_force_apply(QPHIX_info_t *info, GaugeT<T> *G, ForceT<T> *F, QPHIX_gauge_coeffs_t *c, double eps) {
  info->status = QPHIX_FAIL;
  using namespace std::chrono;
  const double t0 = duration<double>(steady_clock::now().time_since_epoch()).count();
  const Geometry &gg = ctx().g;
  GFArgs<T> a;
  a.u[0] = G->u[0];
  a.u[1] = G->u[1];
  a.h[0] = F->h[0];
  a.h[1] = F->h[1];
  a.g = GFGeo{{gg.nx, gg.ny, gg.nz, gg.nt}, gg.nxh, gg.vh};
  // the walks read the two-row copy when the field has one
  const bool rows2 = G->u12[0] && G->u12[1];
  const T *w[2] = {rows2 ? G->u12[0] : G->u[0], rows2 ? G->u12[1] : G->u[1]};
  a.w0 = w[0];
  a.pdiff = reinterpret_cast<const char *>(w[1]) - reinterpret_cast<const char *>(w[0]);
  QPHIX_REQUIRE(a.pdiff % (ptrdiff_t)(2 * sizeof(T)) == 0, "gauge force: misaligned parity arrays");
  a.pdiff /= (ptrdiff_t)(2 * sizeof(T));
  a.c[0] = c->plaquette;
  a.c[1] = c->rectangle;
  a.c[2] = c->parallelogram;
  a.eps = eps;
  if (gauge_force_staples<T>(G, F, c, eps)) {
    info->status = QPHIX_SUCCESS;
    info->final_sec = duration<double>(steady_clock::now().time_since_epoch()).count() - t0;
    info->final_flop = 61092.0;  // the reference's constant, gauge_force_imp.cpp:197
    return;
  }
  if (!gf_step_tables(a)) return;
  static const int order = [] {
    const char *e = std::getenv("QPHIX_GF_ORDER");
    return e && std::atoi(e) == 0 ? 0 : 1;
  }();
  a.order = order;
  const dim3 grid((unsigned)((gg.vh + GF_BLOCK - 1) / GF_BLOCK * 8));
  if (rows2)
    hipLaunchKernelGGL((k_gauge_force<T, 6, 1>), grid, dim3(GF_BLOCK), 0, ctx().stream, a);
  else
    hipLaunchKernelGGL((k_gauge_force<T, 9, 1>), grid, dim3(GF_BLOCK), 0, ctx().stream, a);
  HIP_CHECK(hipGetLastError());
  info->status = QPHIX_SUCCESS;
  info->final_sec = duration<double>(steady_clock::now().time_since_epoch()).count() - t0;
  info->final_flop = 61092.0;  // the reference's constant, gauge_force_imp.cpp:197
}
template void gauge_force_apply<double>(QPHIX_info_t *, GaugeT<double> *, ForceT<double> *,
                                        QPHIX_gauge_coeffs_t *, double);
template void gauge_force_apply<float>(QPHIX_info_t *, GaugeT<float> *, ForceT<float> *,
                                       QPHIX_gauge_coeffs_t *, double);

template <typename T>
static void gauge_force(QPHIX_info_t *info, GaugeT<T> *G, ForceT<T> *F, QPHIX_gauge_coeffs_t *c, double eps) {
  using namespace std::chrono;
  const double t0 = duration<double>(steady_clock::now().time_since_epoch()).count();
  gauge_force_apply<T>(info, G, F, c, eps);
  HIP_CHECK(hipStreamSynchronize(ctx().stream));
  info->final_sec = duration<double>(steady_clock::now().time_since_epoch()).count() - t0;
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
