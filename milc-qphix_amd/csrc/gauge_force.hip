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

// Steps of the 48 loop walks through one link: 6 plaquettes x 3 steps,
// 18 rectangles x 5, 24 parallelograms x 5.
constexpr int GF_MAX_STEPS = 6 * 3 + 18 * 5 + 24 * 5;
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
  // per direction mu, one byte per step (see step_link), four to a word: the
  // kernel reads them with uniform indices, as scalar dword loads
  // (GF_PAD padding steps after the last: the pipeline's final prefetches load them)
  uint32_t tab[4][(GF_MAX_STEPS + GF_PAD + 3) / 4];
};

template <typename T> __device__ __forceinline__ unsigned gf_step(const GFArgs<T> &a, int mu, int k) {
  return (a.tab[mu][k >> 2] >> (8 * (k & 3))) & 0xffu;
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

// Link at the walker's vertex in direction d, of parity p (uniform: every lane
// of the wavefront is at a vertex of the same parity).  dag = false: U_d;
// dag = true: the transpose of U_d (the elements of U_d^dag with their
// imaginary parts not yet negated; the step applies that sign when the link is
// used, so nothing consumes the loaded values early and the loads stay in
// flight).  Each element is one load from a uniform base (scalar) plus the
// lane's site offset.  NE = 6: rows 0-1 only are stored; the elements of row 2
// of U (row 2 of m, or column 2 when dag) are left for recon_row2.
template <int NE, typename T>
__device__ __forceinline__ void load_link_t(const GFArgs<T> &a, const Walker &w, int d, int p, bool dag,
                                            DM3 &m) {
  using T2 = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
  const T2 *base = reinterpret_cast<const T2 *>(a.w0) + (ptrdiff_t)p * a.pdiff +
                   (size_t)(d * NE) * (size_t)a.g.vh;
  const unsigned i = (unsigned)w.r >> 1;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      if (NE == 6 && (dag ? cc : r) == 2) continue;
      const int src = dag ? 3 * cc + r : 3 * r + cc;
      const T2 v = (base + (size_t)src * (size_t)a.g.vh)[i];
      m.re[3 * r + cc] = (double)v.x;
      m.im[3 * r + cc] = (double)v.y;
    }
}

// Row 2 of U from rows 0-1: U2 = conj(U0 x U1), on m as load_link_t<6>
// leaves it (raw elements: m = U, or m = U^T when dag).
template <bool DAG> __device__ __forceinline__ void recon_row2_t(DM3 &m) {
  // element (row a, column b) of U in m (compile-time indices keep m in registers)
  auto at = [](int a, int b) { return DAG ? 3 * b + a : 3 * a + b; };
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
    const double ar = m.re[at(0, c1)], ai = m.im[at(0, c1)], br = m.re[at(1, c2)], bi = m.im[at(1, c2)];
    const double er = m.re[at(0, c2)], ei = m.im[at(0, c2)], fr = m.re[at(1, c1)], fi = m.im[at(1, c1)];
    m.re[at(2, c)] = (ar * br - ai * bi) - (er * fr - ei * fi);
    m.im[at(2, c)] = -((ar * bi + ai * br) - (er * fi + ei * fr));
  }
}
__device__ __forceinline__ void recon_row2(DM3 &m, bool dag) {
  if (dag) recon_row2_t<true>(m);
  else recon_row2_t<false>(m);
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

// One step of a loop walk, as a byte of the per-direction step table:
// bits 0-1 direction d, bit 2 backward, bit 3 first step of a walk (the walk
// starts at x + mu), bit 4 last step (the walk's product is added to the
// staple sum), bits 5-6 index of the walk's coefficient, bit 7 parity of the
// step's link relative to x (the number of unit moves from x to the link's
// site, mod 2).
enum : unsigned { GF_BACK = 4, GF_FIRST = 8, GF_LAST = 16, GF_PODD = 128 };

// Link of one step from the walker's vertex (x's parity cb), as load_link_t
// leaves it; the walker moves to the step's end.
template <int NE, typename T>
__device__ __forceinline__ void step_link(const GFArgs<T> &a, Walker &w, unsigned e, int cb, DM3 &m) {
  const int d = e & 3, p = cb ^ ((e & GF_PODD) ? 1 : 0);
  if (e & GF_BACK) {
    walker_move(w, d, -1, a.g);
    load_link_t<NE>(a, w, d, p, true, m);
  } else {
    load_link_t<NE>(a, w, d, p, false, m);
    walker_move(w, d, 1, a.g);
  }
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
  // every rest path starts at x + mu
  Walker y{{x[0], x[1], x[2], x[3]}, x[0] + a.g.L[0] * (x[1] + a.g.L[1] * (x[2] + a.g.L[2] * x[3]))};
  walker_move(y, mu, 1, a.g);
  __shared__ double sacc[18 * GF_BLOCK];
  double *S = sacc + threadIdx.x;
#pragma unroll
  for (int e = 0; e < 18; ++e) S[e * GF_BLOCK] = 0;
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
    step_link<NE>(a, w, e, cb, buf);
    return e;
  };
  // apply a step's link (table byte e) to the walk's running product r
  auto apply = [&](unsigned e, DM3 &r, DM3 &cur) {
    if constexpr (NE == 6) recon_row2(cur, (e & GF_BACK) != 0);
    if (e & GF_FIRST) {
      const double sg = (e & GF_BACK) ? -1.0 : 1.0;
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        r.re[q] = cur.re[q];
        r.im[q] = sg * cur.im[q];
      }
    } else {
      if (e & GF_BACK) {  // uniform branch: conjugate the link in place
#pragma unroll
        for (int q = 0; q < 9; ++q) cur.im[q] = -cur.im[q];
      }
      dm_rmul_t<false>(r, cur);
    }
    if (e & GF_LAST) {
      const double c = a.c[(e >> 5) & 3];
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        S[(2 * q) * GF_BLOCK] += c * r.re[q];
        S[(2 * q + 1) * GF_BLOCK] += c * r.im[q];
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
// back to x; for the link's own direction m = mu + 1 and another direction n:
//   plaquette   n, -m, -n
//   rectangles  m, n, -m, -m, -n  /  n, -m, -m, -n, m  /  n, n, -m, -n, -n
// for n = +-(nu+1), and for two other directions A, B (ordered, signed):
//   parallelogram  A, B, -m, -A, -B.
// Walks whose coefficient is zero are left out.
template <typename T> static bool gf_step_tables(GFArgs<T> &a) {
  std::memset(a.tab, 0, sizeof(a.tab));
  int n = 0;
  for (int mu = 0; mu < 4; ++mu) {
    int k = 0;
    auto walk = [&](std::initializer_list<int> steps, unsigned ci) {
      const int len = (int)steps.size();
      int j = 0;
      for (int st : steps) {
        unsigned e = (unsigned)((st > 0 ? st : -st) - 1) | (ci << 5);
        if (st < 0) e |= GF_BACK;
        // unit moves from x to the link's site: 1 (to x + mu) + j, + 1 more
        // for a backward step (its link sits at the step's end)
        if ((1 + j + (st < 0 ? 1 : 0)) & 1) e |= GF_PODD;
        if (j == 0) e |= GF_FIRST;
        if (j == len - 1) e |= GF_LAST;
        a.tab[mu][k >> 2] |= e << (8 * (k & 3));
        ++k;
        ++j;
      }
    };
    const int m = mu + 1;
    for (int nu = 0; nu < 4; ++nu) {
      if (nu == mu) continue;
      for (int sg = -1; sg <= 1; sg += 2) {
        const int d = sg * (nu + 1);
        if (a.c[0] != 0) walk({d, -m, -d}, 0);
        if (a.c[1] != 0) {
          walk({m, d, -m, -m, -d}, 1);
          walk({d, -m, -m, -d, m}, 1);
          walk({d, d, -m, -d, -d}, 1);
        }
      }
    }
    if (a.c[2] != 0) {
      for (int da = 0; da < 4; ++da) {
        if (da == mu) continue;
        for (int db = 0; db < 4; ++db) {
          if (db == mu || db == da) continue;
          for (int sa = -1; sa <= 1; sa += 2)
            for (int sb = -1; sb <= 1; sb += 2) {
              const int A = sa * (da + 1), B = sb * (db + 1);
              walk({A, B, -m, -A, -B}, 2);
            }
        }
      }
    }
    // padding steps: forward steps from x + mu (one move from x) along
    // direction 0, loaded by the last prefetches and never applied
    for (int j = 0; j < GF_PAD; ++j)
      a.tab[mu][(k + j) >> 2] |= (GF_FIRST | GF_PODD) << (8 * ((k + j) & 3));
    n = k;
  }
  QPHIX_REQUIRE(n % 6 == 0 && n <= GF_MAX_STEPS, "gauge force: bad step table", false);
  a.nsteps = n;
  return true;
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
