// dslash_mrhs.h -- Dslash applied to NR right-hand sides in one launch.
//
// Links are read from HBM once per site and applied to NR spinors, so link
// traffic (96% of a single-RHS Dslash's bytes) is shared: bytes per site become
// links + NR x 2 x spinor instead of NR x (links + 2 x spinor).  Spinors of
// the NR right-hand sides are interleaved per site (index site * NR + rhs,
// FP32 or FP64, stored by Spin<T> tile-major over 64 such indices), so one
// neighbour gather fetches a site's sources from one contiguous stretch per
// colour.  Each RHS carries its own gate (its solver state's `active`);
// inactive RHS are not stored.
//
// Thread mapping: a work-item handles K of the NR sources of one site, so
// G = NR / K lanes share a site.  Links reach each lane's registers once per
// K sources: L1-to-register traffic per (site, source) is links / K + the
// gathered spinors, while registers grow with K (K accumulators and K
// gathers in flight).  K = 1 is one lane per (site, source); the autotuner
// picks NR and K per lattice.
#pragma once
#include "dslash_kernel.h"

namespace qphix {

template <typename T> struct MrhsArgsT {
  const typename vec16<T>::type *fat, *lng;
  const uint8_t *lsign;
  const T *in;       // [vh][NR][6]
  T *out;            // [vh][NR][6]
  const T *aux;      // EPI_PKP: p, [vh][NR][6]
  Geo g;
  int cb;
  T s;               // EPI_PKP: out = D in + s * p
  double *partials;  // EPI_PKP: [NR][gridDim.x]
  const SolverState *states;  // per-RHS gates (states[r].active)
  int nlog;          // logical 256-item blocks (vh * NR / K / 256, rounded up)
  SiteOrder order;   // traversal order of the sites
};

template <int NR> __device__ __forceinline__ unsigned rhs_mask(const SolverState *st) {
  unsigned m = 0;
#pragma unroll
  for (int r = 0; r < NR; ++r) m |= (st[r].active ? 1u : 0u) << r;
  return m;
}

// The grid is resident-sized (a multiple of N_XCD): the nlog logical 256-item
// blocks are split into N_XCD contiguous slabs, XCD x = blockIdx % N_XCD walks
// slab x with its gridDim.x / N_XCD blocks (neighbour gathers stay in that
// XCD's L2), and reducing epilogues leave one partial per resident block.
struct XcdWalk {
  int q, per, span, base, nlog;
  __device__ XcdWalk(int nlog_) : nlog(nlog_) {
    const int x = blockIdx.x % N_XCD;
    per = gridDim.x / N_XCD;
    q = blockIdx.x / N_XCD;
    span = (nlog + N_XCD - 1) / N_XCD;
    base = x * span;
  }
  // logical block of step k, or -1 when this block's walk is over (uniform per block)
  __device__ int at(int k) const {
    const int qq = q + k * per;
    if (qq >= span) return -1;
    const int blk = base + qq;
    return blk < nlog ? blk : -1;
  }
};

// One direction's link pair u (forward link, then the back link; NC = 18, or
// 12 = two rows each) applied to K sources (r0 .. r0+K-1) of the neighbours nf
// (forward) and nb (back).
template <typename T, int NC, int NR, int K>
__device__ __forceinline__ void apply_links_k(cplx<T> *u, const T *in, int nf, int nb, int r0, unsigned sg_f,
                                              unsigned sg_b, cplx<T> (*acc)[3]) {
  cplx<T> vf[K][3], vb[K][3];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    Spin<T>::load(in, nf * NR + r0 + k, vf[k]);
    Spin<T>::load(in, nb * NR + r0 + k, vb[k]);
  }
  if constexpr (NC == 18) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      matvec_acc(u, vf[k], acc[k]);
      matvec_acc(u + 9, vb[k], acc[k]);
    }
  } else {
    cplx<T> w[9];
#pragma unroll
    for (int e = 0; e < 6; ++e) w[e] = u[e];
    recon_row2(w, sg_f ? (T)-1 : (T)1);
#pragma unroll
    for (int k = 0; k < K; ++k) matvec_acc(w, vf[k], acc[k]);
#pragma unroll
    for (int e = 0; e < 6; ++e) w[e] = u[6 + e];
    recon_row2(w, sg_b ? (T)-1 : (T)1);
#pragma unroll
    for (int k = 0; k < K; ++k) matvec_acc(w, vb[k], acc[k]);
  }
}

// The same with the link pair loaded from the link array (link_chunk layout;
// base = link_base + direction offset, chunks LINK_CSTRIDE apart).
template <typename T, int NC, int NR, int K>
__device__ __forceinline__ void apply_pair_k(const typename vec16<T>::type *base, int ii, const T *in,
                                             int nf, int nb, int r0, unsigned sg_f, unsigned sg_b,
                                             cplx<T> (*acc)[3]) {
  cplx<T> u[NC];
  PairLoad<T, NC, true>::load(base, LINK_CSTRIDE, ii, u);
  apply_links_k<T, NC, NR, K>(u, in, nf, nb, r0, sg_f, sg_b, acc);
}

// Per-source sums over the block: lane l holds K values of source group
// l % G (lanes l and l + m*G share a group).  Thread q < G*K returns the sum
// for source q = group*K + k.
template <int G, int K, int BLOCK> __device__ __forceinline__ double group_block_sum(double (&v)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int off = 32; off >= G; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  __shared__ double sh[BLOCK / 64][G * K];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane < G) {
#pragma unroll
    for (int k = 0; k < K; ++k) sh[w][lane * K + k] = v[k];
  }
  __syncthreads();
  double sum = 0.0;
  if (threadIdx.x < G * K) {
#pragma unroll
    for (int m = 0; m < BLOCK / 64; ++m) sum += sh[m][threadIdx.x];
  }
  return sum;
}

template <typename T, int NL, int NR, int K, int EPI>
__global__ void __launch_bounds__(DSLASH_BLOCK) k_dslash_mrhs(const MrhsArgsT<T> a) {
  static_assert(NR % K == 0, "K must divide NR");
  constexpr int G = NR / K;
  constexpr int CPV = vec16<T>::ncplx;
  constexpr int NCH_F = 18 / CPV, NCH_L = NL / CPV;
  const unsigned mask = rhs_mask<NR>(a.states);
  if (mask == 0) return;  // uniform over the grid
  const int vh = a.g.vh;
  const XcdWalk wk(a.nlog);
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.0;
  for (int step = 0;; ++step) {
    const int blk = wk.at(step);
    if (blk < 0) break;
    const int t = blk * DSLASH_BLOCK + threadIdx.x;
    const int L = t / G, r0 = (t % G) * K;
    const unsigned m = (mask >> r0) & ((1u << K) - 1u);
    if (L >= vh || m == 0) continue;
    const int i = ordered_site(a.g, a.order, L);
    cplx<T> acc[K][3];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[k][c] = cplx<T>{0, 0};
    SiteNbr n;
    site_neighbours(a.g, i, a.cb, n);
    unsigned sg = 0;
    if constexpr (NL == 12) sg = a.lsign[i];
    const size_t fb = link_base(i, NCH_F), lb = link_base(i, NCH_L);
    const int ii = i % LINK_TILE;
    // With K > 1 an unrolled direction loop lets the scheduler hoist all 16 x K
    // gathers and overflow the register file; one direction at a time keeps
    // 2 link pairs + 4 K gathers in flight.
    constexpr int MU_UNROLL = K == 1 ? 4 : 1;
#pragma unroll MU_UNROLL
    for (int mu = 0; mu < 4; ++mu) {
      apply_pair_k<T, 18, NR, K>(a.fat + fb + (size_t)mu * NCH_F * LINK_CSTRIDE, ii, a.in, n.f1[mu],
                                 n.b1[mu], r0, 0, 0, acc);
      apply_pair_k<T, NL, NR, K>(a.lng + lb + (size_t)mu * NCH_L * LINK_CSTRIDE, ii, a.in, n.f3[mu],
                                 n.b3[mu], r0, (sg >> (2 * mu + 1)) & 1u, (sg >> (2 * mu)) & 1u, acc);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (!((m >> k) & 1u)) continue;
      const int slot = i * NR + r0 + k;
      if constexpr (EPI == EPI_PKP) {
        cplx<T> p[3];
        Spin<T>::load(a.aux, slot, p);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          acc[k][c].re = fma(a.s, p[c].re, acc[k][c].re);
          acc[k][c].im = fma(a.s, p[c].im, acc[k][c].im);
          v[k] = fma((double)p[c].re, (double)acc[k][c].re, v[k]);
          v[k] = fma((double)p[c].im, (double)acc[k][c].im, v[k]);
        }
      }
      Spin<T>::store(a.out, slot, acc[k]);
    }
  }
  if constexpr (EPI == EPI_PKP) {
    const double sum = group_block_sum<G, K, DSLASH_BLOCK>(v);
    if (threadIdx.x < NR) a.partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = sum;
  }
}

// Per-source block sums for one-work-item-per-(site, source) kernels (K = 1).
template <int NR, int BLOCK> __device__ __forceinline__ double lane_block_sum(double v) {
  double w[1] = {v};
  return group_block_sum<NR, 1, BLOCK>(w);
}

// One 16-byte link chunk, streamed (nontemporal) load.
__device__ __forceinline__ float4 nt_chunk(const float4 *p) {
  const nf4 d = __builtin_nontemporal_load(reinterpret_cast<const nf4 *>(p));
  return make_float4(d.x, d.y, d.z, d.w);
}
__device__ __forceinline__ double2 nt_chunk(const double2 *p) {
  const nd2 d = __builtin_nontemporal_load(reinterpret_cast<const nd2 *>(p));
  return make_double2(d.x, d.y);
}

// Orders a wavefront's LDS accesses across its lanes (its LDS operations
// execute in order; this keeps the compiler from moving them across).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One source per work-item (K = 1) with the links staged through LDS.  With
// NR lanes per site, loading links per lane fetches every link NR times from
// L1 into registers, and that per-lane data return, not HBM, bounds the
// kernel.  Here each wavefront covers SW = 64 / NR consecutive (logical) sites;
// per direction it loads their fat and Naik link pairs once, as 16-byte chunks
// spread over its 64 lanes (coalesced), stores them in its own LDS slice, and
// every lane reads its site's pair from there (SW distinct addresses per read:
// broadcasts).  The next direction's chunks are loaded into registers while
// the current one is applied.  The slice is private to the wavefront, so no
// block barrier is needed.  Arithmetic and its order are those of
// k_dslash_mrhs<T, NL, NR, 1, EPI>.
template <typename T, int NL, int NR, int EPI>
__global__ void __launch_bounds__(DSLASH_BLOCK) k_dslash_mrhs_lds(const MrhsArgsT<T> a) {
  static_assert(NR >= 2 && 64 % NR == 0, "NR lanes per site, whole sites per wavefront");
  using V = typename vec16<T>::type;
  constexpr int CPV = vec16<T>::ncplx;
  constexpr int NCH_F = 18 / CPV, NCH_L = NL / CPV, NCH = NCH_F + NCH_L;  // chunks per site and direction
  constexpr int SW = 64 / NR;                                             // sites per wavefront
  constexpr int CW = SW * NCH;                                            // chunks per wavefront and direction
  constexpr int QL = (CW + 63) / 64;                                      // chunks loaded per lane
  __shared__ V lds[DSLASH_BLOCK / 64][CW];
  const unsigned mask = rhs_mask<NR>(a.states);
  if (mask == 0) return;  // uniform over the grid
  const int vh = a.g.vh;
  const int lane = threadIdx.x & 63;
  V *wl = lds[threadIdx.x >> 6];
  const XcdWalk wk(a.nlog);
  double v = 0.0;
  for (int step = 0;; ++step) {
    const int blk = wk.at(step);
    if (blk < 0) break;
    const int L0 = (blk * DSLASH_BLOCK + (threadIdx.x & ~63)) / NR;  // first site of the wavefront
    if (L0 >= vh) continue;                                           // uniform per wavefront
    // chunks this lane loads: c = lane + 64 q -> chunk k of wavefront site s
    // (LDS slot k * SW + s); site -1 = none
    int lsite[QL], lk[QL];
#pragma unroll
    for (int q = 0; q < QL; ++q) {
      const int c = lane + 64 * q, k = c / SW, Ls = L0 + c % SW;
      lk[q] = k;
      lsite[q] = (c < CW && Ls < vh) ? ordered_site(a.g, a.order, Ls) : -1;
    }
    auto fetch = [&](int mu, V *pre) {
#pragma unroll
      for (int q = 0; q < QL; ++q) {
        const int i = lsite[q], k = lk[q];
        if (i < 0) continue;
        pre[q] = k < NCH_F ? nt_chunk(a.fat + link_chunk(mu, k, i, NCH_F))
                           : nt_chunk(a.lng + link_chunk(mu, k - NCH_F, i, NCH_L));
      }
    };
    // this lane's (site, source)
    const int s = lane / NR, r0 = lane % NR, L = L0 + s;
    const bool active = L < vh && ((mask >> r0) & 1u);
    int i = 0;
    SiteNbr n{};
    unsigned sg = 0;
    if (active) {
      i = ordered_site(a.g, a.order, L);
      site_neighbours(a.g, i, a.cb, n);
      if constexpr (NL == 12) sg = a.lsign[i];
    }
    cplx<T> acc[1][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[0][c] = cplx<T>{0, 0};
    V pre[QL];
    fetch(0, pre);
    // unrolled: the neighbour indices are selected with constant indices
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      wave_lds_sync();  // the previous direction's reads are done
#pragma unroll
      for (int q = 0; q < QL; ++q)
        if (lsite[q] >= 0) wl[lk[q] * SW + (lane + 64 * q) % SW] = pre[q];
      wave_lds_sync();
      if (mu < 3) fetch(mu + 1, pre);
      if (active) {
        // one link at a time (fewer live registers than a whole pair); complex
        // k of a pair starting at chunk c0 sits in LDS chunk c0 + k / CPV.
        // Same accumulation order as apply_links_k: forward, then back.
        auto lds_c = [&](int c0, int k) -> cplx<T> {
          const V d = wl[(c0 + k / CPV) * SW + s];
          if constexpr (CPV == 2) return (k % 2) ? cplx<T>{d.z, d.w} : cplx<T>{d.x, d.y};
          else return cplx<T>{d.x, d.y};
        };
        cplx<T> w[9], v[3];
#pragma unroll
        for (int e = 0; e < 9; ++e) w[e] = lds_c(0, e);
        Spin<T>::load(a.in, n.f1[mu] * NR + r0, v);
        matvec_acc(w, v, acc[0]);
#pragma unroll
        for (int e = 0; e < 9; ++e) w[e] = lds_c(0, 9 + e);
        Spin<T>::load(a.in, n.b1[mu] * NR + r0, v);
        matvec_acc(w, v, acc[0]);
        if constexpr (NL == 18) {
#pragma unroll
          for (int e = 0; e < 9; ++e) w[e] = lds_c(NCH_F, e);
          Spin<T>::load(a.in, n.f3[mu] * NR + r0, v);
          matvec_acc(w, v, acc[0]);
#pragma unroll
          for (int e = 0; e < 9; ++e) w[e] = lds_c(NCH_F, 9 + e);
          Spin<T>::load(a.in, n.b3[mu] * NR + r0, v);
          matvec_acc(w, v, acc[0]);
        } else {
#pragma unroll
          for (int e = 0; e < 6; ++e) w[e] = lds_c(NCH_F, e);
          recon_row2(w, ((sg >> (2 * mu + 1)) & 1u) ? (T)-1 : (T)1);
          Spin<T>::load(a.in, n.f3[mu] * NR + r0, v);
          matvec_acc(w, v, acc[0]);
#pragma unroll
          for (int e = 0; e < 6; ++e) w[e] = lds_c(NCH_F, 6 + e);
          recon_row2(w, ((sg >> (2 * mu)) & 1u) ? (T)-1 : (T)1);
          Spin<T>::load(a.in, n.b3[mu] * NR + r0, v);
          matvec_acc(w, v, acc[0]);
        }
      }
    }
    if (active) {
      const int slot = i * NR + r0;
      if constexpr (EPI == EPI_PKP) {
        cplx<T> p[3];
        Spin<T>::load(a.aux, slot, p);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          acc[0][c].re = fma(a.s, p[c].re, acc[0][c].re);
          acc[0][c].im = fma(a.s, p[c].im, acc[0][c].im);
          v = fma((double)p[c].re, (double)acc[0][c].re, v);
          v = fma((double)p[c].im, (double)acc[0][c].im, v);
        }
      }
      Spin<T>::store(a.out, slot, acc[0]);
    }
  }
  if constexpr (EPI == EPI_PKP) {
    const double sum = lane_block_sum<NR, DSLASH_BLOCK>(v);
    if (threadIdx.x < NR) a.partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = sum;
  }
}

// Sources r0 .. r0 + K - 1 of interleaved index j = site * NR + r0 (r0 a
// multiple of K, K dividing 64): neighbours in one tile of the tile-major
// layout, so for FP32 each colour of two sources is one 16-byte load.
template <typename T, int K> struct SpinK {
  static __device__ __forceinline__ void load(const T *p, int j, cplx<T> (*v)[3]) {
#pragma unroll
    for (int k = 0; k < K; ++k) Spin<T>::load(p, j + k, v[k]);
  }
};
template <int K> struct SpinK<float, K> {
  static_assert(K % 2 == 0, "pairs of sources");
  static __device__ __forceinline__ void load(const float *p, int j, cplx<float> (*v)[3]) {
    const float2 *q = reinterpret_cast<const float2 *>(p) + spin_cidx((size_t)j, 0);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int h = 0; h < K / 2; ++h) {
        const float4 d = reinterpret_cast<const float4 *>(q + SPIN_TILE * c)[h];
        v[2 * h][c] = cplx<float>{d.x, d.y};
        v[2 * h + 1][c] = cplx<float>{d.z, d.w};
      }
  }
};

// K = 2 or 4 sources per work-item with the links staged through LDS as in
// k_dslash_mrhs_lds: G = NR / K lanes per site, SW = 64 / G sites per
// wavefront.  Every link element a lane reads from LDS serves its K sources,
// and a lane's sources are neighbours in the interleaved layout (FP32: one
// 16-byte load per gathered colour for each pair).  Per source the arithmetic
// and its order are those of k_dslash_mrhs<T, NL, NR, K, EPI>.
template <typename T, int NL, int NR, int K, int EPI>
__global__ void __launch_bounds__(DSLASH_BLOCK) k_dslash_mrhs_ldsk(const MrhsArgsT<T> a) {
  static_assert((K == 2 || K == 4) && NR >= 2 * K && 64 % (NR / K) == 0, "G >= 2 lanes per site");
  using V = typename vec16<T>::type;
  constexpr int CPV = vec16<T>::ncplx;
  constexpr int G = NR / K;                                               // lanes per site
  constexpr int NCH_F = 18 / CPV, NCH_L = NL / CPV, NCH = NCH_F + NCH_L;  // chunks per site and direction
  constexpr int SW = 64 / G;                                              // sites per wavefront
  constexpr int CW = SW * NCH;                                            // chunks per wavefront and direction
  constexpr int QL = (CW + 63) / 64;                                      // chunks loaded per lane
  __shared__ V lds[DSLASH_BLOCK / 64][CW];
  const unsigned mask = rhs_mask<NR>(a.states);
  if (mask == 0) return;  // uniform over the grid
  const int vh = a.g.vh;
  const int lane = threadIdx.x & 63;
  V *wl = lds[threadIdx.x >> 6];
  const XcdWalk wk(a.nlog);
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.0;
  for (int step = 0;; ++step) {
    const int blk = wk.at(step);
    if (blk < 0) break;
    const int L0 = (blk * DSLASH_BLOCK + (threadIdx.x & ~63)) / G;  // first site of the wavefront
    if (L0 >= vh) continue;                                          // uniform per wavefront
    int lsite[QL], lk[QL];
#pragma unroll
    for (int q = 0; q < QL; ++q) {
      const int c = lane + 64 * q, k = c / SW, Ls = L0 + c % SW;
      lk[q] = k;
      lsite[q] = (c < CW && Ls < vh) ? ordered_site(a.g, a.order, Ls) : -1;
    }
    auto fetch = [&](int mu, V *pre) {
#pragma unroll
      for (int q = 0; q < QL; ++q) {
        const int i = lsite[q], k = lk[q];
        if (i < 0) continue;
        pre[q] = k < NCH_F ? nt_chunk(a.fat + link_chunk(mu, k, i, NCH_F))
                           : nt_chunk(a.lng + link_chunk(mu, k - NCH_F, i, NCH_L));
      }
    };
    // this lane's site and sources r0 .. r0 + K - 1
    const int s = lane / G, r0 = (lane % G) * K, L = L0 + s;
    const unsigned m = (mask >> r0) & ((1u << K) - 1u);
    const bool active = L < vh && m != 0;
    int i = 0;
    SiteNbr n{};
    unsigned sg = 0;
    if (active) {
      i = ordered_site(a.g, a.order, L);
      site_neighbours(a.g, i, a.cb, n);
      if constexpr (NL == 12) sg = a.lsign[i];
    }
    cplx<T> acc[K][3];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[k][c] = cplx<T>{0, 0};
    V pre[QL];
    fetch(0, pre);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      wave_lds_sync();  // the previous direction's reads are done
#pragma unroll
      for (int q = 0; q < QL; ++q)
        if (lsite[q] >= 0) wl[lk[q] * SW + (lane + 64 * q) % SW] = pre[q];
      wave_lds_sync();
      if (mu < 3) fetch(mu + 1, pre);
      if (active) {
        auto lds_c = [&](int c0, int k) -> cplx<T> {
          const V d = wl[(c0 + k / CPV) * SW + s];
          if constexpr (CPV == 2) return (k % 2) ? cplx<T>{d.z, d.w} : cplx<T>{d.x, d.y};
          else return cplx<T>{d.x, d.y};
        };
        // one link and one neighbour at a time, applied to the K sources
        auto apply = [&](const cplx<T> *w, int nb) {
          cplx<T> vk[K][3];
          SpinK<T, K>::load(a.in, nb * NR + r0, vk);
#pragma unroll
          for (int k = 0; k < K; ++k) matvec_acc(w, vk[k], acc[k]);
        };
        cplx<T> w[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) w[e] = lds_c(0, e);
        apply(w, n.f1[mu]);
#pragma unroll
        for (int e = 0; e < 9; ++e) w[e] = lds_c(0, 9 + e);
        apply(w, n.b1[mu]);
        if constexpr (NL == 18) {
#pragma unroll
          for (int e = 0; e < 9; ++e) w[e] = lds_c(NCH_F, e);
          apply(w, n.f3[mu]);
#pragma unroll
          for (int e = 0; e < 9; ++e) w[e] = lds_c(NCH_F, 9 + e);
          apply(w, n.b3[mu]);
        } else {
#pragma unroll
          for (int e = 0; e < 6; ++e) w[e] = lds_c(NCH_F, e);
          recon_row2(w, ((sg >> (2 * mu + 1)) & 1u) ? (T)-1 : (T)1);
          apply(w, n.f3[mu]);
#pragma unroll
          for (int e = 0; e < 6; ++e) w[e] = lds_c(NCH_F, 6 + e);
          recon_row2(w, ((sg >> (2 * mu)) & 1u) ? (T)-1 : (T)1);
          apply(w, n.b3[mu]);
        }
      }
      // the next direction's gathers stay after this one's: hoisted, the
      // loads of all four directions (K sources each) are live together
      __builtin_amdgcn_sched_barrier(0);
    }
    if (active) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (!((m >> k) & 1u)) continue;
        const int slot = i * NR + r0 + k;
        if constexpr (EPI == EPI_PKP) {
          cplx<T> p[3];
          Spin<T>::load(a.aux, slot, p);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            acc[k][c].re = fma(a.s, p[c].re, acc[k][c].re);
            acc[k][c].im = fma(a.s, p[c].im, acc[k][c].im);
            v[k] = fma((double)p[c].re, (double)acc[k][c].re, v[k]);
            v[k] = fma((double)p[c].im, (double)acc[k][c].im, v[k]);
          }
        }
        Spin<T>::store(a.out, slot, acc[k]);
      }
    }
  }
  if constexpr (EPI == EPI_PKP) {
    const double sum = group_block_sum<G, K, DSLASH_BLOCK>(v);
    if (threadIdx.x < NR) a.partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = sum;
  }
}

} // namespace qphix
