// dslash_mrhs.h -- Dslash applied to NR right-hand sides in one launch.
//
// Links are loaded once per site and applied to NR spinors, so link traffic
// (96% of a single-RHS Dslash's bytes) is shared: bytes per site become
// links + NR x 2 x 24 B instead of NR x (links + 48 B).  Spinors of the NR
// right-hand sides are interleaved per site ([site][rhs][colour][re,im], FP32),
// so one neighbour gather fetches all NR of them contiguously.  Each RHS
// carries its own gate (its solver state's `active`); inactive RHS skip their
// stores.  Two thread mappings: one work-item per site looping over the NR
// sources (k_dslash_mrhs), or one per (site, source) (k_dslash_mrhs_lane).
#pragma once
#include "dslash_kernel.h"

namespace qphix {

struct MrhsArgs {
  const float4 *fat, *lng;
  const uint8_t *lsign;
  const float *in;   // [vh][NR][6]
  float *out;        // [vh][NR][6]
  const float *aux;  // EPI_PKP: p, [vh][NR][6]
  Geo g;
  int cb;
  float s;           // EPI_PKP: out = D in + s * p
  double *partials;  // EPI_PKP: [NR][gridDim.x]
  const SolverState *states;  // per-RHS gates (states[r].active)
  int nlog;          // lane mapping: logical 256-item blocks (vh * NR / 256, rounded up)
};

template <int NR> __device__ __forceinline__ unsigned rhs_mask(const SolverState *st) {
  unsigned m = 0;
#pragma unroll
  for (int r = 0; r < NR; ++r) m |= (st[r].active ? 1u : 0u) << r;
  return m;
}

template <int NC, int NR>
__device__ __forceinline__ void apply_pair_mrhs(const float4 *base, int vh, int i, const float *in,
                                                int nf, int nb, unsigned sg_f, unsigned sg_b,
                                                unsigned mask, cplx<float> (*acc)[3]) {
  cplx<float> u[NC];
  PairLoad<float, NC, true>::load(base, vh, i, u);
  cplx<float> wf[9], wb[9];
  if constexpr (NC == 18) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      wf[k] = u[k];
      wb[k] = u[9 + k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      wf[k] = u[k];
      wb[k] = u[6 + k];
    }
    recon_row2(wf, sg_f ? -1.f : 1.f);
    recon_row2(wb, sg_b ? -1.f : 1.f);
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (!((mask >> r) & 1u)) continue;
    cplx<float> vf[3], vb[3];
    Spin<float>::load(in, nf * NR + r, vf);
    Spin<float>::load(in, nb * NR + r, vb);
    matvec_acc(wf, vf, acc[r]);
    matvec_acc(wb, vb, acc[r]);
  }
}

template <int NL, int NR, int EPI>
__global__ void __launch_bounds__(DSLASH_BLOCK) k_dslash_mrhs(const MrhsArgs a) {
  const unsigned mask = rhs_mask<NR>(a.states);
  if (mask == 0) return;  // uniform over the grid
  constexpr int NCH_F = 9, NCH_L = NL / 2;
  const int vh = a.g.vh;
  const int i = block_slot(OPT_XCD) * blockDim.x + threadIdx.x;
  const bool valid = i < vh;
  cplx<float> acc[NR][3];
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[r][c] = cplx<float>{0.f, 0.f};
  if (valid) {
    SiteNbr n;
    site_neighbours(a.g, i, a.cb, n);
    unsigned sg = 0;
    if constexpr (NL == 12) sg = a.lsign[i];
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      const size_t tile = (size_t)(i / LINK_TILE);
      const int ii = i % LINK_TILE;
      apply_pair_mrhs<18, NR>(a.fat + (tile * 4 * NCH_F + mu * NCH_F) * LINK_TILE, LINK_TILE, ii,
                              a.in, n.f1[mu], n.b1[mu], 0, 0, mask, acc);
      apply_pair_mrhs<NL, NR>(a.lng + (tile * 4 * NCH_L + mu * NCH_L) * LINK_TILE, LINK_TILE, ii,
                              a.in, n.f3[mu], n.b3[mu], (sg >> (2 * mu + 1)) & 1u,
                              (sg >> (2 * mu)) & 1u, mask, acc);
    }
  }
  if constexpr (EPI == EPI_PLAIN) {
    if (valid) {
#pragma unroll
      for (int r = 0; r < NR; ++r)
        if ((mask >> r) & 1u) Spin<float>::store(a.out, i * NR + r, acc[r]);
    }
  } else {
    double v[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) v[r] = 0.0;
    if (valid) {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if (!((mask >> r) & 1u)) continue;
        cplx<float> p[3];
        Spin<float>::load(a.aux, i * NR + r, p);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          acc[r][c].re = acc[r][c].re + a.s * p[c].re;
          acc[r][c].im = acc[r][c].im + a.s * p[c].im;
          v[r] += (double)p[c].re * acc[r][c].re + (double)p[c].im * acc[r][c].im;
        }
        Spin<float>::store(a.out, i * NR + r, acc[r]);
      }
    }
    block_sum<NR>(v);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int r = 0; r < NR; ++r) a.partials[(size_t)r * gridDim.x + blockIdx.x] = v[r];
    }
  }
}


// Lane-per-(site, source) variant: work-item t handles site t / NR for source
// t % NR, so a wavefront covers 64 / NR sites.  The NR lanes of a site load
// the same link addresses (merged by the memory pipeline, so HBM link traffic
// is still shared NR ways) while each lane keeps the register footprint of
// a single-source Dslash; inactive sources are masked per lane and all loads
// of an active lane are issued together, as in k_dslash.
//
// The grid is resident-sized (a multiple of N_XCD): the nlog logical 256-item
// blocks are split into N_XCD contiguous slabs, XCD x = blockIdx % N_XCD walks
// slab x with its gridDim.x / N_XCD blocks (neighbour gathers stay in that
// XCD's L2), and EPI_PKP leaves one partial per resident block and source.
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

template <int NL, int NR>
__device__ __forceinline__ void mrhs_lane_site(const MrhsArgs &a, int i, int r, cplx<float> *acc) {
  constexpr int NCH_F = 9, NCH_L = NL / 2;
  const int vh = a.g.vh;
  SiteNbr n;
  site_neighbours(a.g, i, a.cb, n);
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    n.f1[mu] = n.f1[mu] * NR + r;
    n.b1[mu] = n.b1[mu] * NR + r;
    n.f3[mu] = n.f3[mu] * NR + r;
    n.b3[mu] = n.b3[mu] * NR + r;
  }
  unsigned sg = 0;
  if constexpr (NL == 12) sg = a.lsign[i];
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    const size_t tile = (size_t)(i / LINK_TILE);
    const int ii = i % LINK_TILE;
    apply_pair<float, 18, float, true>(a.fat + (tile * 4 * NCH_F + mu * NCH_F) * LINK_TILE, LINK_TILE,
                                       ii, a.in, n.f1[mu], n.b1[mu], 0, 0, acc);
    apply_pair<float, NL, float, true>(a.lng + (tile * 4 * NCH_L + mu * NCH_L) * LINK_TILE, LINK_TILE,
                                       ii, a.in, n.f3[mu], n.b3[mu], (sg >> (2 * mu + 1)) & 1u,
                                       (sg >> (2 * mu)) & 1u, acc);
  }
}

// Per-source sums of v over the block (lanes l and l + k*NR of a wavefront
// share a source); thread r < NR returns source r's sum.
template <int NR, int BLOCK> __device__ __forceinline__ double lane_block_sum(double v) {
#pragma unroll
  for (int off = 32; off >= NR; off >>= 1) v += __shfl_down(v, off, 64);
  __shared__ double sh[BLOCK / 64][NR];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane < NR) sh[w][lane] = v;
  __syncthreads();
  double sum = 0.0;
  if (threadIdx.x < NR) {
#pragma unroll
    for (int k = 0; k < BLOCK / 64; ++k) sum += sh[k][threadIdx.x];
  }
  return sum;
}

template <int NL, int NR, int EPI>
__global__ void __launch_bounds__(DSLASH_BLOCK) k_dslash_mrhs_lane(const MrhsArgs a) {
  const unsigned mask = rhs_mask<NR>(a.states);
  if (mask == 0) return;  // uniform over the grid
  const int vh = a.g.vh;
  const XcdWalk wk(a.nlog);
  double v = 0.0;
  for (int k = 0;; ++k) {
    const int blk = wk.at(k);
    if (blk < 0) break;
    const int t = blk * DSLASH_BLOCK + threadIdx.x;
    const int i = t / NR, r = t % NR;
    if (i >= vh || !((mask >> r) & 1u)) continue;
    cplx<float> acc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] = cplx<float>{0.f, 0.f};
    mrhs_lane_site<NL, NR>(a, i, r, acc);
    if constexpr (EPI == EPI_PKP) {
      cplx<float> p[3];
      Spin<float>::load(a.aux, t, p);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        acc[c].re = acc[c].re + a.s * p[c].re;
        acc[c].im = acc[c].im + a.s * p[c].im;
        v += (double)p[c].re * acc[c].re + (double)p[c].im * acc[c].im;
      }
    }
    Spin<float>::store(a.out, t, acc);
  }
  if constexpr (EPI == EPI_PKP) {
    const double sum = lane_block_sum<NR, DSLASH_BLOCK>(v);
    if (threadIdx.x < NR) a.partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = sum;
  }
}

} // namespace qphix
