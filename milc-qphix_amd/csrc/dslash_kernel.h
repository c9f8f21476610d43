// dslash_kernel.h -- the Dslash kernel template (one work-item per output site).
#pragma once
#include "reduce.h"

namespace qphix {

constexpr int DSLASH_BLOCK = 256;
constexpr int N_XCD = 8;  // XCDs of an MI355X; blocks are dispatched round-robin over them

// Site traversal order (multi-RHS kernels).  In site order a site's t +- 1, 3 neighbours are
// read 1-3 t-slices earlier or later (3 MB per slice for 8 interleaved FP32
// sources at 32^3x64), beyond a 4 MB XCD L2, so they are re-fetched.  With
// `on`, XCD x's slab of t-slices is walked in (y, z) patches of py x pz rows:
// all t of the slab for one patch, then the next patch, so the neighbours a
// patch needs (its +-3 halo over 7 t-slices) stay in L2.  A pure permutation
// of which lane computes which site: results are unchanged.
// on = 2: XCD x takes the z-planes [x nz / 8, (x + 1) nz / 8) instead of a
// t-slab and walks them t-slice by t-slice, so that the 8 XCDs advance through
// t together: the +-3 t-slice window every XCD gathers from is then 7 slices of
// its own z-region plus a 3-plane halo (FP64, 8 sources, 32^3x64: 14 MB per
// XCD, 110 MB in all, inside the 256 MB Infinity Cache) instead of 7 whole
// t-slices (44 MB per XCD, 352 MB in all).
struct SiteOrder {
  int on, py, pz, tslab;  // tslab = nt / N_XCD
};

// Logical index L -> site for on = 2 (nz a multiple of N_XCD).
__device__ __forceinline__ int zregion_site(const Geo &g, int L) {
  const int nzl = g.nz / N_XCD;
  const int per_xcd = g.nxy * nzl * g.nt;
  const int x = L / per_xcd;
  int r = L - x * per_xcd;
  const int slice = g.nxy * nzl;  // sites of the region in one t-slice
  const int t = r / slice;
  r -= t * slice;
  const int z = x * nzl + r / g.nxy;
  return (r % g.nxy) + g.nxy * (z + g.nz * t);
}

__device__ __forceinline__ int ordered_site(const Geo &g, const SiteOrder &o, int L) {
  if (!o.on) return L;
  if (o.on == 2) return zregion_site(g, L);
  const int per_xcd = g.nxyz * o.tslab;
  const int x = L / per_xcd;
  int r = L - x * per_xcd;
  const int patch = g.nxh * o.py * o.pz;  // sites of one patch in one t-slice
  const int column = patch * o.tslab;
  const int pidx = r / column;
  r -= pidx * column;
  const int tt = r / patch;
  r -= tt * patch;
  const int xh = r % g.nxh;
  r /= g.nxh;
  const int yy = r % o.py, zz = r / o.py;
  const int npy = g.ny / o.py;
  const int y = (pidx % npy) * o.py + yy, z = (pidx / npy) * o.pz + zz;
  const int t = x * o.tslab + tt;
  return xh + g.nxh * (y + g.ny * (z + g.nz * t));
}



template <typename TL, typename SI, typename SO> struct DslashKArgs {
  const typename vec16<TL>::type *fat;  // output parity's links, tile-major
  const typename vec16<TL>::type *lng;
  const uint8_t *lsign;                 // 12- and 8-real signs (bit 2mu+1 fwd, 2mu back)
  const SI *in;
  SO *out;
  Geo g;
  int cb;
  typename link_compute<TL>::type s;  // epilogue scalar
  const SI *aux;
  const SI *b;
  SO *r, *p;
  ReduceCtl rc;
  const int *gate;
};

// Kernel variants (compile-time bit mask):
//   OPT_XCD  blocks are renumbered so that each of the 8 XCDs (blocks are
//            dispatched round-robin over XCDs) sweeps one contiguous slab of
//            sites, keeping its neighbour gathers inside its own L2
//   OPT_NTL  links are loaded non-temporally (streamed, not cached)
//   OPT_NTS  the output is stored non-temporally
enum { OPT_XCD = 1, OPT_NTL = 2, OPT_NTS = 4 };

__device__ __forceinline__ int block_slot(int opt) {
  const int nb = gridDim.x, b = blockIdx.x;
  if ((opt & OPT_XCD) && nb % N_XCD == 0) return (b % N_XCD) * (nb / N_XCD) + b / N_XCD;
  return b;
}

// Applies one direction's link pair (forward link on in(fwd), back link on
// in(bwd)).  NC = 18: both links complete; NC = 12: rows 0-1 of each, row 2
// rebuilt as sign * conj(r0 x r1); NC = 8 (FP32 only): each link's 8-real form
// (recon8).
// The arithmetic of one pair: acc += U_f vf + U_b vb, u the pair as loaded.
// Shared by every kernel that must give the same sums bit for bit.
template <int NC, typename TC>
__device__ __forceinline__ void pair_compute(const cplx<TC> *u, const cplx<TC> *vf, const cplx<TC> *vb,
                                             unsigned sg_f, unsigned sg_b, cplx<TC> *acc) {
  if constexpr (NC == 18) {
    matvec_acc(u, vf, acc);
    matvec_acc(u + 9, vb, acc);
  } else if constexpr (NC == 8) {
    static_assert(std::is_same<TC, float>::value, "8-real links are FP32 only");
    cplx<TC> w[9];
    recon8(u, sg_f ? -1.f : 1.f, w);
    matvec_acc(w, vf, acc);
    recon8(u + 4, sg_b ? -1.f : 1.f, w);
    matvec_acc(w, vb, acc);
  } else {
    cplx<TC> w[9];
#pragma unroll
    for (int k = 0; k < 6; ++k) w[k] = u[k];
    recon_row2(w, sg_f ? (TC)-1 : (TC)1);
    matvec_acc(w, vf, acc);
#pragma unroll
    for (int k = 0; k < 6; ++k) w[k] = u[6 + k];
    recon_row2(w, sg_b ? (TC)-1 : (TC)1);
    matvec_acc(w, vb, acc);
  }
}

template <typename TL, int NC, typename SI, bool NT>
__device__ __forceinline__ void apply_pair(const typename vec16<TL>::type *base, int vh, int i,
                                           const SI *in, int nf, int nb, unsigned sg_f,
                                           unsigned sg_b, cplx<typename link_compute<TL>::type> *acc) {
  using TC = typename link_compute<TL>::type;
  cplx<TC> u[NC];
  PairLoad<TL, NC, NT>::load(base, vh, i, u);
  cplx<TC> vf[3], vb[3];
  Spin<SI>::load(in, nf, vf);
  Spin<SI>::load(in, nb, vb);
  pair_compute<NC, TC>(u, vf, vb, sg_f, sg_b, acc);
}

// Occupancy target (waves per SIMD) for k_dslash, for measurements; unset, the
// compiler's register allocation sets it.
#ifdef QPHIX_DSLASH_WAVES_PER_EU
#define QPHIX_DSLASH_OCC __attribute__((amdgpu_waves_per_eu(QPHIX_DSLASH_WAVES_PER_EU)))
#else
#define QPHIX_DSLASH_OCC
#endif

template <typename TL, int NL, typename SI, typename SO, int EPI, int OPT>
__global__ void __launch_bounds__(DSLASH_BLOCK) QPHIX_DSLASH_OCC k_dslash(const DslashKArgs<TL, SI, SO> a) {
  if (a.gate && *a.gate == 0) {  // uniform over the grid
    if constexpr (EPI == EPI_PKP_FS) {
      if (blockIdx.x == 0 && threadIdx.x == 0) a.rc.state->fs_gate = 0;  // the step is off
    }
    return;
  }
  if constexpr (EPI == EPI_PKP_FS) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      SolverState *st = a.rc.state;
      st->fs_gate = 1;
      st->fs_iter = st->iteration;
      st->fs_rsq = st->rsq;
      st->fs_mxrsq = st->mx_rsq;
    }
  }
  using TC = typename link_compute<TL>::type;
  constexpr int NCH_F = link_nch<TL>(18), NCH_L = link_nch<TL>(NL);
  constexpr bool NTL = (OPT & OPT_NTL) != 0;
  const int vh = a.g.vh;
  const int i = block_slot(OPT) * blockDim.x + threadIdx.x;
  const bool valid = i < vh;
  cplx<TC> acc[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) acc[c] = cplx<TC>{0, 0};
  if (valid) {
    SiteNbr n;
    site_neighbours(a.g, i, a.cb, n);
    unsigned sg = 0;
    if constexpr (NL == 12 || NL == 8) sg = a.lsign[i];
    // grouped tile-major links (link_chunk): chunk k of direction mu at
    // link_base + (mu * NCH + k) * LINK_CSTRIDE + i % LINK_TILE
    const typename vec16<TL>::type *fb = a.fat + link_base(i, NCH_F);
    const typename vec16<TL>::type *lb = a.lng + link_base(i, NCH_L);
    const int ii = i % LINK_TILE;
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      apply_pair<TL, 18, SI, NTL>(fb + (size_t)mu * NCH_F * LINK_CSTRIDE, LINK_CSTRIDE, ii, a.in,
                                  n.f1[mu], n.b1[mu], 0, 0, acc);
      apply_pair<TL, NL, SI, NTL>(lb + (size_t)mu * NCH_L * LINK_CSTRIDE, LINK_CSTRIDE, ii, a.in,
                                  n.f3[mu], n.b3[mu], (sg >> (2 * mu + 1)) & 1u,
                                  (sg >> (2 * mu)) & 1u, acc);
    }
  }
  if constexpr (EPI == EPI_PLAIN) {
    if (valid) {
      if constexpr ((OPT & OPT_NTS) != 0) Spin<SO>::store_nt(a.out, i, acc);
      else Spin<SO>::store(a.out, i, acc);
    }
  } else if constexpr (EPI == EPI_PKP) {
    double v[1] = {0.0};
    if (valid) {
      cplx<TC> p[3];
      Spin<SI>::load(a.aux, i, p);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        acc[c].re = acc[c].re + a.s * p[c].re;
        acc[c].im = acc[c].im + a.s * p[c].im;
        v[0] += (double)p[c].re * (double)acc[c].re + (double)p[c].im * (double)acc[c].im;
      }
      Spin<SO>::store(a.out, i, acc);
    }
    grid_reduce<1>(v, a.rc);
  } else if constexpr (EPI == EPI_PKP_FS) {
    // ks_d_congrad_fn.cpp:150-158 (FEWSUMS): with ttt = D in + s p, the sums
    // p.ttt, Re<ttt|r>, <ttt|ttt> and |r|^2 in one reduction
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (valid) {
      cplx<TC> p[3], r[3];
      Spin<SI>::load(a.aux, i, p);
      Spin<SO>::load(a.r, i, r);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        acc[c].re = acc[c].re + a.s * p[c].re;
        acc[c].im = acc[c].im + a.s * p[c].im;
        v[0] += (double)p[c].re * (double)acc[c].re + (double)p[c].im * (double)acc[c].im;
        v[1] += (double)r[c].re * (double)acc[c].re + (double)r[c].im * (double)acc[c].im;
        v[2] += (double)acc[c].re * (double)acc[c].re + (double)acc[c].im * (double)acc[c].im;
        v[3] += (double)r[c].re * (double)r[c].re + (double)r[c].im * (double)r[c].im;
      }
      Spin<SO>::store(a.out, i, acc);
    }
    grid_reduce<4>(v, a.rc);
  } else {  // EPI_RESTART
    double v[2] = {0.0, 0.0};
    if (valid) {
      cplx<TC> x[3], b[3];
      Spin<SI>::load(a.aux, i, x);
      Spin<SI>::load(a.b, i, b);
      double num = 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const TC tr = acc[c].re + a.s * x[c].re;
        const TC ti = acc[c].im + a.s * x[c].im;
        acc[c].re = b[c].re + tr;
        acc[c].im = b[c].im + ti;
        num += (double)acc[c].re * acc[c].re + (double)acc[c].im * acc[c].im;
      }
      Spin<SO>::store(a.r, i, acc);
      Spin<SO>::store(a.p, i, acc);
      v[0] = num;
      v[1] = rel_term(site_norm2_rounded(acc), site_norm2_rounded(x));
    }
    grid_reduce<2>(v, a.rc);
  }
}

} // namespace qphix
