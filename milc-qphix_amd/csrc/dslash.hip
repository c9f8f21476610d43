// dslash.hip -- even-odd asqtad/HISQ Dslash for gfx950.
//
// Replaces the reference's per-parity site loop (ks_long_dslash.cpp:564-668)
// and its absent codegen kernel body (ks_dslash_complete_specialization.cpp:14-54):
//   out(x) = sum_mu [ F_{2mu+1}(x) in(x+mu) + F_{2mu}(x) in(x-mu)
//                   + N_{2mu+1}(x) in(x+3mu) + N_{2mu}(x) in(x-3mu) ]
// for every site x of the output parity, with the 16 links applied as stored
// (1146 flops/site).  One work-item per output site; neighbour indices come from
// coordinate arithmetic with periodic wrap (no neighbour table, no halo: the
// reference's boundary.cpp / ks_boundary face exchange has no counterpart on one
// GPU).  Links stream from HBM once per application in 16-byte coalesced
// chunks; the 16 neighbour spinor reads per site hit L2 / Infinity Cache.
//
// Fused epilogues remove whole vector passes from the solvers:
//   EPI_PKP     out = D in + s*p, reduce p.out      (calc_pkp, ks_d_congrad_fn.cpp:159-171)
//   EPI_RESTART r = b + D in + s*x, p = r, reduce |r|^2 and the relative residue
//                                                   (calc_rsq, ks_d_congrad_fn.cpp:82-99;
//                                                    relative_residue, ks_blas_utils_c.cpp:110-149)
//   EPI_PKP_FS  out = D in + s*p, reduce p.out, r.out, out.out and r.r: the one
//               reduction of a single-reduction (FEWSUMS) CG step
//                                                   (ks_d_congrad_fn.cpp:150-158, 450-541)
#include "blas.h"
#include "dslash_kernel.h"

namespace qphix {

// Kernel variant of every launch (see dslash_kernel.h).  Measured on MI355X
// at 32^4 (Naik 12, tile-major links): XCD slabs + streaming link loads take
// FP64 from 230 to 201 us and FP32 from 104 to 95 us; streaming output stores
// take them to 188 and 93 us.  The output stores stream even where the next
// kernel reads the result (a solver's tmp = D p): that pair of Dslashes is 3%
// faster than with cached stores (FP32 190 vs 196 us, FP64 390 vs 403 us).
constexpr int DSLASH_OPT = OPT_XCD | OPT_NTL | OPT_NTS;

// (Measured and dropped: each XCD walking a z-region t-slice by t-slice
// instead of a t-slab, as the multi-RHS kernels do -- FP64 169.5 vs 170.0 us,
// FP32 87.7 vs 87.3 us at 32^4: one source's +-3 t-slice window already fits.)
template <typename TL, int NL, typename SI, typename SO, int EPI>
static void launch_one(const DslashKArgs<TL, SI, SO> &a, int nblocks) {
  hipLaunchKernelGGL((k_dslash<TL, NL, SI, SO, EPI, DSLASH_OPT>), dim3(nblocks), dim3(DSLASH_BLOCK), 0,
                     ctx().stream, a);
}

template <typename TL, typename SI, typename SO>
static void dispatch(const DslashKArgs<TL, SI, SO> &a, int recon, int epi, bool consumer = false) {
  const int nb = (a.g.vh + DSLASH_BLOCK - 1) / DSLASH_BLOCK;
  if (recon == 8) {  // built for the FP32 sloppy operator only
    if constexpr (std::is_same<TL, float>::value) {
      if (epi == EPI_PLAIN) launch_one<TL, 8, SI, SO, EPI_PLAIN>(a, nb);
      else if (epi == EPI_PKP) launch_one<TL, 8, SI, SO, EPI_PKP>(a, nb);
      else if (epi == EPI_PKP_FS) launch_one<TL, 8, SI, SO, EPI_PKP_FS>(a, nb);
      else launch_one<TL, 8, SI, SO, EPI_RESTART>(a, nb);
    } else {
      QPHIX_REQUIRE(false, "Dslash: 8-real long links are FP32 only");
    }
  } else if (recon == 12) {
    if (epi == EPI_PLAIN) launch_one<TL, 12, SI, SO, EPI_PLAIN>(a, nb);
    else if (epi == EPI_PKP) launch_one<TL, 12, SI, SO, EPI_PKP>(a, nb);
    else if (epi == EPI_PKP_FS) launch_one<TL, 12, SI, SO, EPI_PKP_FS>(a, nb);
    else launch_one<TL, 12, SI, SO, EPI_RESTART>(a, nb);
  } else {
    if (epi == EPI_PLAIN) launch_one<TL, 18, SI, SO, EPI_PLAIN>(a, nb);
    else if (epi == EPI_PKP) launch_one<TL, 18, SI, SO, EPI_PKP>(a, nb);
    else if (epi == EPI_PKP_FS) launch_one<TL, 18, SI, SO, EPI_PKP_FS>(a, nb);
    else launch_one<TL, 18, SI, SO, EPI_RESTART>(a, nb);
  }
  HIP_CHECK(hipGetLastError());
  // the single-reduction step always sums its partials in the next kernel
  if (consumer || epi == EPI_PKP_FS) return;  // the next kernel sums the partials itself
  if (epi == EPI_PKP) reduce_finish(1, nb, a.rc, a.gate);
  else if (epi == EPI_RESTART) reduce_finish(2, nb, a.rc, a.gate);
}

static Geo make_geo() {
  const Geometry &g = ctx().g;
  return Geo{g.nxh, g.ny, g.nz, g.nt, g.vh, g.nxy, g.nxyz};
}

template <typename TL, typename SI, typename SO>
static DslashKArgs<TL, SI, SO> make_args(const LinksT<TL> &L, int cb, SO *out, const SI *in,
                                         const DslashExtra &ex) {
  DslashKArgs<TL, SI, SO> a{};
  a.fat = reinterpret_cast<const typename vec16<TL>::type *>(L.fat[cb]);
  a.lng = reinterpret_cast<const typename vec16<TL>::type *>(L.lng[cb]);
  a.lsign = L.lsign[cb];
  a.in = in;
  a.out = out;
  a.g = make_geo();
  a.cb = cb;
  a.s = (typename link_compute<TL>::type)ex.s;
  a.aux = reinterpret_cast<const SI *>(ex.aux);
  a.b = reinterpret_cast<const SI *>(ex.b);
  a.r = reinterpret_cast<SO *>(ex.r);
  a.p = reinterpret_cast<SO *>(ex.p);
  a.rc = reduce_ctl(reinterpret_cast<SolverState *>(ex.state), ex.scalar_op);
  a.gate = ex.gate;
  return a;
}

template <typename T>
void launch_dslash(const LinksT<T> &L, int cb, T *out, const T *in, int epi, const DslashExtra &ex) {
  QPHIX_REQUIRE(L.fat[cb] && L.lng[cb] && in && out, "Dslash: links or vectors of this parity missing");
  dispatch<T, T, T>(make_args<T, T, T>(L, cb, out, in, ex), L.recon, epi, ex.consumer_finalize);
}

template void launch_dslash<double>(const LinksT<double> &, int, double *, const double *, int,
                                    const DslashExtra &);
template void launch_dslash<float>(const LinksT<float> &, int, float *, const float *, int,
                                   const DslashExtra &);

// FP16-stored spinors with FP32 links and arithmetic.  out_h (half) or
// out_f (float) receives the result.
void launch_dslash_half(const LinksT<float> &L, int cb, HalfSite *out_h, float *out_f,
                        const HalfSite *in, int epi, const DslashExtra &ex) {
  QPHIX_REQUIRE(L.fat[cb] && L.lng[cb] && in && (out_h || out_f), "Dslash: links or vectors of this parity missing");
  if (out_h)
    dispatch<float, HalfSite, HalfSite>(make_args<float, HalfSite, HalfSite>(L, cb, out_h, in, ex), L.recon, epi,
                                        ex.consumer_finalize);
  else
    dispatch<float, HalfSite, float>(make_args<float, HalfSite, float>(L, cb, out_f, in, ex), L.recon, epi,
                                     ex.consumer_finalize);
}

void launch_dslash_q16(const LinksT<Q16> &L, int cb, float *out, const float *in, int epi,
                       const DslashExtra &ex) {
  QPHIX_REQUIRE(L.fat[cb] && L.lng[cb] && in && out, "Dslash: links or vectors of this parity missing");
  dispatch<Q16, float, float>(make_args<Q16, float, float>(L, cb, out, in, ex), L.recon, epi, ex.consumer_finalize);
}

void launch_dslash_q16_half(const LinksT<Q16> &L, int cb, HalfSite *out_h, float *out_f, const HalfSite *in,
                            int epi, const DslashExtra &ex) {
  QPHIX_REQUIRE(L.fat[cb] && L.lng[cb] && in && (out_h || out_f), "Dslash: links or vectors of this parity missing");
  if (out_h)
    dispatch<Q16, HalfSite, HalfSite>(make_args<Q16, HalfSite, HalfSite>(L, cb, out_h, in, ex), L.recon, epi,
                                      ex.consumer_finalize);
  else
    dispatch<Q16, HalfSite, float>(make_args<Q16, HalfSite, float>(L, cb, out_f, in, ex), L.recon, epi,
                                   ex.consumer_finalize);
}

} // namespace qphix
