// fermion_force.hip -- HISQ/asqtad fermion-force outer products.
//
// No counterpart in the reference (milc-qphix stops at the solvers and the
// gauge force).  For the pseudofermion action
//     S = sum_k c_k Re phi_k^dag X_k,   X_k = (sigma_k - D_eo D_oe)^{-1} phi_k
// (even sites; several right-hand sides are just more terms), the derivative
// with respect to a forward fat link U(y,mu) -- back links being the MILC
// double-store -U(y-mu,mu)^dag -- is dS = 2 Re Tr[dU(y,mu) O1(y,mu)] with
//     O1(y,mu) = eps(y) sum_k c_k psi_k(y+mu) psi_k(y)^dag,
//     psi_k = X_k on even sites, Y_k = D_oe X_k on odd sites,
//     eps(y) = +1 on even sites, -1 on odd sites,
// and for a Naik link the same with y+3mu (O3).  In PyTorch's convention for
// complex leaves, dS/dRe U + i dS/dIm U = 2 O^dag.
//
// Device work per call: Y_k by the FP64 (FP32) Dslash (groups of 2..8 terms
// as one multi-RHS Dslash, then unpacked), then k_fermion_oprod,
// one work-item per (site, mu, parity), accumulating groups of OPROD_GROUP
// terms in FP64 registers; outputs stay in HBM in a coalesced layout
// [parity][element][site][mu] (double2) and are exported to the raw MILC
// layout only when the caller asks for host arrays.
#include "dslash_kernel.h"
#include "handles.h"
#include "qphix_mi355x.h"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace qphix {

constexpr int OPROD_GROUP = 8;

template <typename T> struct OprodArgs {
  const T *X[OPROD_GROUP];  // even halves
  const T *Y[OPROD_GROUP];  // odd halves, D_oe X
  double c[OPROD_GROUP];
  int nt;                   // terms in this group (1..OPROD_GROUP)
  int accumulate;           // add to the outputs instead of overwriting them
  double2 *o1, *o3;         // oprod_index layout
  Geo g;
  SiteOrder order;          // traversal order of the sites
};

// Raw opposite-parity indices of y + mu and y + 3mu for site i of parity cb.
__device__ __forceinline__ void fwd_hops(const Geo &g, int i, int cb, int mu, int &n1, int &n3) {
  int xh, y, z, t;
  site_coords(g, i, xh, y, z, t);
  if (mu == 0) {
    const int xodd = (y + z + t + cb) & 1;
    const int row = i - xh;
    n1 = row + wrapc(xh + xodd, g.nxh);
    n3 = row + wrapc(xh + xodd + 1, g.nxh);
    return;
  }
  const int c = mu == 1 ? y : mu == 2 ? z : t;
  const int L = mu == 1 ? g.ny : mu == 2 ? g.nz : g.nt;
  const int s = mu == 1 ? g.nxh : mu == 2 ? g.nxy : g.nxyz;
  const int P = mu == 1 ? g.nxy : mu == 2 ? g.nxyz : g.vh;
  n1 = i + s - (c + 1 >= L ? P : 0);
  n3 = i + 3 * s - (c + 3 >= L ? P : 0);
}

// Work-item t of parity cb handles site t / 4 and direction t % 4: the four
// directions of a site sit in neighbouring lanes, so psi(y) is loaded once per
// site from L1 and the four outputs of each element are one contiguous 64 B
// store.
// Layout [cb][tile of 16 sites][element][site % 16][mu]: the 64 work-items of
// 16 consecutive sites write one contiguous 9 KB record per output.
constexpr int OPROD_TS = 16;
__host__ __device__ __forceinline__ size_t oprod_tiles(int vh) { return (size_t)(vh + OPROD_TS - 1) / OPROD_TS; }
__device__ __forceinline__ size_t oprod_index(int cb, int e, int i, int mu, int vh) {
  return (((size_t)cb * oprod_tiles(vh) + i / OPROD_TS) * 9 + e) * (4 * OPROD_TS) + (i % OPROD_TS) * 4 + mu;
}
// double2 elements of one output (both parities)
static inline size_t oprod_len(int vh) { return 2 * oprod_tiles(vh) * 9 * 4 * OPROD_TS; }

// Outputs are written once and read much later (the next group or the export),
// so they are stored non-temporally.
__device__ __forceinline__ void oprod_store(double2 *o, size_t k, const cplx<double> &v) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  __builtin_nontemporal_store(d2{v.re, v.im}, reinterpret_cast<d2 *>(o + k));
}

template <typename T>
__global__ void __launch_bounds__(256) k_fermion_oprod(const OprodArgs<T> a) {
  const int vh = a.g.vh;
  // XCD x (blocks are dispatched round-robin over the 8 XCDs) sweeps the x-th
  // contiguous slab of t-slices; with a.order on it walks the slab one z-plane
  // column at a time (every t of the slab, then the next plane), so the
  // t + 1, t + 3 spinors a site gathers were loaded a few planes earlier and
  // are still in that XCD's L2
  const int slot = block_slot(OPT_XCD);
  const int t = slot * blockDim.x + threadIdx.x;
  const int L = t >> 2, mu = t & 3, cb = blockIdx.y;
  if (L >= vh) return;
  const int i = ordered_site(a.g, a.order, L);
  int n1, n3;
  fwd_hops(a.g, i, cb, mu, n1, n3);
  cplx<double> o1[9], o3[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    if (a.accumulate) {
      const double2 u = a.o1[oprod_index(cb, e, i, mu, vh)], w = a.o3[oprod_index(cb, e, i, mu, vh)];
      o1[e] = cplx<double>{u.x, u.y};
      o3[e] = cplx<double>{w.x, w.y};
    } else {
      o1[e] = cplx<double>{0.0, 0.0};
      o3[e] = cplx<double>{0.0, 0.0};
    }
  }
  const double eps = cb ? -1.0 : 1.0;
#pragma unroll
  for (int k = 0; k < OPROD_GROUP; ++k) {
    if (k >= a.nt) break;
    const T *here = cb ? a.Y[k] : a.X[k];
    const T *other = cb ? a.X[k] : a.Y[k];
    cplx<T> b[3], f1[3], f3[3];
    Spin<T>::load(here, i, b);
    Spin<T>::load(other, n1, f1);
    Spin<T>::load(other, n3, f3);
    const double c = eps * a.c[k];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      // c * conj(psi(y)_q)
      const double br = c * (double)b[q].re, bi = -c * (double)b[q].im;
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const double ar = f1[r].re, ai = f1[r].im, dr = f3[r].re, di = f3[r].im;
        o1[3 * r + q].re += ar * br - ai * bi;
        o1[3 * r + q].im += ar * bi + ai * br;
        o3[3 * r + q].re += dr * br - di * bi;
        o3[3 * r + q].im += dr * bi + di * br;
      }
    }
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    oprod_store(a.o1, oprod_index(cb, e, i, mu, vh), o1[e]);
    oprod_store(a.o3, oprod_index(cb, e, i, mu, vh), o3[e]);
  }
}

// SoA outputs of one parity -> raw MILC order [site][mu][row][col][re,im].
template <typename T>
__global__ void __launch_bounds__(256) k_oprod_export(const double2 *o, int cb, int vh, T *raw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  for (int mu = 0; mu < 4; ++mu)
    for (int e = 0; e < 9; ++e) {
      const double2 v = o[oprod_index(cb, e, i, mu, vh)];
      raw[72 * (size_t)i + 18 * mu + 2 * e] = (T)v.x;
      raw[72 * (size_t)i + 18 * mu + 2 * e + 1] = (T)v.y;
    }
}

// z-plane columns through each XCD's t-slab when nt divides over the XCDs
// (QPHIX_OPROD_ORDER=0: plain site order).
static SiteOrder oprod_order() {
  const Geometry &g = ctx().g;
  SiteOrder o{0, g.ny, 1, 0};
  const char *e = std::getenv("QPHIX_OPROD_ORDER");
  if ((e && std::atoi(e) == 0) || g.nt % N_XCD) return o;
  o.tslab = g.nt / N_XCD;
  o.on = 1;
  return o;
}

// Device part: Y_k and the outer products, into o1/o3 (SoA, both parities).
// Returns false (error recorded) on allocation failure.
template <typename T>
static bool oprod_device(const LinksT<T> &L, int nterms, const double *coef,
                         typename Handles<T>::V *const *X, double2 *o1, double2 *o3, float *kernel_ms) {
  const Geometry &gg = ctx().g;
  const int vh = gg.vh;
  const size_t nv = 6 * (size_t)vh;
  // packed inputs and outputs of the multi-RHS Dslash, then its source gates
  T *ybuf = (T *)ws2_get(2 * OPROD_GROUP * nv * sizeof(T) + OPROD_GROUP * sizeof(SolverState) + 256);
  if (!ybuf) return false;
  T *ypack = ybuf + OPROD_GROUP * nv;
  SolverState *yst = (SolverState *)(((uintptr_t)(ypack + OPROD_GROUP * nv) + 255) & ~(uintptr_t)255);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (kernel_ms) {
    *kernel_ms = 0.f;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
  }
  if (nterms == 0) {
    HIP_CHECK(hipMemsetAsync(o1, 0, oprod_len(vh) * sizeof(double2), ctx().stream));
    HIP_CHECK(hipMemsetAsync(o3, 0, oprod_len(vh) * sizeof(double2), ctx().stream));
  }
  for (int k0 = 0; k0 < nterms; k0 += OPROD_GROUP) {
    const int nt = std::min(OPROD_GROUP, nterms - k0);
    OprodArgs<T> a{};
    for (int k = 0; k < nt; ++k) {
      a.X[k] = X[k0 + k]->half[0];
      a.c[k] = coef[k0 + k];
    }
    if (nt == 1) {  // Y = D_oe X
      DslashExtra ex;
      launch_dslash<T>(L, 1, ybuf, a.X[0], EPI_PLAIN, ex);
      a.Y[0] = ybuf;
    } else {
      // the group's Y_k as one multi-RHS Dslash (links read once for all
      // terms), then back to separate vectors in the packed inputs' buffer:
      // read interleaved, a wavefront's Y loads at one term would stride
      // over the sources (measured: 48^3x96, 64 terms, force 152 -> 109 ms)
      const int nr = mrhs_dslash_packed<T>(L, 1, a.X, nt, ypack, ybuf, yst);
      mrhs_unpack<T>(ybuf, ypack, nt, nr);
      for (int k = 0; k < nt; ++k) a.Y[k] = ypack + (size_t)k * nv;
    }
    a.nt = nt;
    a.accumulate = k0 > 0;
    a.o1 = o1;
    a.o3 = o3;
    a.g = Geo{gg.nxh, gg.ny, gg.nz, gg.nt, gg.vh, gg.nxy, gg.nxyz};
    a.order = oprod_order();
    if (kernel_ms) HIP_CHECK(hipEventRecord(e0, ctx().stream));
    hipLaunchKernelGGL(k_fermion_oprod<T>, dim3((unsigned)((4 * (size_t)vh + 255) / 256), 2), dim3(256), 0,
                       ctx().stream, a);
    HIP_CHECK(hipGetLastError());
    if (kernel_ms) {
      HIP_CHECK(hipEventRecord(e1, ctx().stream));
      HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      *kernel_ms += ms;
    }
  }
  if (kernel_ms) {
    HIP_CHECK(hipEventDestroy(e0));
    HIP_CHECK(hipEventDestroy(e1));
  }
  return true;
}

template <typename T>
static bool oprod_check(typename Handles<T>::L *L, int nterms, const double *coef,
                        typename Handles<T>::V *const *X) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called: asqtad_fermion_force", false);
  QPHIX_REQUIRE(L && nterms >= 0 && (nterms == 0 || (coef && X)), "asqtad_fermion_force: bad arguments", false);
  QPHIX_REQUIRE(L->fat[0] && L->fat[1], "asqtad_fermion_force: both link parities are required", false);
  for (int k = 0; k < nterms; ++k)
    QPHIX_REQUIRE(X[k] && X[k]->half[0], "asqtad_fermion_force: solution lacks its even half", false);
  return true;
}

template <typename T>
static void fermion_force(QPHIX_info_t *info, typename Handles<T>::L *L, int nterms, const double *coef,
                          typename Handles<T>::V *const *X, T *oprod_fat, T *oprod_long) {
  const auto t0 = std::chrono::steady_clock::now();
  if (info) info->status = QPHIX_FAIL;
  if (!oprod_check<T>(L, nterms, coef, X)) return;
  const int vh = ctx().g.vh;
  double2 *o = (double2 *)dev_alloc(2 * oprod_len(vh) * sizeof(double2));
  if (!o) {
    if (info) info->status = QPHIX_MEM_ERROR;
    return;
  }
  double2 *o1 = o, *o3 = o + oprod_len(vh);
  if (!oprod_device<T>(*L, nterms, coef, X, o1, o3, nullptr)) {
    dev_free(o);
    if (info) info->status = QPHIX_MEM_ERROR;
    return;
  }
  const size_t nraw = 72 * (size_t)vh;
  T *stage = (T *)dev_alloc(nraw * sizeof(T));
  if (!stage) {
    dev_free(o);
    if (info) info->status = QPHIX_MEM_ERROR;
    return;
  }
  for (int which = 0; which < 2; ++which) {
    T *dst = which ? oprod_long : oprod_fat;
    if (!dst) continue;
    for (int cb = 0; cb < 2; ++cb) {
      hipLaunchKernelGGL(k_oprod_export<T>, dim3((vh + 255) / 256), dim3(256), 0, ctx().stream,
                         which ? o3 : o1, cb, vh, stage);
      HIP_CHECK(hipMemcpyAsync(dst + (cb ? nraw : 0), stage, nraw * sizeof(T), hipMemcpyDeviceToHost,
                               ctx().stream));
      HIP_CHECK(hipStreamSynchronize(ctx().stream));
    }
  }
  HIP_CHECK(hipStreamSynchronize(ctx().stream));
  dev_free(stage);
  dev_free(o);
  if (info) {
    info->final_sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // per term and site: 1146 flops of D_oe (odd sites only) + 2 x 4 x 9 complex MACs
    info->final_flop = (double)nterms * vh * (1146.0 + 2.0 * 144.0 * 2.0);
    info->status = QPHIX_SUCCESS;
  }
}

template <typename T>
static double fermion_force_time(typename Handles<T>::L *L, int nterms, const double *coef,
                                 typename Handles<T>::V *const *X, int reps, double *oprod_ms) {
  if (!oprod_check<T>(L, nterms, coef, X)) return -1.0;
  const int vh = ctx().g.vh;
  double2 *o = (double2 *)dev_alloc(2 * oprod_len(vh) * sizeof(double2));
  if (!o) return -1.0;
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  oprod_device<T>(*L, nterms, coef, X, o, o + oprod_len(vh), nullptr);  // warm-up
  HIP_CHECK(hipEventRecord(e0, ctx().stream));
  for (int r = 0; r < reps; ++r) oprod_device<T>(*L, nterms, coef, X, o, o + oprod_len(vh), nullptr);
  HIP_CHECK(hipEventRecord(e1, ctx().stream));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  if (oprod_ms) {
    double acc = 0;
    for (int r = 0; r < reps; ++r) {
      float k = 0.f;
      oprod_device<T>(*L, nterms, coef, X, o, o + oprod_len(vh), &k);
      acc += k;
    }
    *oprod_ms = acc / reps;
  }
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  HIP_CHECK(hipStreamSynchronize(ctx().stream));
  dev_free(o);
  return (double)ms / reps;
}

} // namespace qphix

using namespace qphix;

extern "C" {

#define QPHIX_FF_API(P, T)                                                                       \
  void QPHIX_##P##3_asqtad_fermion_force_oprod(QPHIX_info_t *info,                              \
                                               QPHIX_##P##3_FermionLinksAsqtad *L, int nterms,  \
                                               const double *coef, QPHIX_##P##3_ColorVector *X[], \
                                               T *oprod_fat, T *oprod_long) {                    \
    fermion_force<T>(info, L, nterms, coef, X, oprod_fat, oprod_long);                           \
  }                                                                                              \
  double QPHIX_##P##3_asqtad_fermion_force_time(QPHIX_##P##3_FermionLinksAsqtad *L, int nterms, \
                                                const double *coef, QPHIX_##P##3_ColorVector *X[], \
                                                int reps, double *oprod_ms) {                    \
    return fermion_force_time<T>(L, nterms, coef, X, reps < 1 ? 1 : reps, oprod_ms);            \
  }

QPHIX_FF_API(D, double)
QPHIX_FF_API(F, float)

} // extern "C"
