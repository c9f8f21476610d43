// dslash_faces.hip -- the Dslash of a domain-decomposed lattice, run on one
// GPU with every cut face received from the domain itself (self-loopback).
//
// The reference splits each Dslash into a face exchange and two site sets
// (ks_long_dslash.cpp:564-668 with ks_boundary.cpp): pack_face gathers the
// 3-deep faces of the input parity for each cut direction and side
// (ks_boundary.cpp:440-556, 18 reals per face site triple), the interior sites
// are computed while the faces travel, and unpack_face adds the off-node terms
// to the boundary sites (:558-690).  Its ASSUME_MULTINODE build treats every
// direction as cut and aliases the send and receive buffers, so one process
// exchanges faces with itself (:199, :250-256).
//
// MI355X form:
//   k_face_pack   one work-item per face site: faces are stored site-major
//                 (6 reals), face f = 2 mu + side (side 0: layers x_mu = 0..2,
//                 side 1: L-3..L-1), in layer-major order;
//   k_dslash_box  interior sites (all 16 neighbours local) as the plain
//                 kernel does; then boundary sites, each computed in full by
//                 one work-item, every neighbour across a cut read from the
//                 face the neighbouring domain sends.  A boundary site's sum is
//                 formed in the plain kernel's order (pair_compute), so the
//                 result equals asqtad_dslash bit for bit -- no read-modify-
//                 write of partial sums as in the reference's unpack.
// The site sets are boxes: the interior box (layers 3..L-4 in every cut
// direction), and for each cut direction mu the box of its 6 face layers with
// the earlier cut directions restricted to their interior (disjoint boxes
// covering every boundary site once).  With several GPUs the faces would go
// to the neighbours between the pack and the boundary kernel (RCCL send/recv
// per direction over xGMI), overlapping the interior kernel.
#include "blas.h"
#include "dslash_kernel.h"
#include "handles.h"
#include "qphix_mi355x.h"

namespace qphix {

// Scratch of the no-allocation entry point (QPHIX_*3_asqtad_dslash_faces_noalloc):
// its face buffer is this file-scope device array, one copy per GPU, held
// until the process ends.  In host code `scratch` names a host-side stand-in,
// not device memory: host code uses only the address hipGetSymbolAddress
// returned for the current device (g_scratch_dev), taken once per device at
// QPHIX_init (faces_scratch_setup).  Every call on a GPU shares its array and
// finds the previous call's faces there; the pack kernels overwrite every
// face site before the boundary kernels read one, so nothing is cleared.
constexpr size_t FACE_SCRATCH_BYTES = size_t(64) << 20;
constexpr int FACE_SCRATCH_MAX_DEV = 64;
static __device__ __align__(256) unsigned char scratch[FACE_SCRATCH_BYTES];
static void *g_scratch_dev[FACE_SCRATCH_MAX_DEV] = {};

bool faces_scratch_setup(int device) {
  QPHIX_REQUIRE(device >= 0 && device < FACE_SCRATCH_MAX_DEV, "face scratch: device id out of range", false);
  if (g_scratch_dev[device]) return true;
  int cur = -1;
  QPHIX_REQUIRE(hipGetDevice(&cur) == hipSuccess && cur == device,
                "face scratch: the device is not current", false);
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
    (void)hipGetLastError();
    fail(__FILE__, __LINE__, "face scratch: hipGetSymbolAddress failed");
    return false;
  }
  g_scratch_dev[device] = p;
  return true;
}

struct FaceGeo {
  int L[4];
  int vh, cut;
  int S2[4];         // sites of one parity in one layer of direction mu
  long long off[8];  // first site of face 2 mu + side in the face buffer
};

// One box of sites of the output parity: dimension j spans [lo[j], lo[j] +
// len[j]), except dimension `layers`, which spans the 6 face layers 0, 1, 2,
// L-3, L-2, L-1.  x is enumerated in parity-halved steps when it is an
// interval; as the layers dimension every x is visited and lanes of the
// other parity stop.
struct SiteBox {
  int lo[4], len[4];
  int layers;
};

__device__ __forceinline__ bool box_site(const FaceGeo &g, const SiteBox &b, int cb, long long k, int c[4]) {
  const bool halved = b.layers != 0;
  long long r = k;
  int idx0;
  if (halved) {
    const int n0 = b.len[0] / 2;
    idx0 = (int)(r % n0);
    r /= n0;
  } else {
    idx0 = (int)(r % 6);
    r /= 6;
  }
#pragma unroll
  for (int j = 1; j < 4; ++j) {
    const int n = b.layers == j ? 6 : b.len[j];
    const int v = (int)(r % n);
    r /= n;
    c[j] = b.layers == j ? (v < 3 ? v : g.L[j] - 6 + v) : b.lo[j] + v;
  }
  if (halved) {
    c[0] = b.lo[0] + 2 * idx0 + ((b.lo[0] + c[1] + c[2] + c[3] + cb) & 1);
    return true;
  }
  c[0] = idx0 < 3 ? idx0 : g.L[0] - 6 + idx0;
  return ((c[0] + c[1] + c[2] + c[3]) & 1) == cb;
}

__device__ __forceinline__ int lex_half(const FaceGeo &g, const int c[4]) {
  return (c[0] + g.L[0] * (c[1] + g.L[1] * (c[2] + g.L[2] * c[3]))) >> 1;
}

// Position in the face buffer of site c (c[mu] already the layer coordinate
// 0..2 for side 0 or L-3..L-1 for side 1).
__device__ __forceinline__ long long face_pos(const FaceGeo &g, int mu, int side, const int c[4]) {
  int r = 0, st = 1;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (j == mu) continue;
    r += c[j] * st;
    st *= g.L[j];
  }
  const int layer = side ? c[mu] - (g.L[mu] - 3) : c[mu];
  return g.off[2 * mu + side] + (long long)layer * g.S2[mu] + (r >> 1);
}

// Pack face f = 2 mu + side of the input parity pin into the face buffer.
template <typename T>
__global__ void __launch_bounds__(256) k_face_pack(const T *in, T *faces, FaceGeo g, int mu, int side, int pin) {
  const int pos = blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= 3 * g.S2[mu]) return;
  const int layer = pos / g.S2[mu], q = pos - layer * g.S2[mu];
  int c[4];
  int r = 2 * q, first = -1;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (j == mu) continue;
    if (first < 0) first = j;
    c[j] = r % g.L[j];
    r /= g.L[j];
  }
  c[mu] = side ? g.L[mu] - 3 + layer : layer;
  if (((c[0] + c[1] + c[2] + c[3]) & 1) != pin) c[first] += 1;  // the other site of the pair
  using C = typename Spin<T>::C;
  cplx<C> v[3];
  Spin<T>::load(in, lex_half(g, c), v);
  T *o = faces + 6 * (g.off[2 * mu + side] + pos);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    o[2 * k] = v[k].re;
    o[2 * k + 1] = v[k].im;
  }
}

// Neighbour at c + h mu: a local half-site of `in`, or a site of a face.
template <typename T>
__device__ __forceinline__ void load_nbr(const FaceGeo &g, const T *in, const T *faces, const int c[4], int mu, int h,
                                         cplx<typename Spin<T>::C> *v) {
  int y[4] = {c[0], c[1], c[2], c[3]};
  const int L = g.L[mu];
  int t = y[mu] + h;
  if ((g.cut >> mu) & 1) {
    if (t >= L || t < 0) {
      // across the cut: the neighbouring domain's face -- its low layers for
      // t >= L, its high layers for t < 0 (self-loopback: our own faces)
      const int side = t >= L ? 0 : 1;
      y[mu] = t >= L ? t - L : t + L;
      const T *f = faces + 6 * face_pos(g, mu, side, y);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        v[k].re = f[2 * k];
        v[k].im = f[2 * k + 1];
      }
      return;
    }
  } else {
    t = t >= L ? t - L : (t < 0 ? t + L : t);
  }
  y[mu] = t;
  Spin<T>::load(in, lex_half(g, y), v);
}

template <typename T> struct FaceArgs {
  const typename vec16<T>::type *fat, *lng;
  const uint8_t *lsign;
  const T *in, *faces;
  T *out;
  Geo geo;  // for the interior's neighbour arithmetic
  FaceGeo g;
  SiteBox box;
  int cb;
  long long nsites;
};

// The two pairs of direction mu at site i (coordinates c): fat, then Naik,
// the order of the plain kernel.
template <typename T, int NL, bool EXT>
__device__ __forceinline__ void box_dir(const FaceArgs<T> &a, const int c[4], int i, int mu, unsigned sg,
                                        const SiteNbr &n, cplx<T> *acc) {
  constexpr int NCH_F = link_nch<T>(18), NCH_L = link_nch<T>(NL);
  const int ii = i % LINK_TILE;
  const typename vec16<T>::type *fbm = a.fat + link_base(i, NCH_F) + (size_t)mu * NCH_F * LINK_CSTRIDE;
  const typename vec16<T>::type *lbm = a.lng + link_base(i, NCH_L) + (size_t)mu * NCH_L * LINK_CSTRIDE;
  cplx<T> u1[18], vf[3], vb[3];
  PairLoad<T, 18, true>::load(fbm, LINK_CSTRIDE, ii, u1);
  if constexpr (EXT) {
    load_nbr(a.g, a.in, a.faces, c, mu, 1, vf);
    load_nbr(a.g, a.in, a.faces, c, mu, -1, vb);
  } else {
    Spin<T>::load(a.in, n.f1[mu], vf);
    Spin<T>::load(a.in, n.b1[mu], vb);
  }
  pair_compute<18, T>(u1, vf, vb, 0, 0, acc);
  cplx<T> u3[NL];
  PairLoad<T, NL, true>::load(lbm, LINK_CSTRIDE, ii, u3);
  if constexpr (EXT) {
    load_nbr(a.g, a.in, a.faces, c, mu, 3, vf);
    load_nbr(a.g, a.in, a.faces, c, mu, -3, vb);
  } else {
    Spin<T>::load(a.in, n.f3[mu], vf);
    Spin<T>::load(a.in, n.b3[mu], vb);
  }
  pair_compute<NL, T>(u3, vf, vb, (sg >> (2 * mu + 1)) & 1u, (sg >> (2 * mu)) & 1u, acc);
}

// EXT = false: the interior box (neighbours by the plain kernel's index
// arithmetic, none of them across a cut); true: a boundary box (its direction
// loop is kept rolled: unrolled, the neighbour resolution spilled ~1 KB per
// lane).
template <typename T, int NL, bool EXT>
__global__ void __launch_bounds__(DSLASH_BLOCK) k_dslash_box(const FaceArgs<T> a) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.nsites) return;
  int c[4];
  if (!box_site(a.g, a.box, a.cb, k, c)) return;
  const int i = lex_half(a.g, c);
  cplx<T> acc[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) acc[q] = cplx<T>{0, 0};
  unsigned sg = 0;
  if constexpr (NL == 12) sg = a.lsign[i];
  SiteNbr n;
  if constexpr (!EXT) {
    site_neighbours(a.geo, i, a.cb, n);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) box_dir<T, NL, EXT>(a, c, i, mu, sg, n, acc);
  } else {
#pragma unroll 1
    for (int mu = 0; mu < 4; ++mu) box_dir<T, NL, EXT>(a, c, i, mu, sg, n, acc);
  }
  Spin<T>::store_nt(a.out, i, acc);
}

static FaceGeo face_geo(int cut) {
  const Geometry &gg = ctx().g;
  FaceGeo g{};
  g.L[0] = gg.nx; g.L[1] = gg.ny; g.L[2] = gg.nz; g.L[3] = gg.nt;
  g.vh = gg.vh;
  g.cut = cut;
  long long off = 0;
  for (int mu = 0; mu < 4; ++mu) {
    int s = 1;
    for (int j = 0; j < 4; ++j)
      if (j != mu) s *= g.L[j];
    g.S2[mu] = s / 2;
    for (int side = 0; side < 2; ++side) {
      g.off[2 * mu + side] = off;
      if ((cut >> mu) & 1) off += 3LL * g.S2[mu];
    }
  }
  return g;
}

static long long face_sites(const FaceGeo &g) {
  long long n = 0;
  for (int mu = 0; mu < 4; ++mu)
    if ((g.cut >> mu) & 1) n += 6LL * g.S2[mu];
  return n;
}

static long long box_lanes(const SiteBox &b) {
  long long n = b.layers == 0 ? 6 : b.len[0] / 2;
  for (int j = 1; j < 4; ++j) n *= b.layers == j ? 6 : b.len[j];
  return n;
}

template <typename T, int NL>
static void launch_faces(const LinksT<T> &L, int cb, T *out, const T *in, T *faces, int cut) {
  const FaceGeo g = face_geo(cut);
  const Geometry &gg = ctx().g;
  for (int mu = 0; mu < 4; ++mu) {
    if (!((cut >> mu) & 1)) continue;
    for (int side = 0; side < 2; ++side)
      hipLaunchKernelGGL(k_face_pack<T>, dim3((unsigned)((3 * g.S2[mu] + 255) / 256)), dim3(256), 0,
                         ctx().stream, in, faces, g, mu, side, 1 - cb);
  }
  FaceArgs<T> a{};
  a.fat = reinterpret_cast<const typename vec16<T>::type *>(L.fat[cb]);
  a.lng = reinterpret_cast<const typename vec16<T>::type *>(L.lng[cb]);
  a.lsign = L.lsign[cb];
  a.in = in;
  a.faces = faces;
  a.out = out;
  a.geo = Geo{gg.nxh, gg.ny, gg.nz, gg.nt, gg.vh, gg.nxy, gg.nxyz};
  a.g = g;
  a.cb = cb;
  // interior
  SiteBox in_box{};
  in_box.layers = -1;
  bool empty = false;
  for (int j = 0; j < 4; ++j) {
    const bool c = (cut >> j) & 1;
    in_box.lo[j] = c ? 3 : 0;
    in_box.len[j] = c ? g.L[j] - 6 : g.L[j];
    if (in_box.len[j] <= 0) empty = true;
  }
  if (!empty) {
    a.box = in_box;
    a.nsites = box_lanes(in_box);
    hipLaunchKernelGGL((k_dslash_box<T, NL, false>), dim3((unsigned)((a.nsites + DSLASH_BLOCK - 1) / DSLASH_BLOCK)),
                       dim3(DSLASH_BLOCK), 0, ctx().stream, a);
  }
  // boundary boxes
  for (int d = 0; d < 4; ++d) {
    if (!((cut >> d) & 1)) continue;
    SiteBox b{};
    b.layers = d;
    bool none = false;
    for (int j = 0; j < 4; ++j) {
      const bool inner = j < d && ((cut >> j) & 1);
      b.lo[j] = inner ? 3 : 0;
      b.len[j] = inner ? g.L[j] - 6 : g.L[j];
      if (j != d && b.len[j] <= 0) none = true;
    }
    if (none) continue;
    a.box = b;
    a.nsites = box_lanes(b);
    hipLaunchKernelGGL((k_dslash_box<T, NL, true>), dim3((unsigned)((a.nsites + DSLASH_BLOCK - 1) / DSLASH_BLOCK)),
                       dim3(DSLASH_BLOCK), 0, ctx().stream, a);
  }
  HIP_CHECK(hipGetLastError());
}

// Argument checks shared by both entry points (errors recorded).
template <typename T>
static bool faces_args_ok(LinksT<T> *L, VectorT<T> *out, VectorT<T> *in, int parity, int cut) {
  QPHIX_REQUIRE(L && out && in, "asqtad_dslash_faces: null argument", false);
  QPHIX_REQUIRE(parity == QPHIX_EVEN || parity == QPHIX_ODD, "asqtad_dslash_faces: parity must be EVEN or ODD",
                false);
  const int cb = parity == QPHIX_EVEN ? 0 : 1;
  QPHIX_REQUIRE(out->half[cb] && in->half[1 - cb] && L->fat[cb] && L->lng[cb],
                "asqtad_dslash_faces: vectors or links of the needed parity missing", false);
  const Geometry &gg = ctx().g;
  const int Ls[4] = {gg.nx, gg.ny, gg.nz, gg.nt};
  for (int mu = 0; mu < 4; ++mu)
    QPHIX_REQUIRE(!((cut >> mu) & 1) || Ls[mu] >= 8,
                  "asqtad_dslash_faces: a cut direction needs an extent of at least 8 (two disjoint 3-deep faces)",
                  false);
  return true;
}

template <typename T> static void run_faces(LinksT<T> *L, VectorT<T> *out, VectorT<T> *in, int cb, T *faces, int cut) {
  if (L->recon == 12) launch_faces<T, 12>(*L, cb, out->half[cb], in->half[1 - cb], faces, cut);
  else launch_faces<T, 18>(*L, cb, out->half[cb], in->half[1 - cb], faces, cut);
}

template <typename T>
static double dslash_faces(LinksT<T> *L, VectorT<T> *out, VectorT<T> *in, int parity, int cut, int reps) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called: asqtad_dslash_faces", -1.0);
  QPHIX_REQUIRE(cut >= 0 && cut <= 15, "asqtad_dslash_faces: cut mask must be within 0..15", -1.0);
  if (!faces_args_ok(L, out, in, parity, cut)) return -1.0;
  const int cb = parity == QPHIX_EVEN ? 0 : 1;
  const FaceGeo g = face_geo(cut);
  T *faces = nullptr;
  if (face_sites(g) > 0) {
    faces = (T *)ws2_get(sizeof(T) * 6 * (size_t)face_sites(g));
    if (!faces) return -1.0;  // error recorded
  }
  auto run = [&]() { run_faces<T>(L, out, in, cb, faces, cut); };
  if (reps <= 0) {
    run();
    sync_if_needed();
    return 0.0;
  }
  run();  // warm-up
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  HIP_CHECK(hipEventRecord(e0, ctx().stream));
  for (int r = 0; r < reps; ++r) run();
  HIP_CHECK(hipEventRecord(e1, ctx().stream));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  return (double)ms / reps;
}

// The same Dslash with the face buffer in the device array `scratch`: no
// device memory is allocated or freed, no library routine that allocates is
// called, and nothing is timed.  Faces beyond FACE_SCRATCH_BYTES are rejected
// before anything is launched.
template <typename T>
static int dslash_faces_noalloc(LinksT<T> *L, VectorT<T> *out, VectorT<T> *in, int parity, int cut) {
  QPHIX_REQUIRE(require_init(), "QPHIX_init has not been called: asqtad_dslash_faces_noalloc", (int)QPHIX_FAIL);
  QPHIX_REQUIRE(cut >= 0 && cut <= 15, "asqtad_dslash_faces_noalloc: cut mask must be within 0..15",
                (int)QPHIX_FAIL);
  const FaceGeo g = face_geo(cut);
  const size_t need = sizeof(T) * 6 * (size_t)face_sites(g);
  if (need > FACE_SCRATCH_BYTES) {
    char msg[200];
    std::snprintf(msg, sizeof(msg),
                  "asqtad_dslash_faces_noalloc: the faces of this lattice and cut mask need %zu bytes, more than "
                  "the %zu-byte scratch (use asqtad_dslash_faces)",
                  need, FACE_SCRATCH_BYTES);
    fail(__FILE__, __LINE__, msg);
    return QPHIX_FAIL;
  }
  if (!faces_args_ok(L, out, in, parity, cut)) return QPHIX_FAIL;
  const int dev = ctx().device;
  QPHIX_REQUIRE(dev >= 0 && dev < FACE_SCRATCH_MAX_DEV && g_scratch_dev[dev],
                "asqtad_dslash_faces_noalloc: no scratch address for this device (QPHIX_init failed?)",
                (int)QPHIX_FAIL);
  run_faces<T>(L, out, in, parity == QPHIX_EVEN ? 0 : 1, static_cast<T *>(g_scratch_dev[dev]), cut);
  sync_if_needed();
  return QPHIX_SUCCESS;
}

} // namespace qphix

using namespace qphix;

extern "C" {

int QPHIX_D3_asqtad_dslash_faces_noalloc(QPHIX_D3_FermionLinksAsqtad *L, QPHIX_D3_ColorVector *out,
                                         QPHIX_D3_ColorVector *in, QPHIX_evenodd_t parity, int cut_mask) {
  return dslash_faces_noalloc<double>(L, out, in, parity, cut_mask);
}
int QPHIX_F3_asqtad_dslash_faces_noalloc(QPHIX_F3_FermionLinksAsqtad *L, QPHIX_F3_ColorVector *out,
                                         QPHIX_F3_ColorVector *in, QPHIX_evenodd_t parity, int cut_mask) {
  return dslash_faces_noalloc<float>(L, out, in, parity, cut_mask);
}

double QPHIX_D3_asqtad_dslash_faces(QPHIX_D3_FermionLinksAsqtad *L, QPHIX_D3_ColorVector *out,
                                    QPHIX_D3_ColorVector *in, QPHIX_evenodd_t parity, int cut_mask, int reps) {
  return dslash_faces<double>(L, out, in, parity, cut_mask, reps);
}
double QPHIX_F3_asqtad_dslash_faces(QPHIX_F3_FermionLinksAsqtad *L, QPHIX_F3_ColorVector *out,
                                    QPHIX_F3_ColorVector *in, QPHIX_evenodd_t parity, int cut_mask, int reps) {
  return dslash_faces<float>(L, out, in, parity, cut_mask, reps);
}

} // extern "C"
