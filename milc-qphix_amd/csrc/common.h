// common.h -- internal declarations shared by the HIP sources of
// libqphixmilc_hip.so.  Not installed; callers see include/qphix*.h only.
#pragma once
#include <type_traits>

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

extern "C" {
#include <qphix_int.h>
#include <qphix_f3.h>
#include <qphix_d3.h>
}

namespace qphix {

// ---------------------------------------------------------------------------
// Error handling.  Contract violations and HIP failures either abort the
// process (the reference's MYASSERT behaviour, ks_boundary.h:35) or, when the
// caller switched abort-on-error off, are recorded for QPHIX_MI355X_last_error.
// ---------------------------------------------------------------------------
void fail(const char *file, int line, const char *msg);
bool failed();

#define QPHIX_CHECK(cond, msg)                                   \
  do {                                                           \
    if (!(cond)) ::qphix::fail(__FILE__, __LINE__, msg);         \
  } while (0)

// Fails and returns from the calling function (optional return value).
#define QPHIX_REQUIRE(cond, msg, ...)                            \
  do {                                                           \
    if (!(cond)) {                                               \
      ::qphix::fail(__FILE__, __LINE__, msg);                    \
      return __VA_ARGS__;                                        \
    }                                                            \
  } while (0)

#define HIP_CHECK(expr)                                          \
  do {                                                           \
    hipError_t e_ = (expr);                                      \
    if (e_ != hipSuccess) ::qphix::fail(__FILE__, __LINE__, hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// Geometry of the single global lattice (one lattice per process, as in the
// reference, ks_globals.h:14-43).  Site index within a parity is MILC's raw
// index i = xh + nxh*(y + ny*(z + nz*t)); full x = 2*xh + ((y+z+t+cb)&1).
// ---------------------------------------------------------------------------
struct Geometry {
  int nx, ny, nz, nt, nxh;
  int vh;   // sites per parity
  int nxy;  // nxh*ny   (stride of z)
  int nxyz; // nxh*ny*nz (stride of t)
};

constexpr size_t PINNED_BYTES = 1 << 20;  // pinned host staging for solver state reads

struct Context {
  bool initialised = false;
  Geometry g{};
  int device = 0;
  hipStream_t stream = nullptr;
  bool sync_calls = true;    // every API call synchronises before returning
  bool graphs = true;        // batched solvers replay captured hipGraphs
  bool abort_on_error = true;
  int n_cu = 0;
  // reduction scratch (device): per-block partial sums
  double *partials = nullptr;
  int max_blocks = 0;
  size_t partials_len = 0;   // doubles in `partials`
  // pinned host mirror for small scalar reads
  void *pinned = nullptr;
  // workspace cache for solver temporaries
  void *ws = nullptr;
  size_t ws_bytes = 0;
  void *ws2 = nullptr;
  size_t ws2_bytes = 0;
};

Context &ctx();
bool require_init();
void sync_if_needed();
void *ws_get(size_t bytes);  // grows; valid until the next ws_get with more bytes
void dev_free(void *p);
void *dev_alloc(size_t bytes);
// For optional buffers the caller can do without: null on failure, nothing
// recorded (so no abort under abort-on-error).
void *dev_try_alloc(size_t bytes);
int fewsums_mode();
// Most steps per update of the multi-shift solvers' shifted systems
// (QPHIX_MI355X_set_msx_fuse; 1 = every step).
int msx_fuse();
void set_fewsums_mode(int mode);
// Takes the device address of the no-allocation face scratch for `device`
// (which must be current); once per device, at init.  False: error recorded.
bool faces_scratch_setup(int device);

// ---------------------------------------------------------------------------
// Device field storage.
//
// Colour vector: per parity, vh sites x 3 complex numbers, tile-major over 64
// sites (Spin<T> and spin_cidx in device.h); create/extract reorder MILC's raw
// [site][colour][re,im] half on the device.
//
// Asqtad links of output parity cb: for each direction mu the forward link
// (slot 2mu+1) and the back link (slot 2mu) form a "pair" of NC complex numbers
// (NC = 18, or 12 when only rows 0-1 of each long link are stored, or 8 for
// the FP32 sloppy copy's long links in their 8-real form, recon8).  A pair is
// cut into 16-byte chunks (1 complex double or 2 complex float).  LINK_TILE
// consecutive sites form a tile (one wavefront's sites), and LINK_GROUP
// consecutive tiles a group; inside a group chunk k of direction mu is stored
// for all its tiles in turn, each as a row of LINK_TILE chunks:
//     elem(mu, k, i) at ((g * 4*NCH + mu*NCH + k) * LINK_GROUP + t) * LINK_TILE + i % LINK_TILE,
//     g = i / (LINK_GROUP * LINK_TILE), t = (i / LINK_TILE) % LINK_GROUP.
// Consecutive lanes load consecutive 16 B (full coalesced wavefront
// transactions), and the wavefronts of neighbouring tiles, which run at the
// same time, read chunk k of the same direction from one contiguous stretch.
// Measured at 32^4 against one group per tile (LINK_GROUP = 1, each
// wavefront streaming its own 4*NCH KiB): FP64 Dslash 172.2 -> 167.8 us,
// FP32 93.5 -> 88.9 us; groups of 4 / 16 / 256 tiles: 169.8 / 168.4 / 168.7 us.
// ---------------------------------------------------------------------------
template <typename T> struct vec16;
template <> struct vec16<double> { using type = double2; static constexpr int ncplx = 1; };
template <> struct vec16<float> { using type = float4; static constexpr int ncplx = 2; };

// 16-bit fixed-point links (storage of the sloppy operators of the mixed
// solvers, arithmetic in FP32).  Each link has a scale s = max |element| (over
// the real and imaginary parts of its stored rows) and elements
// q = round(x / s * 32767), x = q * s / 32767: |error| <= s / 65534 per
// element.  A pair (forward link, then back link) of NC complex elements is
// cut into 16-byte chunks: the 2 NC int16 values in order, 8 to a chunk, and
// the two float scales in the last two dwords of the last chunk (NC = 18:
// 5 chunks, 80 B; NC = 12: 4 chunks, 64 B).
struct Q16 {};
template <> struct vec16<Q16> { using type = uint4; static constexpr int ncplx = 0; };
// chunks of a link pair with nc complex elements
template <typename TL> __host__ __device__ constexpr int link_nch(int nc) {
  return std::is_same<TL, Q16>::value ? (nc == 12 ? 4 : 5) : nc / vec16<TL>::ncplx;
}
// arithmetic type of a link storage type
template <typename TL> struct link_compute { using type = TL; };
template <> struct link_compute<Q16> { using type = float; };

#ifndef QPHIX_LINK_TILE
#define QPHIX_LINK_TILE 64
#endif
constexpr int LINK_TILE = QPHIX_LINK_TILE;

constexpr int LINK_GROUP = 64;
// chunks between chunk k and k + 1 of one site
constexpr int LINK_CSTRIDE = LINK_GROUP * LINK_TILE;

// 16-byte chunk index of chunk k of direction mu at site i (nch chunks per direction).
__host__ __device__ __forceinline__ size_t link_chunk(int mu, int k, int i, int nch) {
  const size_t g = (size_t)(i / LINK_CSTRIDE);
  const int t = (i / LINK_TILE) % LINK_GROUP;
  return ((g * (4 * nch) + mu * nch + k) * LINK_GROUP + t) * LINK_TILE + (i % LINK_TILE);
}
// Chunk index of chunk 0 of direction 0 for lane 0 of site i's tile: chunk k of
// direction mu of site i is at link_base(i, nch) + (mu * nch + k) * LINK_CSTRIDE + i % LINK_TILE.
__host__ __device__ __forceinline__ size_t link_base(int i, int nch) {
  return link_chunk(0, 0, i, nch) - (i % LINK_TILE);
}
// Sites of link storage per parity (vh rounded up to whole groups).
__host__ __device__ __forceinline__ size_t link_sites(int vh) {
  return (size_t)((vh + LINK_CSTRIDE - 1) / LINK_CSTRIDE) * LINK_CSTRIDE;
}

template <typename T> struct LinksT {
  T *fat[2] = {nullptr, nullptr};
  T *lng[2] = {nullptr, nullptr};
  uint8_t *lsign[2] = {nullptr, nullptr};  // 12- and 8-real row-2 sign, bit (2mu+fwd)
  int recon = 18;                          // 18 or 12 (long links); 8: FP32 sloppy copy only
  int parity = 0;                          // QPHIX_EVEN/ODD/EVENODD supplied
  LinksT<float> *sloppy = nullptr;         // lazily built FP32 copy (D3 only)
  LinksT<Q16> *sloppy16 = nullptr;         // lazily built 16-bit copy (D3 only)
  LinksT<float> *sloppy8 = nullptr;        // lazily built FP32 copy, 8-real long links (D3 only)
  bool no8 = false;                        // some long link has no 8-real form
};

template <typename T> struct VectorT {
  T *half[2] = {nullptr, nullptr};  // [0] even sites, [1] odd sites
  int parity = 0;
};

// Internal gauge field (the reference keeps 8 slots per even site,
// qphix_d3.cpp:177-343) and momentum (8 reals per link, C13 mapping): the
// gauge field is complex-major per parity (gauge_elem), the momentum in raw
// per-parity order mom[cb][i][mu][8].
// Gauge links per parity, complex-major: element e (row-major 3x3) of the link
// in direction mu at site i is the complex number at index (mu*9 + e)*vh + i,
// so lanes at consecutive sites read consecutive complexes.
// u12 (null unless every link is SU(3) to rounding) holds rows 0-1 only, the
// same way with 6 elements per link (gauge_elem_n with ne = 6); the force
// kernel rebuilds row 2 as conj(r0 x r1).
template <typename T> struct GaugeT {
  T *u[2] = {nullptr, nullptr};
  T *u12[2] = {nullptr, nullptr};
  int parity = 0;
};
__host__ __device__ __forceinline__ size_t gauge_elem_n(int mu, int e, int i, int vh, int ne) {
  return (size_t)(mu * ne + e) * vh + i;
}
__host__ __device__ __forceinline__ size_t gauge_elem(int mu, int e, int i, int vh) {
  return gauge_elem_n(mu, e, i, vh, 9);
}
template <typename T> struct ForceT {
  T *h[2] = {nullptr, nullptr};
  int parity = 0;
};

// Half-precision colour vector used by the FP16-stored-spinor solver: 6 halves
// of a per-site-normalised vector plus the float norm, 16 bytes per site.
struct alignas(16) HalfSite {
  __half v[6];
  float norm;
};

// ---------------------------------------------------------------------------
// Kernel launchers (defined in the .hip files).
// ---------------------------------------------------------------------------
enum DslashEpilogue {
  EPI_PLAIN = 0,  // out = D in
  EPI_PKP = 1,    // out = D in + s*aux ; reduce aux.out
  EPI_RESTART = 2, // r = b + (D in + s*aux) ; p = r ; reduce |r|^2 (+ rel)
  EPI_PKP_FS = 3   // out = D in + s*aux ; reduce aux.out, r.out, out.out, r.r
                   // (r read-only); block 0 snapshots the gate and scalars
                   // for the single-reduction CG step that follows
};

struct DslashExtra {
  double s = 0;                // scalar of the epilogue (-msq_x4 or -sigma)
  const void *aux = nullptr;   // p (PKP) or x (RESTART)
  const void *b = nullptr;     // RESTART: source
  void *r = nullptr;           // RESTART: residual out; PKP_FS: residual in
  void *p = nullptr;           // RESTART: direction out
  void *state = nullptr;       // device SolverState updated by the last block
  int scalar_op = 0;           // ScalarOp run by the last block
  const int *gate = nullptr;   // kernel is a no-op when *gate == 0
  bool consumer_finalize = false;  // reducing epilogues: leave the partials to the next kernel
};

template <typename T>
void launch_dslash(const LinksT<T> &L, int cb, T *out, const T *in, int epi, const DslashExtra &ex);

void launch_dslash_half(const LinksT<float> &L, int cb, HalfSite *out_h, float *out_f,
                        const HalfSite *in, int epi, const DslashExtra &ex);

// 16-bit fixed-point links (Q16) with FP32 or FP16-stored spinors and FP32
// arithmetic: the sloppy operator of the mixed solvers.
void launch_dslash_q16(const LinksT<Q16> &L, int cb, float *out, const float *in, int epi,
                       const DslashExtra &ex);
void launch_dslash_q16_half(const LinksT<Q16> &L, int cb, HalfSite *out_h, float *out_f,
                            const HalfSite *in, int epi, const DslashExtra &ex);

// FP32 copy of FP64 links in the same layout (fields.hip); false (error
// recorded, nothing left allocated) when device memory runs out
bool links_to_float(const LinksT<double> &Ld, LinksT<float> &Lf);
// 16-bit fixed-point copy of FP64 links (fields.hip); false as above
bool links_to_q16(const LinksT<double> &Ld, LinksT<Q16> &Lq);
// The lazily built sloppy copies of a D3 link set (fields.hip); null (error
// recorded) when they cannot be allocated
const LinksT<float> *sloppy_links(LinksT<double> *L);
const LinksT<float> *sloppy_links_solver(LinksT<double> *L);
const LinksT<Q16> *sloppy_links16(LinksT<double> *L);
// Link bits of the mixed solvers' sloppy operator: 32 (default) or 16.
int sloppy_link_bits();

struct SolverState;
// D on output parity cb of n (2..8) vectors src[r] (the other parity's halves)
// as one multi-RHS Dslash (multishift_batch.hip): the inputs are packed into
// pin and the results left in pout, both [vh][nr][6] interleaved with nr = 2,
// 4 or 8 >= n (source r of site i at colour-vector index i * nr + r); st:
// nr device states used as the sources' gates.  Returns nr.  Queued only.
template <typename T>
int mrhs_dslash_packed(const LinksT<T> &L, int cb, const T *const *src, int n, T *pin, T *pout, SolverState *st);
// The first n sources of an interleaved [vh][nr][6] buffer into n vectors
// dst + r * 6 vh (multishift_batch.hip).  Queued only.
template <typename T> void mrhs_unpack(const T *src, T *dst, int n, int nr);

// Mixed-precision CG on one parity system (mixed.hip); out is the initial guess.
int mixed_solve(QPHIX_info_t *info, LinksT<double> *L, QPHIX_invert_arg_t *inv,
                QPHIX_resid_arg_t *res, double mass, VectorT<double> *out, VectorT<double> *in,
                int sloppy_bits);

// A batch of solver iterations captured once as a hipGraph and replayed.
// The enqueue function must launch kernels only (no memsets or copies), with
// arguments that stay fixed for the graph's lifetime; every scalar it needs
// lives in device memory.  With ctx().graphs off, run() just enqueues.
struct BatchGraph {
  hipGraphExec_t exec = nullptr;
  BatchGraph() = default;
  BatchGraph(const BatchGraph &) = delete;
  BatchGraph &operator=(const BatchGraph &) = delete;
  // use_graph = false: plain launches whatever ctx().graphs says (capture
  // and instantiation allocate; the no-allocation solver must not)
  template <typename F> void run(F &&enqueue, bool use_graph = true) {
    if (!ctx().graphs || !use_graph) {
      enqueue();
      return;
    }
    if (!exec) {
      hipGraph_t g = nullptr;
      HIP_CHECK(hipStreamBeginCapture(ctx().stream, hipStreamCaptureModeThreadLocal));
      enqueue();
      HIP_CHECK(hipStreamEndCapture(ctx().stream, &g));
      HIP_CHECK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
      HIP_CHECK(hipGraphDestroy(g));
    }
    HIP_CHECK(hipGraphLaunch(exec, ctx().stream));
  }
  void reset() {
    if (exec) HIP_CHECK(hipGraphExecDestroy(exec));
    exec = nullptr;
  }
  ~BatchGraph() { reset(); }
};

// Second workspace, for drivers that call other solvers while holding theirs.
void *ws2_get(size_t bytes);
// The same, null without a recorded error when the memory is not there.
void *ws2_try_get(size_t bytes);
int gauge_force_staples_mode();
template <typename T>
bool gauge_force_staples(GaugeT<T> *G, ForceT<T> *F, const QPHIX_gauge_coeffs_t *c, double eps);

// Symanzik force step mom <- mom - eps TA(W) queued on the library stream
// (gauge_force.hip); info->status = QPHIX_SUCCESS once queued.
template <typename T>
void gauge_force_apply(QPHIX_info_t *info, GaugeT<T> *G, ForceT<T> *F, QPHIX_gauge_coeffs_t *c, double eps);


} // namespace qphix
