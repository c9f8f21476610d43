// context.hip -- process-global state, init/finalize, errors, device memory.
//
// Replaces the reference's QPHIX_init/QPHIX_finalize (qphix_misc.cpp:427-594):
// the reference's thread pool, barriers and neighbour tables become one HIP
// device, one stream and on-the-fly neighbour index arithmetic.
#include "common.h"
#include <algorithm>
#include "qphix_mi355x.h"

#include <cstring>
#include <mutex>

namespace qphix {

static Context g_ctx;
static char g_last_error[512] = {0};
static int g_error_count = 0;

Context &ctx() { return g_ctx; }

void fail(const char *file, int line, const char *msg) {
  std::snprintf(g_last_error, sizeof(g_last_error), "%s:%d: %s", file, line, msg);
  ++g_error_count;
  if (g_ctx.abort_on_error) {
    std::fprintf(stderr, "QPHIX ERROR %s\n", g_last_error);
    std::fflush(stderr);
    std::exit(1);
  }
}

bool failed() { return g_error_count != 0; }

bool require_init() { return g_ctx.initialised; }

void sync_if_needed() {
  if (g_ctx.sync_calls) HIP_CHECK(hipStreamSynchronize(g_ctx.stream));
}

// Allocation-failure injection for the error-path tests: after `n` more
// successful allocations every dev_alloc fails as if device memory were
// exhausted (negative: never).
static long g_fail_alloc_after = -1;

void *dev_alloc(size_t bytes) {
  void *p = nullptr;
  if (bytes == 0) return nullptr;
  if (g_fail_alloc_after == 0) {
    fail(__FILE__, __LINE__, "hipMalloc failed (device memory exhausted?) [injected]");
    return nullptr;
  }
  if (g_fail_alloc_after > 0) --g_fail_alloc_after;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    fail(__FILE__, __LINE__, "hipMalloc failed (device memory exhausted?)");
    return nullptr;
  }
  return p;
}

// Single-reduction CG steps: bit 0 the one-precision CG, bit 1 the mixed
// solver's sloppy steps.  QPHIX_FEWSUMS in the environment sets the initial
// value; QPHIX_MI355X_set_fewsums wins.
static int g_fewsums = -1;
int fewsums_mode() {
  if (g_fewsums < 0) {
    const char *e = std::getenv("QPHIX_FEWSUMS");
    g_fewsums = e ? (std::atoi(e) & 3) : 2;
  }
  return g_fewsums;
}
void set_fewsums_mode(int mode) { g_fewsums = mode & 3; }

void *dev_try_alloc(size_t bytes) {
  void *p = nullptr;
  if (bytes == 0 || g_fail_alloc_after == 0) return nullptr;
  if (g_fail_alloc_after > 0) --g_fail_alloc_after;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();  // clear the sticky error: the caller falls back
    return nullptr;
  }
  return p;
}

void dev_free(void *p) {
  if (p) HIP_CHECK(hipFree(p));
}

void *ws_get(size_t bytes) {
  Context &c = g_ctx;
  if (bytes > c.ws_bytes) {
    HIP_CHECK(hipStreamSynchronize(c.stream));
    dev_free(c.ws);
    c.ws = dev_alloc(bytes);
    c.ws_bytes = c.ws ? bytes : 0;
  }
  return c.ws;
}

void *ws2_get(size_t bytes) {
  Context &c = g_ctx;
  if (bytes > c.ws2_bytes) {
    HIP_CHECK(hipStreamSynchronize(c.stream));
    dev_free(c.ws2);
    c.ws2 = dev_alloc(bytes);
    c.ws2_bytes = c.ws2 ? bytes : 0;
  }
  return c.ws2;
}

void *ws2_try_get(size_t bytes) {
  Context &c = g_ctx;
  if (bytes > c.ws2_bytes) {
    HIP_CHECK(hipStreamSynchronize(c.stream));
    dev_free(c.ws2);
    c.ws2 = dev_try_alloc(bytes);
    c.ws2_bytes = c.ws2 ? bytes : 0;
  }
  return c.ws2;
}

// Gauge force kernel: 1 (default) the staple-product kernel where the links
// are SU(3) to rounding, 0 always the loop-walk kernel.
static int g_gf_staples = 1;
int gauge_force_staples_mode() { return g_gf_staples; }

static QPHIX_status_t init_common(QPHIX_layout_t *layout) {
  Context &c = g_ctx;
  if (c.initialised) return QPHIX_SUCCESS;  // F and D share one lattice (C14)
  // QPHIX_GRAPHS=0 starts the process with graph replay off (read once; a
  // later QPHIX_MI355X_set_graphs call wins)
  static bool env_read = false;
  if (!env_read) {
    env_read = true;
    if (const char *e = std::getenv("QPHIX_GRAPHS")) c.graphs = std::atoi(e) != 0;
  }
  if (!layout || layout->latdim != 4 || !layout->latsize) {
    fail(__FILE__, __LINE__, "QPHIX_init: need a 4-d layout");
    return QPHIX_FAIL;
  }
  if (layout->machdim != layout->latdim) {
    fail(__FILE__, __LINE__, "QPHIX_init: machine dimension must equal lattice dimension");
    return QPHIX_FAIL;
  }
  for (int d = 0; d < layout->machdim; ++d)
    if (layout->machsize && layout->machsize[d] != 1) {
      fail(__FILE__, __LINE__, "QPHIX_init: this build runs one lattice on one GPU (machsize must be 1)");
      return QPHIX_FAIL;
    }
  const int *L = layout->latsize;
  // qphix_misc.cpp:494-497: Nxh%2, Ny%4, Nz%4, Nt%4 (minimum lattice 4^4)
  if (L[0] % 4 || L[1] % 4 || L[2] % 4 || L[3] % 4 || L[0] <= 0) {
    fail(__FILE__, __LINE__, "QPHIX_init: every lattice extent must be a positive multiple of 4");
    return QPHIX_FAIL;
  }
  long long vol = 1LL * L[0] * L[1] * L[2] * L[3];
  if (vol / 2 > (1LL << 30)) {
    fail(__FILE__, __LINE__, "QPHIX_init: lattice too large for 32-bit site indices");
    return QPHIX_FAIL;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    fail(__FILE__, __LINE__, "QPHIX_init: no HIP device visible");
    return QPHIX_FAIL;
  }
  const char *dv = std::getenv("QPHIX_DEVICE");
  c.device = dv ? std::atoi(dv) : 0;
  if (c.device < 0 || c.device >= ndev) c.device = 0;
  if (hipSetDevice(c.device) != hipSuccess) {
    fail(__FILE__, __LINE__, "QPHIX_init: hipSetDevice failed");
    return QPHIX_FAIL;
  }
  if (!faces_scratch_setup(c.device)) return QPHIX_FAIL;
  hipDeviceProp_t prop;
  HIP_CHECK(hipGetDeviceProperties(&prop, c.device));
  c.n_cu = prop.multiProcessorCount;
  HIP_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));

  Geometry &g = c.g;
  g.nx = L[0]; g.ny = L[1]; g.nz = L[2]; g.nt = L[3];
  g.nxh = g.nx / 2;
  g.vh = (int)(vol / 2);
  g.nxy = g.nxh * g.ny;
  g.nxyz = g.nxy * g.nz;
  if (layout->even_sites_on_node && (long long)layout->even_sites_on_node != vol / 2) {
    fail(__FILE__, __LINE__, "QPHIX_init: even_sites_on_node disagrees with latsize");
    return QPHIX_FAIL;
  }
  // Reduction scratch: enough partials for one block per 64 sites, 4 values
  // each, and for the multi-RHS kernels (8 sources, one 256-lane block per 32
  // sites): vh / 4 values.
  c.max_blocks = (g.vh + 63) / 64 + 1024;
  c.partials_len = std::max((size_t)4 * c.max_blocks, (size_t)g.vh / 4 + 1024);
  c.partials = (double *)dev_alloc(sizeof(double) * c.partials_len);
  if (!c.partials) return QPHIX_MEM_ERROR;
  HIP_CHECK(hipHostMalloc(&c.pinned, PINNED_BYTES, hipHostMallocDefault));
  c.initialised = true;
  return QPHIX_SUCCESS;
}

static QPHIX_status_t finalize_common() {
  Context &c = g_ctx;
  if (!c.initialised) return QPHIX_SUCCESS;
  HIP_CHECK(hipStreamSynchronize(c.stream));
  dev_free(c.partials);
  dev_free(c.ws);
  dev_free(c.ws2);
  if (c.pinned) HIP_CHECK(hipHostFree(c.pinned));
  HIP_CHECK(hipStreamDestroy(c.stream));
  const bool keep_abort = c.abort_on_error;
  const bool keep_sync = c.sync_calls;
  const bool keep_graphs = c.graphs;
  c = Context{};
  c.abort_on_error = keep_abort;
  c.sync_calls = keep_sync;
  c.graphs = keep_graphs;
  return QPHIX_SUCCESS;
}

} // namespace qphix

using namespace qphix;

extern "C" {

// qphix_misc.cpp:427-547 (F and D share the same lattice and device state)
QPHIX_status_t QPHIX_init_F(QPHIX_layout_t *layout) { return init_common(layout); }
QPHIX_status_t QPHIX_init_D(QPHIX_layout_t *layout) { return init_common(layout); }
// qphix_misc.cpp:583-586
QPHIX_status_t QPHIX_init(QPHIX_layout_t *layout) {
  const QPHIX_status_t f = QPHIX_init_F(layout);
  if (f != QPHIX_SUCCESS) return f;  // one report per failed init
  return QPHIX_init_D(layout);
}
QPHIX_status_t QPHIX_finalize_F() { return finalize_common(); }
QPHIX_status_t QPHIX_finalize_D() { return finalize_common(); }
QPHIX_status_t QPHIX_finalize() { return (QPHIX_status_t)(QPHIX_finalize_F() | QPHIX_finalize_D()); }

void QPHIX_MI355X_set_abort_on_error(int on) { g_ctx.abort_on_error = on != 0; }

int QPHIX_MI355X_last_error(char *buf, int len) {
  if (buf && len > 0) {
    std::strncpy(buf, g_last_error, (size_t)len - 1);
    buf[len - 1] = 0;
  }
  return g_error_count;
}

void QPHIX_MI355X_clear_error(void) {
  g_error_count = 0;
  g_last_error[0] = 0;
}

void QPHIX_MI355X_set_sync(int on) { g_ctx.sync_calls = on != 0; }
void QPHIX_MI355X_debug_fail_alloc(int after) { g_fail_alloc_after = after; }
void QPHIX_MI355X_set_fewsums(int mode) { qphix::set_fewsums_mode(mode); }
void QPHIX_MI355X_set_graphs(int on) { g_ctx.graphs = on != 0; }
void QPHIX_MI355X_set_gauge_force_staples(int on) { g_gf_staples = on != 0; }

void QPHIX_MI355X_sync(void) {
  if (g_ctx.initialised) HIP_CHECK(hipStreamSynchronize(g_ctx.stream));
}

void *QPHIX_MI355X_stream(void) { return (void *)g_ctx.stream; }

int QPHIX_MI355X_device_info(int *n_cu, double *hbm_gib) {
  if (!g_ctx.initialised) return 1;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, g_ctx.device) != hipSuccess) return 1;
  if (n_cu) *n_cu = prop.multiProcessorCount;
  if (hbm_gib) *hbm_gib = (double)prop.totalGlobalMem / (1024.0 * 1024.0 * 1024.0);
  return 0;
}

} // extern "C"
