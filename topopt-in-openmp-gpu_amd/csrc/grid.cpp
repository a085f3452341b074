// Library context (error state, stream, present table), grid set-up and the
// device-side grid registry.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

#include "t3d_internal.hpp"

namespace t3d {

// ---------------------------------------------------------------------------
// error state
namespace {
thread_local int g_err = TOP3D_OK;
thread_local char g_msg[512] = "";
hipStream_t g_stream = nullptr;
bool g_stream_set = false;  // the null stream is a valid choice
bool g_own_stream = false;
std::mutex g_mu;
}  // namespace

void set_error(int code, const char *fmt, ...) {
  g_err = code;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
  if (std::getenv("TOP3D_VERBOSE_ERRORS")) std::fprintf(stderr, "top3d: %s\n", g_msg);
}

bool hip_ok(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  set_error(TOP3D_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  return false;
}

// stream of the enclosing StreamScope on this thread, if any
thread_local hipStream_t t_scope_stream = nullptr;
thread_local bool t_scope_on = false;

StreamScope::StreamScope(hipStream_t s) : prev(t_scope_stream), prev_on(t_scope_on) {
  t_scope_stream = s;
  t_scope_on = true;
}
StreamScope::~StreamScope() {
  t_scope_stream = prev;
  t_scope_on = prev_on;
}

hipStream_t stream() {
  if (t_scope_on) return t_scope_stream;
  if (!g_stream_set) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_stream_set) {
      hip_ok(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking),
             "hipStreamCreate");
      g_own_stream = true;
      g_stream_set = true;
    }
  }
  return g_stream;
}

// Device memory held by the library (every allocation goes through
// dalloc/dfree), for top3d_memory_info.
namespace {
std::mutex g_amu;
std::map<uintptr_t, size_t> g_alloc;
uint64_t g_held = 0, g_peak = 0;
}  // namespace

static void track_alloc(void *p, size_t bytes) {
  std::lock_guard<std::mutex> lk(g_amu);
  g_alloc[(uintptr_t)p] = bytes;
  g_held += bytes;
  if (g_held > g_peak) g_peak = g_held;
}

static void untrack_alloc(void *p) {
  std::lock_guard<std::mutex> lk(g_amu);
  auto it = g_alloc.find((uintptr_t)p);
  if (it == g_alloc.end()) return;
  g_held -= it->second;
  g_alloc.erase(it);
}

void *dalloc_uninit(size_t bytes) {
  void *p = nullptr;
  if (bytes == 0) bytes = 8;
  if (!hip_ok(hipMalloc(&p, bytes), "hipMalloc")) return nullptr;
  track_alloc(p, bytes);
  return p;
}

void *dalloc(size_t bytes) {
  void *p = dalloc_uninit(bytes);
  if (!p) return nullptr;
  if (!hip_ok(hipMemsetAsync(p, 0, bytes ? bytes : 8, stream()), "hipMemsetAsync")) return nullptr;
  return p;
}

void dfree(void *p) {
  if (!p) return;
  untrack_alloc(p);
  hip_ok(hipFree(p), "hipFree");
}

// ---------------------------------------------------------------------------
// present table: host address -> device mirror
namespace {
struct Mapping {
  void *dev;
  uint64_t bytes;
};
std::map<uintptr_t, Mapping> g_present;  // keyed by host start address
std::mutex g_pmu;

Mapping *find_mapping(const void *host, uintptr_t *offset) {
  const uintptr_t a = (uintptr_t)host;
  auto it = g_present.upper_bound(a);
  if (it == g_present.begin()) return nullptr;
  --it;
  if (a >= it->first && a < it->first + it->second.bytes) {
    if (offset) *offset = a - it->first;
    return &it->second;
  }
  return nullptr;
}
}  // namespace

bool is_device_ptr(const void *p) {
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice;
}

void *dptr(const void *p, const char *what) {
  if (!p) return nullptr;
  {
    std::lock_guard<std::mutex> lk(g_pmu);
    uintptr_t off = 0;
    Mapping *m = find_mapping(p, &off);
    if (m) return (char *)m->dev + off;
  }
  if (is_device_ptr(p)) return const_cast<void *>(p);
  set_error(TOP3D_ERR_UNMAPPED, "%s: host pointer %p is not mapped (call enter_data/top3d_map)",
            what, p);
  return nullptr;
}

// ---------------------------------------------------------------------------
// padding: y is padded so that nely+1+pad is a multiple of kYAlign; x and z
// carry one halo node on each side plus one, as in gpu_definitions.c:67-92.
namespace {
constexpr int kYAlign = 1;
// supports that setFixedDof_halo (and so setupGC) builds
int g_default_support = TOP3D_SUPPORT_STRIPS;
std::map<const void *, DevGrid *> g_grids;  // keyed by gc->precomputedKE
std::mutex g_gmu;
}  // namespace

void compute_level(const struct gridContext *gc, int l, Level *o) {
  const int n = 1 << l;
  o->l = l;
  o->ncell = n;
  o->nelx = (int)gc->nelx / n;
  o->nely = (int)gc->nely / n;
  o->nelz = (int)gc->nelz / n;
  o->nx = o->nelx + 1;
  o->ny = o->nely + 1;
  o->nz = o->nelz + 1;
  const int pady = (kYAlign - ((o->nely + 1) % kYAlign)) % kYAlign;
  o->wx = o->nelx + 3;
  o->wy = o->nely + pady + 3;
  o->wz = o->nelz + 3;
  o->ndof = 3 * (int64_t)o->wx * o->wy * o->wz;
}

LevelArgs DevGrid::args(int l) const {
  LevelArgs a;
  a.nelx = lv[l].nelx;
  a.nely = lv[l].nely;
  a.nelz = lv[l].nelz;
  a.wy = lv[l].wy;
  a.wz = lv[l].wz;
  a.ncell = lv[l].ncell;
  a.few = ewy;
  a.fez = ewz;
  return a;
}

DevGrid *dev_grid(const struct gridContext *gc) {
  if (!gc) {
    set_error(TOP3D_ERR_ARG, "null gridContext");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_gmu);
  auto it = g_grids.find((const void *)gc->precomputedKE);
  if (it == g_grids.end()) {
    set_error(TOP3D_ERR_ARG, "gridContext was not set up by this library (setupGC)");
    return nullptr;
  }
  return it->second;
}

DevGrid *dev_grid_node(const struct gpuNode *node) {
  if (!node) {
    set_error(TOP3D_ERR_ARG, "null gpuNode");
    return nullptr;
  }
  return dev_grid(node->gc);
}

bool dev_grid_quiet(const struct gridContext *gc) {
  std::lock_guard<std::mutex> lk(g_gmu);
  return gc->precomputedKE && g_grids.count((const void *)gc->precomputedKE);
}

void release_dev_grid(const struct gridContext *gc) {
  DevGrid *g = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_gmu);
    auto it = g_grids.find((const void *)gc->precomputedKE);
    if (it == g_grids.end()) return;
    g = it->second;
    g_grids.erase(it);
  }
  hipStreamSynchronize(stream());
  for (int l = 0; l < kMaxLevels; l++) {
    dfree(g->fixmask[l]);
    dfree(g->fixplane[l]);
    dfree(g->fixed[l]);
  }
  for (int i = 0; i < 2; i++) {
    dfree(g->KE[i]);
    dfree(g->KEf[i]);
  }
  dfree(g->KE2);
  dfree(g->KE3);
  for (int l = 0; l < kMaxLevels; l++) dfree(g->KEdiag[l]);
  for (int i = 0; i < DevGrid::kScratchSlots; i++) dfree(g->scratch[i]);
  delete g;
}

void *DevGrid::work(int slot, size_t bytes) {
  if (scratch[slot] && scratch_bytes[slot] >= bytes) return scratch[slot];
  if (scratch[slot]) {
    hipStreamSynchronize(stream());
    dfree(scratch[slot]);
  }
  scratch[slot] = dalloc(bytes);
  scratch_bytes[slot] = scratch[slot] ? bytes : 0;
  return scratch[slot];
}

// Upload fixed DOFs of level l (list + per-node mask) to the device.
static bool upload_fixed(const struct gridContext *gc, DevGrid *g, int l) {
  dfree(g->fixmask[l]);
  dfree(g->fixplane[l]);
  dfree(g->fixed[l]);
  const Level &L = g->lv[l];
  const int64_t n = (int64_t)gc->fixedDofs[l].n;
  std::vector<uint8_t> mask(L.nodes(), 0);
  std::vector<int> plane(L.wx, 0);
  std::vector<int64_t> idx(n > 0 ? n : 1);
  for (int64_t f = 0; f < n; f++) {
    const int64_t d = (int64_t)gc->fixedDofs[l].idx[f];
    if (d < 0 || d >= L.ndof) {
      set_error(TOP3D_ERR_ARG, "fixed DOF %lld outside level %d", (long long)d, l);
      return false;
    }
    idx[f] = d;
    mask[d / 3] |= (uint8_t)(1u << (d % 3));
    plane[(d / 3) / ((int64_t)L.wy * L.wz)] = 1;
  }
  g->nfixed[l] = n;
  g->fixmask[l] = (uint8_t *)dalloc(mask.size());
  g->fixplane[l] = (int *)dalloc(sizeof(int) * plane.size());
  g->fixed[l] = (int64_t *)dalloc(sizeof(int64_t) * idx.size());
  if (!g->fixmask[l] || !g->fixplane[l] || !g->fixed[l]) return false;
  return T3D_HIP(hipMemcpyAsync(g->fixmask[l], mask.data(), mask.size(),
                                hipMemcpyHostToDevice, stream())) &&
         T3D_HIP(hipMemcpyAsync(g->fixplane[l], plane.data(), sizeof(int) * plane.size(),
                                hipMemcpyHostToDevice, stream())) &&
         T3D_HIP(hipMemcpyAsync(g->fixed[l], idx.data(), sizeof(int64_t) * n,
                                hipMemcpyHostToDevice, stream())) &&
         T3D_HIP(hipStreamSynchronize(stream()));
}

static DevGrid *build_dev_grid(struct gridContext *gc, int nl) {
  DevGrid *g = new DevGrid();
  g->nl = nl;
  g->nelx = (int)gc->nelx;
  g->nely = (int)gc->nely;
  g->nelz = (int)gc->nelz;
  g->E0 = gc->E0;
  g->Emin = gc->Emin;
  for (int l = 0; l < nl; l++) compute_level(gc, l, &g->lv[l]);
  g->ewy = (int)gc->wrapy - 1;
  g->ewz = (int)gc->wrapz - 1;
  g->nelem = (int64_t)(gc->wrapx - 1) * g->ewy * g->ewz;
  for (int i = 0; i < 2 && i < nl; i++) {
    const int n = 1 << i;
    const size_t cnt = 576 * (size_t)(n * n * n);
    g->KE[i] = (double *)dalloc(sizeof(double) * cnt);
    g->KEf[i] = (float *)dalloc(sizeof(float) * cnt);
    std::vector<float> f(cnt);
    for (size_t k = 0; k < cnt; k++) f[k] = (float)gc->precomputedKE[i][k];
    T3D_HIP(hipMemcpyAsync(g->KE[i], gc->precomputedKE[i], sizeof(double) * cnt,
                           hipMemcpyHostToDevice, stream()));
    T3D_HIP(hipMemcpyAsync(g->KEf[i], f.data(), sizeof(float) * cnt,
                           hipMemcpyHostToDevice, stream()));
    T3D_HIP(hipStreamSynchronize(stream()));
  }
  if (nl >= 3) {  // level-2 sub-cell matrices: coarse element matrices as one GEMM
    const size_t cnt = 576 * 64;
    g->KE2 = (double *)dalloc(sizeof(double) * cnt);
    T3D_HIP(hipMemcpyAsync(g->KE2, gc->precomputedKE[2], sizeof(double) * cnt,
                           hipMemcpyHostToDevice, stream()));
    T3D_HIP(hipStreamSynchronize(stream()));
  }
  if (nl >= 4) {  // level-3 sub-cell matrices: level-3 element matrices as one GEMM
    const size_t cnt = 576 * 512;
    g->KE3 = (double *)dalloc(sizeof(double) * cnt);
    T3D_HIP(hipMemcpyAsync(g->KE3, gc->precomputedKE[3], sizeof(double) * cnt,
                           hipMemcpyHostToDevice, stream()));
    T3D_HIP(hipStreamSynchronize(stream()));
  }
  // diagonals of the grid's own sub-cell matrices for the matrix-free
  // inverse diagonals (levels whose 24 x 8^l table stays small)
  for (int l = 0; l < nl && l < 5; l++) {
    const size_t ncell = (size_t)1 << (3 * l);
    std::vector<double> dg(24 * ncell);
    for (size_t c = 0; c < ncell; c++)
      for (int r = 0; r < 24; r++) dg[24 * c + r] = gc->precomputedKE[l][576 * c + 25 * r];
    g->KEdiag[l] = (double *)dalloc(sizeof(double) * dg.size());
    T3D_HIP(hipMemcpyAsync(g->KEdiag[l], dg.data(), sizeof(double) * dg.size(),
                           hipMemcpyHostToDevice, stream()));
    T3D_HIP(hipStreamSynchronize(stream()));
  }
  g->wh0 = hadamard_blocks(gc->precomputedKE[0], g->M0);
  for (int l = 0; l < nl; l++) upload_fixed(gc, g, l);
  std::lock_guard<std::mutex> lk(g_gmu);
  g_grids[(const void *)gc->precomputedKE] = g;
  return g;
}

// Re-upload level-l BCs after a caller changed gc->fixedDofs[l].
void refresh_fixed(struct gridContext *gc, int l) {
  DevGrid *g = dev_grid(gc);
  if (g) upload_fixed(gc, g, l);
}

}  // namespace t3d

using namespace t3d;

// ===========================================================================
// C ABI
extern "C" {

int top3d_last_error(void) { return g_err; }
void top3d_set_default_support(int mode) {
  g_default_support = mode == TOP3D_SUPPORT_FACE ? TOP3D_SUPPORT_FACE : TOP3D_SUPPORT_STRIPS;
}
int top3d_get_default_support(void) { return g_default_support; }
const char *top3d_last_error_string(void) { return g_msg; }
void top3d_clear_error(void) {
  g_err = TOP3D_OK;
  g_msg[0] = 0;
}

void top3d_set_stream(void *s) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_own_stream && g_stream) {
    hipStreamSynchronize(g_stream);
    hipStreamDestroy(g_stream);
  }
  g_stream = (hipStream_t)s;
  g_stream_set = true;
  g_own_stream = false;
}
void *top3d_get_stream(void) { return (void *)stream(); }
int top3d_synchronize(void) {
  return T3D_HIP(hipStreamSynchronize(stream())) ? TOP3D_OK : TOP3D_ERR_HIP;
}

static int map_impl(const void *host, uint64_t bytes, bool copy) {
  if (!host || !bytes) {
    set_error(TOP3D_ERR_ARG, "top3d_map: null/empty range");
    return TOP3D_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_pmu);
  uintptr_t off;
  Mapping *m = find_mapping(host, &off);
  if (m) {  // already present: OpenMP semantics keep the existing mirror
    if (off + bytes > m->bytes) {
      set_error(TOP3D_ERR_ARG, "top3d_map: range extends an existing mapping");
      return TOP3D_ERR_ARG;
    }
    return TOP3D_OK;
  }
  void *d = dalloc_uninit(bytes);
  if (!d) return TOP3D_ERR_HIP;
  if (copy) {
    if (!T3D_HIP(hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, stream())) ||
        !T3D_HIP(hipStreamSynchronize(stream()))) {
      dfree(d);
      return TOP3D_ERR_HIP;
    }
  } else {
    T3D_HIP(hipMemsetAsync(d, 0, bytes, stream()));
  }
  g_present[(uintptr_t)host] = Mapping{d, bytes};
  return TOP3D_OK;
}

int top3d_memory_info(struct Top3dMemoryInfo *out, int reset_peak) {
  if (!out) {
    set_error(TOP3D_ERR_ARG, "top3d_memory_info: null output");
    return TOP3D_ERR_ARG;
  }
  {
    std::lock_guard<std::mutex> lk(g_amu);
    out->library_bytes = g_held;
    out->library_peak_bytes = g_peak;
    if (reset_peak) g_peak = g_held;
  }
  size_t fr = 0, tot = 0;
  if (!T3D_HIP(hipMemGetInfo(&fr, &tot))) return TOP3D_ERR_HIP;
  out->device_used_bytes = (uint64_t)(tot - fr);
  out->device_total_bytes = (uint64_t)tot;
  return TOP3D_OK;
}

int top3d_map(const void *host, uint64_t bytes) { return map_impl(host, bytes, true); }
int top3d_map_alloc(const void *host, uint64_t bytes) { return map_impl(host, bytes, false); }

int top3d_unmap(const void *host) {
  std::lock_guard<std::mutex> lk(g_pmu);
  auto it = g_present.find((uintptr_t)host);
  if (it == g_present.end()) {
    set_error(TOP3D_ERR_UNMAPPED, "top3d_unmap: %p not mapped", host);
    return TOP3D_ERR_UNMAPPED;
  }
  hipStreamSynchronize(stream());
  dfree(it->second.dev);
  g_present.erase(it);
  return TOP3D_OK;
}

int top3d_update_to(const void *host, uint64_t bytes) {
  uintptr_t off;
  void *d;
  {
    std::lock_guard<std::mutex> lk(g_pmu);
    Mapping *m = find_mapping(host, &off);
    if (!m || off + bytes > m->bytes) {
      set_error(TOP3D_ERR_UNMAPPED, "top3d_update_to: %p not mapped", host);
      return TOP3D_ERR_UNMAPPED;
    }
    d = (char *)m->dev + off;
  }
  return (T3D_HIP(hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, stream())) &&
          T3D_HIP(hipStreamSynchronize(stream())))
             ? TOP3D_OK
             : TOP3D_ERR_HIP;
}

int top3d_update_from(void *host, uint64_t bytes) {
  uintptr_t off;
  void *d;
  {
    std::lock_guard<std::mutex> lk(g_pmu);
    Mapping *m = find_mapping(host, &off);
    if (!m || off + bytes > m->bytes) {
      set_error(TOP3D_ERR_UNMAPPED, "top3d_update_from: %p not mapped", host);
      return TOP3D_ERR_UNMAPPED;
    }
    d = (char *)m->dev + off;
  }
  return (T3D_HIP(hipMemcpyAsync(host, d, bytes, hipMemcpyDeviceToHost, stream())) &&
          T3D_HIP(hipStreamSynchronize(stream())))
             ? TOP3D_OK
             : TOP3D_ERR_HIP;
}

void *top3d_device_ptr(const void *p) { return dptr(p, "top3d_device_ptr"); }

// ---- padding / set-up ------------------------------------------------------

void computePadding(const struct gridContext *gc, const int l, int *ncell,
                    int *wrapx, int *wrapy, int *wrapz, uint_fast32_t *ndof) {
  Level L;
  compute_level(gc, l, &L);
  *ncell = L.ncell;
  *wrapx = L.wx;
  *wrapy = L.wy;
  *wrapz = L.wz;
  *ndof = (uint_fast32_t)L.ndof;
}

// ---- the reference's padded layout ------------------------------------------
// The reference pads y so that nely_l + 1 + pad is a multiple of
// STENCIL_SIZE_Y = 128 (gpu_definitions.c:67-92); x and z carry no padding
// (STENCIL_SIZE_X = STENCIL_SIZE_Z = 1, definitions.h:20-21).  This library
// leaves y unpadded.  Both keep node (i, j, k) at i*wy*wz + k*wy + j with the
// same halo, so a copy moves each (i, k) row of y-length min(wy) as a block.
namespace {
struct Strides {
  int64_t wx, wy_ref, wy_lib, wz;
};
Strides layout_strides(const struct gridContext *gc, int l) {
  Level L;
  compute_level(gc, l, &L);
  const int ny = L.nely;
  const int pad = (STENCIL_SIZE_Y - ((ny + 1) % STENCIL_SIZE_Y)) % STENCIL_SIZE_Y;
  return Strides{L.wx, (int64_t)ny + pad + 3, L.wy, L.wz};
}
// rows of `per` values each: (rows_x * rows_z) rows, `len` values copied,
// the destination's tail of each row zeroed
void copy_rows(int64_t nrows, int64_t src_row, int64_t dst_row, int64_t per, const double *src,
               double *dst) {
  const int64_t len = std::min(src_row, dst_row) * per;
  for (int64_t r = 0; r < nrows; r++) {
    std::memcpy(dst + r * dst_row * per, src + r * src_row * per, sizeof(double) * len);
    std::fill(dst + r * dst_row * per + len, dst + (r + 1) * dst_row * per, 0.0);
  }
}
bool layout_args_ok(const struct gridContext *gc, int l, const void *a, const void *b,
                    const char *what) {
  if (!gc || !a || !b || l < 0 || l >= kMaxLevels || (gc->nely >> l) < 1) {
    set_error(TOP3D_ERR_ARG, "%s: invalid arguments", what);
    return false;
  }
  return true;
}
}  // namespace

void top3d_nodes_from_reference_layout(const struct gridContext *gc, int l, const double *ref,
                                       double *lib) {
  if (!layout_args_ok(gc, l, ref, lib, "top3d_nodes_from_reference_layout")) return;
  const Strides s = layout_strides(gc, l);
  copy_rows(s.wx * s.wz, s.wy_ref, s.wy_lib, 3, ref, lib);
}

void top3d_nodes_to_reference_layout(const struct gridContext *gc, int l, const double *lib,
                                     double *ref) {
  if (!layout_args_ok(gc, l, lib, ref, "top3d_nodes_to_reference_layout")) return;
  const Strides s = layout_strides(gc, l);
  copy_rows(s.wx * s.wz, s.wy_lib, s.wy_ref, 3, lib, ref);
}

void top3d_elements_from_reference_layout(const struct gridContext *gc, const double *ref,
                                          double *lib) {
  if (!layout_args_ok(gc, 0, ref, lib, "top3d_elements_from_reference_layout")) return;
  const Strides s = layout_strides(gc, 0);
  copy_rows((s.wx - 1) * (s.wz - 1), s.wy_ref - 1, s.wy_lib - 1, 1, ref, lib);
}

void top3d_elements_to_reference_layout(const struct gridContext *gc, const double *lib,
                                        double *ref) {
  if (!layout_args_ok(gc, 0, lib, ref, "top3d_elements_to_reference_layout")) return;
  const Strides s = layout_strides(gc, 0);
  copy_rows((s.wx - 1) * (s.wz - 1), s.wy_lib - 1, s.wy_ref - 1, 1, lib, ref);
}

void top3d_reference_padding(const struct gridContext *gc, int l, int *wrapx, int *wrapy,
                             int *wrapz, uint64_t *ndof) {
  if (!gc || l < 0 || l >= kMaxLevels) {
    set_error(TOP3D_ERR_ARG, "top3d_reference_padding: invalid arguments");
    return;
  }
  const Strides s = layout_strides(gc, l);
  if (wrapx) *wrapx = (int)s.wx;
  if (wrapy) *wrapy = (int)s.wy_ref;
  if (wrapz) *wrapz = (int)s.wz;
  if (ndof) *ndof = 3 * (uint64_t)s.wx * s.wy_ref * s.wz;
}

// stencil_grid_utility.c:7-65.  TOP3D_SUPPORT_STRIPS (the default) clamps
// the two strips of the x=0 face that the reference builds (:34-57);
// TOP3D_SUPPORT_FACE clamps every node of that face, the variant the
// reference keeps in comments (:19-32).
void setFixedDof_halo(struct gridContext *gc, const int l) {
  Level L;
  compute_level(gc, l, &L);
  if (g_default_support == TOP3D_SUPPORT_FACE) {
    gc->fixedDofs[l].n = 3 * (uint64_t)L.ny * L.nz;
    gc->fixedDofs[l].idx =
        (uint_fast32_t *)std::malloc(sizeof(uint_fast32_t) * gc->fixedDofs[l].n);
    int off = 0;
    for (uint64_t k = 1; k < (uint64_t)L.nz + 1; k++)
      for (uint64_t j = 1; j < (uint64_t)L.ny + 1; j++)
        for (int c = 0; c < 3; c++)
          gc->fixedDofs[l].idx[off++] = 3 * ((uint64_t)L.wy * L.wz + L.wy * k + j) + c;
    if (dev_grid_quiet(gc)) refresh_fixed(gc, l);
    return;
  }
  const int nodelimit = L.nely / 4 + 1;
  gc->fixedDofs[l].n = 3 * L.nz * 2 * nodelimit;
  gc->fixedDofs[l].idx =
      (uint_fast32_t *)std::malloc(sizeof(uint_fast32_t) * gc->fixedDofs[l].n);
  int off = 0;
  const uint64_t i = 1;
  for (uint64_t k = 1; k < (uint64_t)L.nz + 1; k++) {
    for (uint64_t j = 1; j < (uint64_t)nodelimit + 1; j++)
      for (int c = 0; c < 3; c++)
        gc->fixedDofs[l].idx[off++] = 3 * (i * L.wy * L.wz + L.wy * k + j) + c;
    for (uint64_t j = (uint64_t)(L.ny + 1 - nodelimit); j < (uint64_t)L.ny + 1; j++)
      for (int c = 0; c < 3; c++)
        gc->fixedDofs[l].idx[off++] = 3 * (i * L.wy * L.wz + L.wy * k + j) + c;
  }
  if (dev_grid_quiet(gc)) refresh_fixed(gc, l);
}

// stencil_grid_utility.c:67-108
void setupGC(struct gridContext *gc, const int nl, const int nelx,
             const int nely, const int nelz) {
  if (nl < 3 || nl > kMaxLevels) {
    set_error(TOP3D_ERR_ARG, "setupGC: nl=%d outside [3,%d]", nl, kMaxLevels);
    std::memset(gc, 0, sizeof(*gc));
    return;
  }
  const int f = 1 << (nl - 1);
  if (nelx <= 0 || nely <= 0 || nelz <= 0 || nelx % f || nely % f || nelz % f) {
    set_error(TOP3D_ERR_ARG, "setupGC: grid %dx%dx%d not divisible by 2^(nl-1)",
              nelx, nely, nelz);
    std::memset(gc, 0, sizeof(*gc));
    return;
  }
  gc->E0 = 1;
  gc->Emin = 1e-6;
  gc->nu = 0.3;
  gc->penal = 3;  // the exponent 3 is built into the kernels
  gc->elementSizeX = gc->elementSizeY = gc->elementSizeZ = 0.5;
  gc->nelx = nelx;
  gc->nely = nely;
  gc->nelz = nelz;
  Level L0;
  compute_level(gc, 0, &L0);
  gc->wrapx = L0.wx;
  gc->wrapy = L0.wy;
  gc->wrapz = L0.wz;
  gc->precomputedKE = (MTYPE **)std::malloc(sizeof(MTYPE *) * nl);
  gc->fixedDofs = (struct FixedDofs *)std::calloc(nl, sizeof(struct FixedDofs));
  for (int l = 0; l < nl; l++) {
    const size_t n = 1u << l;
    gc->precomputedKE[l] = (MTYPE *)std::malloc(sizeof(MTYPE) * 576 * n * n * n);
    element_matrices(gc->nu, l, gc->precomputedKE[l]);
    setFixedDof_halo(gc, l);
  }
  build_dev_grid(gc, nl);
}

void freeGC(struct gridContext *gc, const int nl) {
  if (!gc->precomputedKE) return;
  release_dev_grid(gc);
  for (int l = 0; l < nl; l++) {
    std::free(gc->precomputedKE[l]);
    std::free(gc->fixedDofs[l].idx);
  }
  std::free(gc->precomputedKE);
  std::free(gc->fixedDofs);
  gc->precomputedKE = nullptr;
  gc->fixedDofs = nullptr;
}

static void alloc_field(const struct gridContext *gc, int l, size_t esz, void **v) {
  Level L;
  compute_level(gc, l, &L);
  *v = std::calloc((size_t)L.ndof, esz);
}

void allocateZeroPaddedStateField(const struct gridContext gc, const int l, CTYPE **v) {
  alloc_field(&gc, l, sizeof(CTYPE), (void **)v);
}
void allocateZeroPaddedStateField_MTYPE(const struct gridContext gc, const int l,
                                        MTYPE **v) {
  alloc_field(&gc, l, sizeof(MTYPE), (void **)v);
}
void allocateZeroPaddedStateField_STYPE(const struct gridContext gc, const int l,
                                        STYPE **v) {
  alloc_field(&gc, l, sizeof(STYPE), (void **)v);
}

// gpu_definitions.c:3-53; one GPU: num_targets must be 1.
void gridContextToGPUGrid(struct gridContext *gc, struct gpuGrid *gpu_gc,
                          const int nl, const int verbose) {
  if (gpu_gc->num_targets != 1) {
    set_error(TOP3D_ERR_ARG,
              "gridContextToGPUGrid: %d targets requested; this build runs on one GPU",
              gpu_gc->num_targets);
    gpu_gc->num_targets = 1;
  }
  gpu_gc->target = (struct gpuNode *)std::calloc(1, sizeof(struct gpuNode));
  struct gpuNode *node = &gpu_gc->target[0];
  node->id = 0;
  node->prev = node->next = nullptr;
  node->offset_x = 0;
  node->x_global = 0;
  node->gc = (struct gridContext *)std::malloc(sizeof(struct gridContext));
  setupGC(node->gc, nl, (int)gc->nelx, (int)gc->nely, (int)gc->nelz);
  if (verbose == 1)
    std::printf("GPU %d:\n\tGlobal index: %d\n\tOffset: %d\n", node->id,
                node->x_global, node->offset_x);
}

void freeGPUGrid(struct gpuGrid *gpu_gc, const int nl) {
  if (!gpu_gc->target) return;
  struct gpuNode *node = &gpu_gc->target[0];
  freeGC(node->gc, nl);
  std::free(node->gc);
  std::free(gpu_gc->target);
  gpu_gc->target = nullptr;
}

}  // extern "C"
