// C entry points of the reference's operator API (stencil_methods.h,
// stencil_solvers.h, stencil_assembly.h, gpu_definitions.h,
// stencil_optimization.h) on the gfx950 kernels, plus the autograd hooks.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "solver_state.hpp"

namespace t3d {

enum { kApplyMode = 0, kJacobiMode = 1, kResidualMode = 2 };
enum { kSlotE64 = 0, kSlotE32 = 1, kSlotTmp = 2, kSlotRed = 3, kSlotKe = 4, kSlotS = 5,
       kSlotVec = 6, kSlotAux = 7 };

static double *young64(DevGrid &g, const double *dx) {
  double *E = (double *)g.work(kSlotE64, sizeof(double) * g.nelem);
  if (E) launch_young<double>(g, dx, E);
  return E;
}

// element matrices of levels 2..l into one scratch block; returns level l's
static double *coarse_elements(DevGrid &g, const double *E, int l) {
  size_t off[kMaxLevels] = {}, total = 0;
  for (int q = 2; q <= l; q++) {
    off[q] = total;
    total += 576 * (size_t)g.lv[q].elements();
  }
  double *K = (double *)g.work(kSlotKe, sizeof(double) * total);
  if (!K) return nullptr;
  for (int q = 2; q <= l; q++) {
    if (q == 2) launch_elem_level2(g, E, K + off[2]);
    else launch_elem_galerkin(g, q, K + off[q - 1], K + off[q]);
  }
  return K + off[l];
}

// identity-row (matrix-free convention) stencil of level l >= 2
static double *mf_stencil(DevGrid &g, const double *E, int l) {
  double *Ke = coarse_elements(g, E, l);
  double *S = (double *)g.work(kSlotS, sizeof(double) * 243 * g.lv[l].interior_nodes());
  if (!Ke || !S) return nullptr;
  launch_node_stencil(g, l, Ke, S, 0);
  return S;
}

static bool sync_ok() { return T3D_HIP(hipStreamSynchronize(stream())); }

// Jacobi sweeps with the fused operator, ping-pong u <-> tmp, result in u.
// The sweep kernels write the real nodes only; tmp starts as a copy of u so
// that both buffers carry u's halo values, which the reference's sweeps leave
// unchanged.
template <typename F>
static void sweeps(int64_t ndof, int nswp, double *u, double *tmp, F apply) {
  if (nswp > 0)
    hipMemcpyAsync(tmp, u, sizeof(double) * ndof, hipMemcpyDeviceToDevice, stream());
  double *cur = u, *oth = tmp;
  for (int s = 0; s < nswp; s++) {
    apply(cur, oth);
    std::swap(cur, oth);
  }
  if (cur != u) hipMemcpyAsync(u, cur, sizeof(double) * ndof, hipMemcpyDeviceToDevice, stream());
}

// ---------------------------------------------------------------------------
// CSR kernels for the host-CSR compatibility API (stencil_assembly.c:262-364)
__global__ void k_csr_apply(int nx, int ny, int nz, int wy, int wz,
                            const int *__restrict__ ro, const int *__restrict__ ci,
                            const double *__restrict__ va, const double *__restrict__ in,
                            double *__restrict__ out, const double *__restrict__ b,
                            const double *__restrict__ invD, double omega, int mode) {
  const int64_t nrows = 3 * (int64_t)nx * ny * nz;
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= nrows) return;
  auto halo = [&](int64_t d) {
    const int64_t p = d / 3;
    const int j = (int)(p % ny), k = (int)((p / ny) % nz), i = (int)(p / ((int64_t)ny * nz));
    return 3 * (((int64_t)(i + 1) * wz + (k + 1)) * wy + (j + 1)) + d % 3;
  };
  double acc = 0.0;
  for (int e = ro[row]; e < ro[row + 1]; e++) acc += in[halo(ci[e])] * va[e];
  const int64_t h = halo(row);
  out[h] = mode == kApplyMode ? acc : in[h] + omega * invD[h] * (b[h] - acc);
}

__global__ void k_copy_interior_elems(int64_t nelem, int nelx, int nely, int nelz, int few,
                                      int fez, const double *__restrict__ src,
                                      double *__restrict__ dst) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  const int j = (int)(e % few), k = (int)((e / few) % fez), i = (int)(e / ((int64_t)few * fez));
  if (i >= 1 && i <= nelx && j >= 1 && j <= nely && k >= 1 && k <= nelz) dst[e] = src[e];
}

__global__ void k_scale(int64_t n, double a, double *__restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] *= a;
}

template <typename T>
__global__ void k_zero_noninterior(int64_t nodes, int nx, int ny, int nz, int wy, int wz,
                                   T *__restrict__ v) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nodes) return;
  const int j = (int)(n % wy), k = (int)((n / wy) % wz), i = (int)(n / ((int64_t)wy * wz));
  if (i >= 1 && i <= nx && j >= 1 && j <= ny && k >= 1 && k <= nz) return;
  v[3 * n] = v[3 * n + 1] = v[3 * n + 2] = T(0);
}

// zero every non-interior node of a level-l vector (the operators write the
// interior nodes only)
template <typename T>
static void zero_noninterior(const DevGrid &g, int l, T *v) {
  const Level &L = g.lv[l];
  const int64_t nodes = L.nodes();
  hipLaunchKernelGGL((k_zero_noninterior<T>), dim3((nodes + 255) / 256), dim3(256), 0, stream(),
                     nodes, L.nx, L.ny, L.nz, L.wy, L.wz, v);
}

// Coarse direct solver object behind initializeCholmod/factorize/solve.
struct CoarseObj {
  int l = 0;
  CoarseFactor f;
};

static void upload_csr(const struct CSRMatrix &M, int **ro, int **ci, double **va) {
  *ro = (int *)dalloc(sizeof(int) * (M.nrows + 1));
  *ci = (int *)dalloc(sizeof(int) * M.nnz);
  *va = (double *)dalloc(sizeof(double) * M.nnz);
  hipMemcpyAsync(*ro, M.rowOffsets, sizeof(int) * (M.nrows + 1), hipMemcpyHostToDevice, stream());
  hipMemcpyAsync(*ci, M.colIndex, sizeof(int) * M.nnz, hipMemcpyHostToDevice, stream());
  hipMemcpyAsync(*va, M.vals, sizeof(double) * M.nnz, hipMemcpyHostToDevice, stream());
}

// Temporary device view of a pointer: mapped/device pointers are used as is;
// unmapped host arrays are copied in (and back if `out`).
struct DevView {
  void *host = nullptr, *dev = nullptr;
  size_t bytes = 0;
  bool owned = false, out = false;
  bool open(const void *p, size_t nbytes, bool copy_in, bool copy_out) {
    host = const_cast<void *>(p);
    bytes = nbytes;
    out = copy_out;
    if (!p) return false;
    void *d = top3d_device_ptr_quiet(p);
    if (d) {
      dev = d;
      return true;
    }
    dev = dalloc(nbytes);
    owned = true;
    if (!dev) return false;
    if (copy_in) hipMemcpyAsync(dev, p, nbytes, hipMemcpyHostToDevice, stream());
    return true;
  }
  static void *top3d_device_ptr_quiet(const void *p) {
    const int e = top3d_last_error();
    char msg[512];
    std::strncpy(msg, top3d_last_error_string(), sizeof(msg) - 1);
    void *d = dptr(p, "view");
    if (!d) set_error(e, "%s", msg);  // restore the previous status
    return d;
  }
  void close() {
    if (owned && dev) {
      if (out) hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, stream());
      hipStreamSynchronize(stream());
      dfree(dev);
    }
    dev = nullptr;
  }
};

}  // namespace t3d

using namespace t3d;

extern "C" {

// ---- matrix-free operators -------------------------------------------------

// stencil_methods.c:7-428
void applyStateOperator_stencil(const struct gpuNode *node, const DTYPE *x, const CTYPE *in,
                                CTYPE *out) {
  DevGrid *g = dev_grid_node(node);
  if (!g) return;
  const double *dx = (const double *)dptr(x, "applyStateOperator_stencil x");
  const double *din = (const double *)dptr(in, "applyStateOperator_stencil in");
  double *dout = (double *)dptr(out, "applyStateOperator_stencil out");
  if (!dx || !din || !dout) return;
  double *E = young64(*g, dx);
  launch_fine<double>(*g, kApplyMode, E, din, dout, nullptr, nullptr, 0.0, 0);
  launch_zero_pad_rows(*g, dout);
  sync_ok();
}

void applyStateOperator_stencil_f32(const struct gpuNode *node, const DTYPE *x, const float *in,
                                    float *out) {
  DevGrid *g = dev_grid_node(node);
  if (!g) return;
  const double *dx = (const double *)dptr(x, "x");
  const float *din = (const float *)dptr(in, "in");
  float *dout = (float *)dptr(out, "out");
  if (!dx || !din || !dout) return;
  float *E = (float *)g->work(kSlotE32, sizeof(float) * g->nelem);
  launch_young<float>(*g, dx, E);
  launch_fine<float>(*g, kApplyMode, E, din, dout, nullptr, nullptr, 0.f, 0);
  sync_ok();
}

// stencil_methods.c:436-519 (any level >= 1; identity BC rows)
void applyStateOperatorSubspace_halo(const struct gpuNode *node, const int l, const DTYPE *x,
                                     CTYPE *in, CTYPE *out) {
  DevGrid *g = dev_grid_node(node);
  if (!g) return;
  if (l < 0 || l >= g->nl) {
    set_error(TOP3D_ERR_ARG, "applyStateOperatorSubspace_halo: level %d", l);
    return;
  }
  const double *dx = (const double *)dptr(x, "x");
  const double *din = (const double *)dptr(in, "in");
  double *dout = (double *)dptr(out, "out");
  if (!dx || !din || !dout) return;
  double *E = young64(*g, dx);
  if (l == 0) {
    launch_fine<double>(*g, kApplyMode, E, din, dout, nullptr, nullptr, 0.0, 0);
    launch_zero_pad_rows(*g, dout);
  } else if (l == 1) {
    launch_level1<double>(*g, kApplyMode, E, din, dout, nullptr, nullptr, 0.0, 0);
    zero_noninterior<double>(*g, 1, dout);
  } else {
    double *S = mf_stencil(*g, E, l);
    if (!S) return;
    launch_stencil<double>(*g, l, kApplyMode, S, din, dout, nullptr, nullptr, 0.0);
  }
  sync_ok();
}

void applyStateOperatorSubspace_halo_f32(const struct gpuNode *node, const int l,
                                         const DTYPE *x, const float *in, float *out) {
  DevGrid *g = dev_grid_node(node);
  if (!g) return;
  if (l != 1) {
    set_error(TOP3D_ERR_ARG, "applyStateOperatorSubspace_halo_f32: level 1 only");
    return;
  }
  const double *dx = (const double *)dptr(x, "x");
  const float *din = (const float *)dptr(in, "in");
  float *dout = (float *)dptr(out, "out");
  if (!dx || !din || !dout) return;
  float *E = (float *)g->work(kSlotE32, sizeof(float) * g->nelem);
  launch_young<float>(*g, dx, E);
  launch_level1<float>(*g, kApplyMode, E, din, dout, nullptr, nullptr, 0.f, 0);
  zero_noninterior<float>(*g, 1, dout);
  sync_ok();
}

// stencil_methods.c:521-590
void projectToFinerGrid_halo(const struct gpuNode *node, const int l, const CTYPE *ucoarse,
                             CTYPE *ufine) {
  DevGrid *g = dev_grid_node(node);
  if (!g) return;
  if (l < 0 || l + 1 >= g->nl) {
    set_error(TOP3D_ERR_ARG, "projectToFinerGrid_halo: level %d", l);
    return;
  }
  const double *dc = (const double *)dptr(ucoarse, "ucoarse");
  double *df = (double *)dptr(ufine, "ufine");
  if (!dc || !df) return;
  launch_prolong<double>(*g, l, dc, df);
  sync_ok();
}

// stencil_methods.c:595-688
void projectToCoarserGrid_halo(const struct gpuNode *node, const int l, const CTYPE *ufine,
                               CTYPE *ucoarse) {
  DevGrid *g = dev_grid_node(node);
  if (!g) return;
  if (l < 0 || l + 1 >= g->nl) {
    set_error(TOP3D_ERR_ARG, "projectToCoarserGrid_halo: level %d", l);
    return;
  }
  const double *df = (const double *)dptr(ufine, "ufine");
  double *dc = (double *)dptr(ucoarse, "ucoarse");
  if (!dc || !df) return;
  launch_restrict<double>(*g, l, df, dc);
  sync_ok();
}

// stencil_methods.c:693-773
void assembleInvertedMatrixDiagonalSubspace_halo(const struct gpuNode *node, const DTYPE *x,
                                                 const int l, MTYPE *diag) {
  DevGrid *g = dev_grid_node(node);
  if (!g) return;
  if (l < 0 || l >= g->nl) {
    set_error(TOP3D_ERR_ARG, "assembleInvertedMatrixDiagonalSubspace_halo: level %d", l);
    return;
  }
  const double *dx = (const double *)dptr(x, "x");
  double *dd = (double *)dptr(diag, "diag");
  if (!dx || !dd) return;
  double *E = young64(*g, dx);
  launch_invdiag_mf<double>(*g, l, E, dd);
  sync_ok();
}

// stencil_methods.c:808-862; c is a host or device scalar
void getComplianceAndSensetivity_halo(const struct gpuNode *node, const DTYPE *x, STYPE *u,
                                      DTYPE *c, DTYPE *dcdx) {
  DevGrid *g = dev_grid_node(node);
  if (!g) return;
  const double *dx = (const double *)dptr(x, "x");
  const double *du = (const double *)dptr(u, "u");
  double *ddc = (double *)dptr(dcdx, "dcdx");
  if (!dx || !du || !ddc) return;
  double *red = (double *)g->work(kSlotRed, sizeof(double) * (1024 + 16));
  launch_sensitivity(*g, dx, du, ddc, red, red + 1024);
  if (is_device_ptr(c))
    hipMemcpyAsync(c, red + 1024, sizeof(double), hipMemcpyDeviceToDevice, stream());
  else
    hipMemcpyAsync(c, red + 1024, sizeof(double), hipMemcpyDeviceToHost, stream());
  sync_ok();
}

// ---- smoothers ---------------------------------------------------------------

// stencil_solvers.c:97-122
void smoothDampedJacobi_halo(const struct gpuNode *node, const DTYPE *x,
                             const uint_fast32_t nswp, const CTYPE omega, const MTYPE *invD,
                             CTYPE *u, const CTYPE *b, CTYPE *tmp) {
  DevGrid *g = dev_grid_node(node);
  if (!g) return;
  const double *dx = (const double *)dptr(x, "x");
  const double *di = (const double *)dptr(invD, "invD");
  double *du = (double *)dptr(u, "u");
  const double *db = (const double *)dptr(b, "b");
  double *dt = (double *)dptr(tmp, "tmp");
  if (!dx || !di || !du || !db || !dt) return;
  double *E = young64(*g, dx);
  sweeps(g->lv[0].ndof, (int)nswp, du, dt, [&](const double *cur, double *nxt) {
    launch_fine<double>(*g, kJacobiMode, E, cur, nxt, db, di, omega, 0);
  });
  sync_ok();
}

// stencil_solvers.c:11-49
void smoothDampedJacobiSubspace_halo(const struct gpuNode *node, const DTYPE *x, const int l,
                                     const uint_fast32_t nswp, const CTYPE omega,
                                     const MTYPE *invD, CTYPE *u, const CTYPE *b, CTYPE *tmp) {
  DevGrid *g = dev_grid_node(node);
  if (!g) return;
  if (l < 1 || l >= g->nl) {
    set_error(TOP3D_ERR_ARG, "smoothDampedJacobiSubspace_halo: level %d", l);
    return;
  }
  const double *dx = (const double *)dptr(x, "x");
  const double *di = (const double *)dptr(invD, "invD");
  double *du = (double *)dptr(u, "u");
  const double *db = (const double *)dptr(b, "b");
  double *dt = (double *)dptr(tmp, "tmp");
  if (!dx || !di || !du || !db || !dt) return;
  double *E = young64(*g, dx);
  if (l == 1) {
    sweeps(g->lv[1].ndof, (int)nswp, du, dt, [&](const double *cur, double *nxt) {
      launch_level1<double>(*g, kJacobiMode, E, cur, nxt, db, di, omega, 0);
    });
  } else {
    double *S = mf_stencil(*g, E, l);
    if (!S) return;
    sweeps(g->lv[l].ndof, (int)nswp, du, dt, [&](const double *cur, double *nxt) {
      launch_stencil<double>(*g, l, kJacobiMode, S, cur, nxt, db, di, omega);
    });
  }
  sync_ok();
}

// ---- assembled levels: host CSR in the reference layout ----------------------

// stencil_assembly.c:8-117 (host pattern)
void allocateSubspaceMatrix(const struct gridContext gc, const int l, struct CSRMatrix *M) {
  Level L;
  compute_level(&gc, l, &L);
  const int nxc = L.nx, nyc = L.ny, nzc = L.nz;
  M->nrows = 3 * nxc * nyc * nzc;
  M->rowOffsets = (int *)std::malloc(sizeof(int) * (M->nrows + 1));
  M->rowOffsets[0] = 0;
  for (int i = 0; i < nxc; i++)
    for (int k = 0; k < nzc; k++)
      for (int j = 0; j < nyc; j++) {
        const int nid = i * nyc * nzc + nyc * k + j;
        const int sz = 3 * (std::min(i + 1, nxc - 1) - std::max(i - 1, 0) + 1) *
                       (std::min(j + 1, nyc - 1) - std::max(j - 1, 0) + 1) *
                       (std::min(k + 1, nzc - 1) - std::max(k - 1, 0) + 1);
        for (int c = 0; c < 3; c++) M->rowOffsets[3 * nid + c + 1] = sz;
      }
  for (int i = 1; i < M->nrows + 1; i++) M->rowOffsets[i] += M->rowOffsets[i - 1];
  M->nnz = (uint64_t)M->rowOffsets[M->nrows];
  M->colIndex = (int *)std::malloc(sizeof(int) * M->nnz);
  M->vals = (MTYPE *)std::calloc(M->nnz, sizeof(MTYPE));
  for (int i = 0; i < nxc; i++)
    for (int k = 0; k < nzc; k++)
      for (int j = 0; j < nyc; j++) {
        const int rn = i * nyc * nzc + nyc * k + j;
        int off[3] = {M->rowOffsets[3 * rn], M->rowOffsets[3 * rn + 1], M->rowOffsets[3 * rn + 2]};
        for (int ii = std::max(i - 1, 0); ii < std::min(i + 2, nxc); ii++)
          for (int kk = std::max(k - 1, 0); kk < std::min(k + 2, nzc); kk++)
            for (int jj = std::max(j - 1, 0); jj < std::min(j + 2, nyc); jj++) {
              const int cn = ii * nyc * nzc + nyc * kk + jj;
              for (int r = 0; r < 3; r++) {
                for (int c = 0; c < 3; c++) M->colIndex[off[r] + c] = 3 * cn + c;
                off[r] += 3;
              }
            }
      }
}

void freeSubspaceMatrix(struct CSRMatrix *M) {
  std::free(M->rowOffsets);
  std::free(M->colIndex);
  std::free(M->vals);
  M->rowOffsets = nullptr;
  M->colIndex = nullptr;
  M->vals = nullptr;
}

// stencil_assembly.c:126-260: assembled on the GPU (Galerkin element
// matrices, symmetric elimination), copied into the caller's host CSR;
// tmp receives the free-DOF mask per CSR row as in the reference.
void assembleSubspaceMatrix(const struct gridContext gc, const int l, const DTYPE *x,
                            struct CSRMatrix M, MTYPE *tmp) {
  DevGrid *g = dev_grid(&gc);
  if (!g) return;
  if (l < 2 || l >= g->nl) {
    set_error(TOP3D_ERR_ARG, "assembleSubspaceMatrix: level %d", l);
    return;
  }
  const double *dx = (const double *)dptr(x, "x");
  if (!dx) return;
  double *E = young64(*g, dx);
  double *Ke = coarse_elements(*g, E, l);
  const Level &L = g->lv[l];
  const int64_t nint = L.interior_nodes();
  double *S = (double *)g->work(kSlotS, sizeof(double) * 243 * nint);
  if (!Ke || !S) return;
  launch_node_stencil(*g, l, Ke, S, 1);
  std::vector<double> hS(243 * (size_t)nint);
  hipMemcpyAsync(hS.data(), S, sizeof(double) * hS.size(), hipMemcpyDeviceToHost, stream());
  std::vector<uint8_t> mask(L.nodes());
  hipMemcpyAsync(mask.data(), g->fixmask[l], mask.size(), hipMemcpyDeviceToHost, stream());
  if (!sync_ok()) return;
  for (int i = 0; i < L.nx; i++)
    for (int k = 0; k < L.nz; k++)
      for (int j = 0; j < L.ny; j++) {
        const int64_t n = ((int64_t)i * L.nz + k) * L.ny + j;
        for (int r = 0; r < 3; r++) {
          int e = M.rowOffsets[3 * n + r];
          for (int dx2 = -1; dx2 <= 1; dx2++)
            for (int dz = -1; dz <= 1; dz++)
              for (int dy = -1; dy <= 1; dy++) {
                const int i2 = i + dx2, k2 = k + dz, j2 = j + dy;
                if (i2 < 0 || i2 >= L.nx || j2 < 0 || j2 >= L.ny || k2 < 0 || k2 >= L.nz) continue;
                const int nb = (dx2 + 1) * 9 + (dz + 1) * 3 + (dy + 1);
                for (int c = 0; c < 3; c++) M.vals[e++] = hS[(size_t)(nb * 9 + 3 * r + c) * nint + n];
              }
          if (tmp) {
            const uint8_t m = mask[((int64_t)(i + 1) * L.wz + (k + 1)) * L.wy + (j + 1)];
            tmp[3 * n + r] = ((m >> r) & 1) ? 0.0 : 1.0;
          }
        }
      }
}

// stencil_assembly.c:262-364
void applyStateOperatorSubspaceMatrix(const struct gridContext gc, const int l,
                                      const struct CSRMatrix M, const CTYPE *in, CTYPE *out) {
  DevGrid *g = dev_grid(&gc);
  if (!g) return;
  const double *din = (const double *)dptr(in, "in");
  double *dout = (double *)dptr(out, "out");
  if (!din || !dout) return;
  const Level &L = g->lv[l];
  int *ro, *ci;
  double *va;
  upload_csr(M, &ro, &ci, &va);
  const int bs = 256;
  zero_noninterior<double>(*g, l, dout);
  hipLaunchKernelGGL(k_csr_apply, dim3((M.nrows + bs - 1) / bs), dim3(bs), 0, stream(), L.nx,
                     L.ny, L.nz, L.wy, L.wz, ro, ci, va, din, dout, nullptr, nullptr, 0.0,
                     (int)kApplyMode);
  sync_ok();
  dfree(ro);
  dfree(ci);
  dfree(va);
}

// stencil_solvers.c:55-90
void smoothDampedJacobiSubspaceMatrix_halo(const struct gridContext *gc,
                                           const struct CSRMatrix M, const int l,
                                           const uint_fast32_t nswp, const CTYPE omega,
                                           const MTYPE *invD, CTYPE *u, const CTYPE *b,
                                           CTYPE *tmp) {
  DevGrid *g = dev_grid(gc);
  if (!g) return;
  const double *di = (const double *)dptr(invD, "invD");
  double *du = (double *)dptr(u, "u");
  const double *db = (const double *)dptr(b, "b");
  double *dt = (double *)dptr(tmp, "tmp");
  if (!di || !du || !db || !dt) return;
  const Level &L = g->lv[l];
  int *ro, *ci;
  double *va;
  upload_csr(M, &ro, &ci, &va);
  const int bs = 256;
  sweeps(L.ndof, (int)nswp, du, dt, [&](const double *cur, double *nxt) {
    hipMemcpyAsync(nxt, cur, sizeof(double) * L.ndof, hipMemcpyDeviceToDevice, stream());
    hipLaunchKernelGGL(k_csr_apply, dim3((M.nrows + bs - 1) / bs), dim3(bs), 0, stream(), L.nx,
                       L.ny, L.nz, L.wy, L.wz, ro, ci, va, cur, nxt, db, di, (double)omega,
                       (int)kJacobiMode);
  });
  sync_ok();
  dfree(ro);
  dfree(ci);
  dfree(va);
}

// stencil_assembly.c:366-408: the coarse solver object replaces CHOLMOD's
void initializeCholmod(const struct gridContext gc, const int l,
                       struct CoarseSolverData *solverData, const struct CSRMatrix M) {
  std::memset(solverData, 0, sizeof(*solverData));
  DevGrid *g = dev_grid(&gc);
  if (!g) return;
  CoarseObj *c = new CoarseObj();
  c->l = l;
  const Level &L = g->lv[l];
  const int bw = 3 * (L.ny * L.nz + L.ny + 1) + 2;
  // the rows are the level's interior DOFs, x-planes of 3 ny nz consecutive
  const int planes = M.nrows == 3 * L.interior_nodes() ? L.nx : 0;
  if (!c->f.allocate(M.nrows, bw, false, planes)) {
    c->f.release();
    delete c;
    return;
  }
  solverData->factoredMatrix = c;
}

void finishCholmod(const struct gridContext gc, const int l, struct CoarseSolverData *solverData,
                   const struct CSRMatrix M, const int verbose) {
  CoarseObj *c = (CoarseObj *)solverData->factoredMatrix;
  if (!c) return;
  if (verbose == 1) std::printf("coarse solver: n=%d, half bandwidth %d\n", c->f.n, c->f.bw);
  hipStreamSynchronize(stream());
  c->f.release();
  delete c;
  solverData->factoredMatrix = nullptr;
  (void)gc; (void)l; (void)M;
}

// CSR rows into the coarse factor's storage (logical DOF d at pos[d])
__global__ void k_csr_lower_to_dense(int64_t nrows, const int *__restrict__ ro,
                                     const int *__restrict__ ci, const double *__restrict__ va,
                                     const int *__restrict__ pos, int64_t np,
                                     double *__restrict__ A) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  for (int e = ro[r]; e < ro[r + 1]; e++) {
    const int64_t c = ci[e];
    if (c <= r) {  // CHOLMOD stype=1 on the CSR-as-CSC reads these entries
      const int64_t pr = pos[r], pc = pos[c];
      A[pr * np + pc] = va[e];
      A[pc * np + pr] = va[e];
    }
  }
}

// stencil_assembly.c:428-437
void factorizeSubspaceMatrix(const struct gridContext gc, const int l,
                             struct CoarseSolverData solverData, const struct CSRMatrix M) {
  CoarseObj *c = (CoarseObj *)solverData.factoredMatrix;
  if (!c) {
    set_error(TOP3D_ERR_ARG, "factorizeSubspaceMatrix: initializeCholmod was not called");
    return;
  }
  int *ro, *ci;
  double *va;
  upload_csr(M, &ro, &ci, &va);
  c->f.begin_assembly();
  const int bs = 256;
  hipLaunchKernelGGL(k_csr_lower_to_dense, dim3((c->f.n + bs - 1) / bs), dim3(bs), 0, stream(),
                     (int64_t)c->f.n, ro, ci, va, c->f.pos, (int64_t)c->f.np, c->f.A);
  c->f.factor();
  dfree(ro);
  dfree(ci);
  dfree(va);
  (void)gc; (void)l;
}

// stencil_assembly.c:439-503
void solveSubspaceMatrix(const struct gridContext gc, const int l,
                         struct CoarseSolverData solverData, const CTYPE *in, CTYPE *out) {
  CoarseObj *c = (CoarseObj *)solverData.factoredMatrix;
  DevGrid *g = dev_grid(&gc);
  if (!c || !g) return;
  const double *din = (const double *)dptr(in, "in");
  double *dout = (double *)dptr(out, "out");
  if (!din || !dout) return;
  launch_coarse_apply<double>(*g, l, c->f, din, dout);
  sync_ok();
}

// ---- design loop -------------------------------------------------------------

// gpu_definitions.c:94-110
DTYPE compute_volume(const struct gridContext gc, DTYPE *xPhys) {
  DevGrid *g = dev_grid(&gc);
  if (!g) return 0;
  DevView v;
  if (!v.open(xPhys, sizeof(double) * g->nelem, true, false)) return 0;
  double *red = (double *)g->work(kSlotRed, sizeof(double) * (1024 + 16));
  launch_volume(*g, (const double *)v.dev, red, red + 1024);
  double h = 0;
  hipMemcpyAsync(&h, red + 1024, sizeof(double), hipMemcpyDeviceToHost, stream());
  sync_ok();
  v.close();
  return h;
}

// gpu_definitions.c:112-125 (x <- xnew; float result as in the reference)
float compute_change(DTYPE *x, DTYPE *xnew, const int nelem) {
  DevView vx, vn;
  if (!vx.open(x, sizeof(double) * nelem, true, true)) return 0;
  if (!vn.open(xnew, sizeof(double) * nelem, true, false)) return 0;
  double *red = (double *)dalloc(sizeof(double) * (1024 + 16));
  launch_change(nelem, (double *)vx.dev, (const double *)vn.dev, red, red + 1024);
  double h = 0;
  hipMemcpyAsync(&h, red + 1024, sizeof(double), hipMemcpyDeviceToHost, stream());
  sync_ok();
  dfree(red);
  vn.close();
  vx.close();
  return (float)h;
}

// gpu_definitions.c:127-157
void update_solution(const DTYPE *x, DTYPE *xnew, const DTYPE *dv, const DTYPE *dc,
                     const int nelem, const DTYPE gval) {
  DevView vx, vn, vv, vc;
  if (!vx.open(x, sizeof(double) * nelem, true, false) ||
      !vn.open(xnew, sizeof(double) * nelem, false, true) ||
      !vv.open(dv, sizeof(double) * nelem, true, false) ||
      !vc.open(dc, sizeof(double) * nelem, true, false))
    return;
  double *scratch = (double *)dalloc(sizeof(double) * kOcScratchDoubles);
  if (scratch) {
    // a failure is recorded by oc_update (HIP error, or no convergence)
    oc_update((const double *)vx.dev, (double *)vn.dev, (const double *)vv.dev,
              (const double *)vc.dev, nelem, gval, scratch);
    dfree(scratch);
  }
  vc.close();
  vv.close();
  vn.close();
  vx.close();
}

// ---- OC update without allocation -------------------------------------------
// update_solution_noalloc keeps the bisection state and the per-workgroup
// partial sums in this file-scope array: one copy per device, fixed size
// whatever nelem is, held until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the symbol);
// it uses the device address from hipGetSymbolAddress, looked up once per
// device.
static __device__ __align__(256) double scratch[kOcScratchDoubles];
constexpr int kNoAllocMaxDevices = 64;
static double *g_scratch_dev[kNoAllocMaxDevices] = {};
static std::mutex g_scratch_mu;

static double *noalloc_scratch() {
  int dev = -1;
  if (!T3D_HIP(hipGetDevice(&dev))) return nullptr;
  if (dev < 0 || dev >= kNoAllocMaxDevices) {
    set_error(TOP3D_ERR_ARG, "update_solution_noalloc: device %d beyond %d", dev,
              kNoAllocMaxDevices);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    if (!T3D_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)))) return nullptr;
    g_scratch_dev[dev] = (double *)p;
  }
  return g_scratch_dev[dev];
}

// gpu_definitions.c:127-157 as update_solution, allocating nothing: the
// arrays must be device pointers or mapped host arrays (no staging), and the
// scratch is the per-device array above.
void update_solution_noalloc(const DTYPE *x, DTYPE *xnew, const DTYPE *dv, const DTYPE *dc,
                             const int nelem, const DTYPE gval) {
  if (nelem < 0) {
    set_error(TOP3D_ERR_ARG, "update_solution_noalloc: nelem %d < 0", nelem);
    return;
  }
  const double *dx = (const double *)dptr(x, "update_solution_noalloc x");
  double *dn = (double *)dptr(xnew, "update_solution_noalloc xnew");
  const double *ddv = (const double *)dptr(dv, "update_solution_noalloc dv");
  const double *ddc = (const double *)dptr(dc, "update_solution_noalloc dc");
  if (!dx || !dn || !ddv || !ddc) return;
  double *s = noalloc_scratch();
  if (!s) return;
  oc_update(dx, dn, ddv, ddc, nelem, gval, s);
}

// stencil_optimization.c:228-285 (runs on the GPU; host arrays are staged)
void applyDensityFilter(const struct gridContext gc, const DTYPE rmin, const DTYPE *rho,
                        DTYPE *out) {
  DevGrid *g = dev_grid(&gc);
  if (!g) return;
  DevView vr, vo;
  if (!vr.open(rho, sizeof(double) * g->nelem, true, false) ||
      !vo.open(out, sizeof(double) * g->nelem, true, true))
    return;
  launch_filter(*g, rmin, (const double *)vr.dev, (double *)vo.dev);
  sync_ok();
  vo.close();
  vr.close();
}

// stencil_optimization.c:290-395
void applyDensityFilterGradient(const struct gridContext gc, const DTYPE rmin, DTYPE *v) {
  DevGrid *g = dev_grid(&gc);
  if (!g) return;
  DevView vv;
  if (!vv.open(v, sizeof(double) * g->nelem, true, true)) return;
  double *tmp = (double *)g->work(kSlotTmp, sizeof(double) * g->nelem);
  double *tmp2 = (double *)g->work(kSlotAux, sizeof(double) * g->nelem);
  launch_filter_grad(*g, rmin, (const double *)vv.dev, tmp, tmp2);
  // interior elements take the filtered value; the halo of v is untouched
  hipLaunchKernelGGL(k_copy_interior_elems, dim3((g->nelem + 255) / 256), dim3(256), 0, stream(),
                     g->nelem, g->nelx, g->nely, g->nelz, g->ewy, g->ewz, tmp2,
                     (double *)vv.dev);
  sync_ok();
  vv.close();
}

// gpu_definitions.c:159-195: map every array the design loop uses
void enter_data(const struct gridContext gridContext, const struct SolverDataBuffer solverData,
                const struct gpuGrid gpu_gc, const DTYPE *xPhys, const DTYPE *x,
                const DTYPE *xnew, const STYPE *U, const CTYPE *F, const DTYPE *dc,
                const int nelem, const int nl) {
  Level L;
  compute_level(&gridContext, 0, &L);
  const uint64_t ne = sizeof(double) * (uint64_t)nelem, nd = sizeof(double) * (uint64_t)L.ndof;
  top3d_map(xPhys, ne);
  top3d_map(x, ne);
  top3d_map(dc, ne);
  top3d_map(xnew, ne);
  top3d_map(U, nd);
  top3d_map(F, nd);
  (void)solverData; (void)gpu_gc; (void)nl;  // solver buffers already live on the device
}

// ---- autograd hooks ----------------------------------------------------------
int top3d_compliance_forward(const struct gpuGrid *gpu_gc, struct SolverDataBuffer *data,
                             int nl, double tol, const double *xPhys_dev, const double *F_dev,
                             double *U_dev, double *c) {
  int it = 0;
  float res = 0;
  top3d_clear_error();
  solveStateMG_halo(gpu_gc, const_cast<double *>(xPhys_dev), 4, nl, tol, data, &it, &res,
                    const_cast<double *>(F_dev), U_dev);
  if (top3d_last_error() != TOP3D_OK && top3d_last_error() != TOP3D_ERR_NO_CONVERGE)
    return top3d_last_error();
  const int solve_status = top3d_last_error();
  DevGrid *g = dev_grid_node(&gpu_gc->target[0]);
  if (!g) return top3d_last_error();
  double *red = (double *)g->work(kSlotRed, sizeof(double) * (1024 + 16));
  double *dc = (double *)g->work(kSlotVec, sizeof(double) * g->nelem);
  launch_sensitivity(*g, xPhys_dev, U_dev, dc, red, red + 1024);
  hipMemcpyAsync(c, red + 1024, sizeof(double), hipMemcpyDeviceToHost, stream());
  if (!sync_ok()) return TOP3D_ERR_HIP;
  return solve_status;
}

int top3d_compliance_backward(const struct gpuGrid *gpu_gc, double rmin,
                              const double *xPhys_dev, const double *U_dev, double grad_c,
                              double *grad_x_dev) {
  DevGrid *g = dev_grid_node(&gpu_gc->target[0]);
  if (!g) return top3d_last_error();
  double *red = (double *)g->work(kSlotRed, sizeof(double) * (1024 + 16));
  double *dc = (double *)g->work(kSlotVec, sizeof(double) * g->nelem);
  launch_fill<double>(g->nelem, dc, 0.0);
  launch_sensitivity(*g, xPhys_dev, U_dev, dc, red, red + 1024);
  hipLaunchKernelGGL(k_scale, dim3((g->nelem + 255) / 256), dim3(256), 0, stream(), g->nelem,
                     grad_c, dc);
  if (rmin > 0) {
    double *tmp = (double *)g->work(kSlotTmp, sizeof(double) * g->nelem);
    launch_fill<double>(g->nelem, grad_x_dev, 0.0);
    launch_filter_grad(*g, rmin, dc, tmp, grad_x_dev);
  } else {
    hipMemcpyAsync(grad_x_dev, dc, sizeof(double) * g->nelem, hipMemcpyDeviceToDevice, stream());
  }
  return sync_ok() ? TOP3D_OK : TOP3D_ERR_HIP;
}

}  // extern "C"
