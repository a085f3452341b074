// SIMP/OC design loop (top3dmgcg, stencil_optimization.c:13-223) on the
// library's entry points.  Every array is mapped once (enter_data) and the
// loop runs on the device; the reference's host filter and its two
// `target update` transfers per iteration are gone.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "t3d_internal.hpp"

namespace {
int g_default_precision = TOP3D_FP64;
int g_default_bc = TOP3D_BC_REFERENCE;
}  // namespace

extern "C" {

void top3d_set_default_precision(int mode) { g_default_precision = mode; }
void top3d_set_default_bc_mode(int mode) { g_default_bc = mode; }

// stencil_optimization.c:39-95 (cantilever load on the z=0 face near x=nelx)
void top3d_cantilever_load(const struct gridContext *gc, CTYPE *F) {
  const int nelx = (int)gc->nelx, nely = (int)gc->nely;
  const uint64_t wy = gc->wrapy, wz = gc->wrapz;
  const double radius = ((double)(nely + 1)) / 5.0, r2 = radius * radius;
  const double cx = (double)nelx, cy = ((double)nely - 1.0) / 2.0;
  int count = 0;
  for (int i = 0; i < nelx; i++)
    for (int j = 0; j < nely; j++) {
      const double dx = i - cx, dy = j - cy;
      if (dx * dx + dy * dy < r2) count++;
    }
  const double f = -1.0 / (4.0 * (double)count);
  const uint64_t kk = 1;
  for (int i = 0; i < nelx; i++)
    for (int j = 0; j < nely; j++) {
      const double dx = i - cx, dy = j - cy;
      if (dx * dx + dy * dy < r2) {
        const uint64_t ii = i + 1, jj = j + 1;
        F[3 * ((ii + 1) * wy * wz + wy * kk + (jj + 1)) + 2] += f;
        F[3 * ((ii + 1) * wy * wz + wy * kk + jj) + 2] += f;
        F[3 * (ii * wy * wz + wy * kk + (jj + 1)) + 2] += f;
        F[3 * (ii * wy * wz + wy * kk + jj) + 2] += f;
      }
    }
}

void top3d_line_load(const struct gridContext *gc, CTYPE *F) {
  top3d_line_load_profile(gc, F, TOP3D_LOAD_CONSTANT);
}

void top3d_line_load_profile(const struct gridContext *gc, CTYPE *F, int profile) {
  const uint64_t wy = gc->wrapy, wz = gc->wrapz;
  const uint64_t i = gc->nelx + 1, k = 1, ny = gc->nely;
  for (uint64_t jj = 0; jj <= ny; jj++) {
    double w = 1.0;
    if (profile == TOP3D_LOAD_END_HALF && (jj == 0 || jj == ny)) w = 0.5;
    if (profile == TOP3D_LOAD_SINE) w = std::sin(M_PI * (double)jj / (double)ny);
    F[3 * (i * wy * wz + k * wy + (jj + 1)) + 2] -= w;
  }
}

void top3dmgcg(const uint_fast32_t nelx, const uint_fast32_t nely, const uint_fast32_t nelz,
               const DTYPE volfrac, const DTYPE rmin, const uint_fast32_t nl, const float cgtol,
               const uint_fast32_t cgmax, const int verbose, const int write_result,
               const int max_iterations) {
  (void)cgmax;  // passed but unused by the reference too
  struct gridContext gridContext;
  setupGC(&gridContext, (int)nl, (int)nelx, (int)nely, (int)nelz);
  if (top3d_last_error()) {
    std::fprintf(stderr, "top3d: %s\n", top3d_last_error_string());
    return;
  }
  struct gpuGrid gpu_gc;
  gpu_gc.num_targets = 1;
  gridContextToGPUGrid(&gridContext, &gpu_gc, (int)nl, verbose);
  const uint64_t nelem = (uint64_t)(gridContext.wrapx - 1) * (gridContext.wrapy - 1) *
                         (gridContext.wrapz - 1);
  CTYPE *F;
  STYPE *U;
  allocateZeroPaddedStateField(gridContext, 0, &F);
  allocateZeroPaddedStateField_STYPE(gridContext, 0, &U);
  top3d_cantilever_load(&gridContext, F);
  DTYPE *dc = (DTYPE *)std::calloc(nelem, sizeof(DTYPE));
  DTYPE *dv = (DTYPE *)std::malloc(sizeof(DTYPE) * nelem);
  DTYPE *x = (DTYPE *)std::calloc(nelem, sizeof(DTYPE));
  DTYPE *xPhys = (DTYPE *)std::calloc(nelem, sizeof(DTYPE));
  DTYPE *xnew = (DTYPE *)std::calloc(nelem, sizeof(DTYPE));
  DTYPE c = 0.0;
  for (uint64_t i = 0; i < nelem; i++) dv[i] = 1.0;
  for (uint64_t i = 1; i < gridContext.nelx + 1; i++)
    for (uint64_t k = 1; k < gridContext.nelz + 1; k++)
      for (uint64_t j = 1; j < gridContext.nely + 1; j++) {
        const uint64_t idx = i * (gridContext.wrapy - 1) * (gridContext.wrapz - 1) +
                             k * (gridContext.wrapy - 1) + j;
        x[idx] = volfrac;
        xPhys[idx] = volfrac;
      }
  struct SolverDataBuffer solverData;
  allocateSolverData(gridContext, (int)nl, &solverData);
  top3d_set_precision(&solverData, g_default_precision);
  top3d_set_bc_mode(&solverData, g_default_bc);
  top3d_map(dv, sizeof(DTYPE) * nelem);
  applyDensityFilterGradient(gridContext, rmin, dv);
  const auto start = std::chrono::steady_clock::now();
  auto wtime = [&]() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
  };
  enter_data(gridContext, solverData, gpu_gc, xPhys, x, xnew, U, F, dc, (int)nelem, (int)nl);
  if (top3d_last_error()) std::fprintf(stderr, "top3d: %s\n", top3d_last_error_string());
  unsigned int loop = 0;
  float change = 1;
  DTYPE vol = 0.0;
  int gciter_total = 0;
  while ((change > 1e-2) && (loop < (unsigned)max_iterations)) {
    loop++;
    int cgiter = 0;
    float cgres = 0;
    solveStateMG_halo(&gpu_gc, xPhys, 4, (int)nl, cgtol, &solverData, &cgiter, &cgres, F, U);
    vol = compute_volume(gridContext, xPhys);
    getComplianceAndSensetivity_halo(&gpu_gc.target[0], xPhys, U, &c, dc);
    applyDensityFilterGradient(gridContext, rmin, dc);
    vol /= (DTYPE)(gridContext.nelx * gridContext.nely * gridContext.nelz);
    const DTYPE g = vol - volfrac;
    update_solution(x, xnew, dv, dc, (int)nelem, g);
    change = compute_change(x, xnew, (int)nelem);
    applyDensityFilter(gridContext, rmin, x, xPhys);
    gciter_total += cgiter;
    if (verbose == 1)
      std::printf("It.:%4i Obj.:%6.3e Vol.:%6.3f ch.:%4.2f relres: %4.2e iters: %4i time: %6.3f \n",
                  loop, c, vol, change, cgres, cgiter, wtime());
  }
  std::printf("%4i %12.6f %6.3f %4.2f %9i %9i %6.3f \n", loop, c, vol, change,
              (int)(nelx * nely * nelz), gciter_total, wtime());
  if (verbose == 1) std::printf("End time: %6.3f \n", wtime());
  if (write_result == 1) {
    top3d_update_from(xPhys, sizeof(DTYPE) * nelem);
    char name1[60], name2[60];
    std::snprintf(name1, 60, "out_%d_%d_%d.vtu", (int)nelx, (int)nely, (int)nelz);
    writeDensityVtkFile((int)nelx, (int)nely, (int)nelz, xPhys, name1);
    std::snprintf(name2, 60, "out_%d_%d_%d_halo.vtu", (int)nelx, (int)nely, (int)nelz);
    writeDensityVtkFileWithHalo((int)nelx, (int)nely, (int)nelz, xPhys, name2);
  }
  for (void *p : {(void *)xPhys, (void *)x, (void *)dc, (void *)xnew, (void *)U, (void *)F,
                  (void *)dv})
    top3d_unmap(p);
  freeSolverData(&solverData, (int)nl);
  freeGC(&gridContext, (int)nl);
  freeGPUGrid(&gpu_gc, (int)nl);
  std::free(F);
  std::free(U);
  std::free(dc);
  std::free(dv);
  std::free(x);
  std::free(xPhys);
  std::free(xnew);
}

}  // extern "C"
