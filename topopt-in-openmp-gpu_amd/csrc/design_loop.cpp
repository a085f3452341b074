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

// The design arrays of the loop, host arrays in the padded element layout
// (mapped to the device by enter_data), freed together.
struct DesignArrays {
  uint64_t count = 0;
  DTYPE *sens = nullptr;     // compliance sensitivity, then its filtered form
  DTYPE *vol_sens = nullptr; // volume sensitivity through the filter
  DTYPE *design = nullptr;   // design variables
  DTYPE *phys = nullptr;     // filtered densities the solve sees
  DTYPE *next = nullptr;     // the OC update's proposal
  bool allocate(uint64_t n, DTYPE fill_value, const struct gridContext &grid) {
    count = n;
    sens = (DTYPE *)std::calloc(n, sizeof(DTYPE));
    vol_sens = (DTYPE *)std::malloc(sizeof(DTYPE) * n);
    design = (DTYPE *)std::calloc(n, sizeof(DTYPE));
    phys = (DTYPE *)std::calloc(n, sizeof(DTYPE));
    next = (DTYPE *)std::calloc(n, sizeof(DTYPE));
    if (!sens || !vol_sens || !design || !phys || !next) return false;
    for (uint64_t e = 0; e < n; e++) vol_sens[e] = 1.0;
    // uniform start on the real elements, the halo at 0 (stencil_optimization.c:103-119)
    const uint64_t sy = grid.wrapy - 1, sz = grid.wrapz - 1;
    for (uint64_t i = 1; i <= grid.nelx; i++)
      for (uint64_t k = 1; k <= grid.nelz; k++)
        for (uint64_t j = 1; j <= grid.nely; j++) {
          design[i * sy * sz + k * sy + j] = fill_value;
          phys[i * sy * sz + k * sy + j] = fill_value;
        }
    return true;
  }
  void release() {
    for (DTYPE *p : {sens, vol_sens, design, phys, next}) std::free(p);
  }
};

void top3dmgcg(const uint_fast32_t nelx, const uint_fast32_t nely, const uint_fast32_t nelz,
               const DTYPE volfrac, const DTYPE rmin, const uint_fast32_t nl, const float cgtol,
               const uint_fast32_t cgmax, const int verbose, const int write_result,
               const int max_iterations) {
  (void)cgmax;  // the reference passes it and never reads it either
  const int levels = (int)nl;
  struct gridContext grid;
  setupGC(&grid, levels, (int)nelx, (int)nely, (int)nelz);
  if (top3d_last_error()) {
    std::fprintf(stderr, "top3d: %s\n", top3d_last_error_string());
    return;
  }
  struct gpuGrid devices;
  devices.num_targets = 1;
  gridContextToGPUGrid(&grid, &devices, levels, verbose);
  const uint64_t n_elem = (uint64_t)(grid.wrapx - 1) * (grid.wrapy - 1) * (grid.wrapz - 1);
  CTYPE *load;
  STYPE *disp;
  allocateZeroPaddedStateField(grid, 0, &load);
  allocateZeroPaddedStateField_STYPE(grid, 0, &disp);
  top3d_cantilever_load(&grid, load);
  DesignArrays arr;
  if (!arr.allocate(n_elem, volfrac, grid)) {
    std::fprintf(stderr, "top3d: out of host memory\n");
    arr.release();
    return;
  }
  struct SolverDataBuffer solver;
  allocateSolverData(grid, levels, &solver);
  top3d_set_precision(&solver, g_default_precision);
  top3d_set_bc_mode(&solver, g_default_bc);
  // the volume sensitivity is filtered once, before the loop
  top3d_map(arr.vol_sens, sizeof(DTYPE) * n_elem);
  applyDensityFilterGradient(grid, rmin, arr.vol_sens);
  const auto t_start = std::chrono::steady_clock::now();
  auto seconds = [&]() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  };
  enter_data(grid, solver, devices, arr.phys, arr.design, arr.next, disp, load, arr.sens,
             (int)n_elem, levels);
  if (top3d_last_error()) std::fprintf(stderr, "top3d: %s\n", top3d_last_error_string());
  const DTYPE n_real = (DTYPE)(grid.nelx * grid.nely * grid.nelz);
  unsigned int iteration = 0;
  float max_change = 1;
  DTYPE mean_density = 0.0, compliance = 0.0;
  int cg_total = 0;
  // solve, volume, sensitivity, filtered sensitivity, OC, change, filter
  // (stencil_optimization.c:148-176)
  while (max_change > 1e-2 && iteration < (unsigned)max_iterations) {
    ++iteration;
    int cg_last = 0;
    float res_last = 0;
    solveStateMG_halo(&devices, arr.phys, 4, levels, cgtol, &solver, &cg_last, &res_last, load,
                      disp);
    mean_density = compute_volume(grid, arr.phys) / n_real;
    getComplianceAndSensetivity_halo(&devices.target[0], arr.phys, disp, &compliance,
                                     arr.sens);
    applyDensityFilterGradient(grid, rmin, arr.sens);
    update_solution(arr.design, arr.next, arr.vol_sens, arr.sens, (int)n_elem,
                    mean_density - volfrac);
    max_change = compute_change(arr.design, arr.next, (int)n_elem);
    applyDensityFilter(grid, rmin, arr.design, arr.phys);
    cg_total += cg_last;
    // the reference's per-iteration line (stencil_optimization.c:181-185)
    if (verbose == 1)
      std::printf("It.:%4i Obj.:%6.3e Vol.:%6.3f ch.:%4.2f relres: %4.2e iters: %4i time: %6.3f \n",
                  iteration, compliance, mean_density, max_change, res_last, cg_last, seconds());
  }
  // the reference's summary line (stencil_optimization.c:193)
  std::printf("%4i %12.6f %6.3f %4.2f %9i %9i %6.3f \n", iteration, compliance, mean_density,
              max_change, (int)(nelx * nely * nelz), cg_total, seconds());
  if (verbose == 1) std::printf("End time: %6.3f \n", seconds());
  if (write_result == 1) {
    top3d_update_from(arr.phys, sizeof(DTYPE) * n_elem);
    char path[64];
    std::snprintf(path, sizeof(path), "out_%d_%d_%d.vtu", (int)nelx, (int)nely, (int)nelz);
    writeDensityVtkFile((int)nelx, (int)nely, (int)nelz, arr.phys, path);
    std::snprintf(path, sizeof(path), "out_%d_%d_%d_halo.vtu", (int)nelx, (int)nely, (int)nelz);
    writeDensityVtkFileWithHalo((int)nelx, (int)nely, (int)nelz, arr.phys, path);
  }
  for (const void *p : {(const void *)arr.phys, (const void *)arr.design, (const void *)arr.sens,
                        (const void *)arr.next, (const void *)disp, (const void *)load,
                        (const void *)arr.vol_sens})
    top3d_unmap(p);
  freeSolverData(&solver, levels);
  freeGC(&grid, levels);
  freeGPUGrid(&devices, levels);
  std::free(load);
  std::free(disp);
  arr.release();
}

}  // extern "C"
