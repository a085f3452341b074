/*
 * top3d_hip.h -- C ABI of the MI355X (gfx950) MGCG state-solve library.
 *
 * Drop-in replacement for the device path of topopt/TopOpt-in-OpenMP-GPU:
 * the reference's structs and entry points keep their names, field order and
 * argument meaning; each declaration cites the reference declaration it
 * replaces.  Element types are the reference's typedef names fixed to the
 * FP64 configuration (definitions.h:25-29 re-typed): every array that crosses
 * this boundary is double.  The mixed-precision solve keeps FP32 copies
 * inside the library (see top3d_set_precision).
 *
 * Memory: the reference hands host pointers to OpenMP target regions and
 * relies on the present table filled by `#pragma omp target enter data`.
 * This library keeps its own present table: enter_data() / top3d_map()
 * create device mirrors of host arrays, top3d_update_to/from() replace
 * `#pragma omp target update`, and every entry point accepts either a mapped
 * host pointer (its mirror is used) or a device pointer.  Unmapped host
 * pointers are rejected (top3d_last_error() reports it).
 *
 * Errors: entry points keep the reference's void signatures; failures are
 * recorded in a thread-local status read with top3d_last_error().
 * All device work is ordered on one HIP stream (top3d_set_stream).
 */
#ifndef TOP3D_HIP_H
#define TOP3D_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* definitions.h:25-29, FP64 configuration */
typedef double MTYPE;
typedef double STYPE;
typedef double CTYPE;
typedef double DTYPE;

/* definitions.h:17,23 */
#define STENCIL_SIZE_Y 128
#define number_of_matrix_free_levels 2

/* definitions.h:31-34 */
struct FixedDofs {
  uint_fast32_t n;
  uint_fast32_t *idx;
};

/* definitions.h:36-53 */
struct gridContext {
  double E0;
  double Emin;
  double nu;
  double elementSizeX;
  double elementSizeY;
  double elementSizeZ;
  uint_fast32_t nelx;
  uint_fast32_t nely;
  uint_fast32_t nelz;
  uint_fast32_t wrapx;
  uint_fast32_t wrapy;
  uint_fast32_t wrapz;
  double penal;
  MTYPE **precomputedKE;
  struct FixedDofs *fixedDofs;
};

/* stencil_assembly.h:8-15 */
struct CSRMatrix {
  uint64_t nnz;
  int32_t nrows;
  int *rowOffsets;
  int *colIndex;
  MTYPE *vals;
};

/* stencil_assembly.h:17-27.  The CHOLMOD handles of the reference become
 * opaque slots; `cholmodCommon` holds this library's coarse-solver object.  */
struct CoarseSolverData {
  void *cholmodCommon;
  void *sparseMatrix;
  void *factoredMatrix;
  void *rhs;
  void *solution;
  void *Y_workspace;
  void *E_workspace;
};

/* gpu_definitions.h:7-23.  r,p,q,z and the per-level arrays hold DEVICE
 * pointers owned by the library (allocateSolverData).                        */
struct SolverDataBuffer {
  CTYPE *r;
  CTYPE *p;
  CTYPE *q;
  CTYPE *z;
  MTYPE **invD;
  CTYPE **dmg;
  CTYPE **rmg;
  CTYPE **zmg;
  struct CSRMatrix *coarseMatrices;
  struct CoarseSolverData bottomSolver;
};

/* gpu_definitions.h:25-37 */
struct gpuNode {
  int id;
  struct gpuNode *prev;
  struct gpuNode *next;
  struct gridContext *gc;
  int offset_x;
  int x_global;
};

struct gpuGrid {
  int num_targets;
  struct gpuNode *target;
};

/* ---- grid set-up (stencil_grid_utility.h:5-17, gpu_definitions.h:39-56) -- */
void setupGC(struct gridContext *gc, const int nl, const int nelx,
             const int nely, const int nelz);
void freeGC(struct gridContext *gc, const int nl);
void setFixedDof_halo(struct gridContext *gc, const int l);
void allocateZeroPaddedStateField(const struct gridContext gc, const int l,
                                  CTYPE **v);
void allocateZeroPaddedStateField_MTYPE(const struct gridContext gc,
                                        const int l, MTYPE **v);
void allocateZeroPaddedStateField_STYPE(const struct gridContext gc,
                                        const int l, STYPE **v);
void computePadding(const struct gridContext *gc, const int l, int *ncell,
                    int *wrapx, int *wrapy, int *wrapz, uint_fast32_t *ndof);
void gridContextToGPUGrid(struct gridContext *gc, struct gpuGrid *gpu_gc,
                          const int nl, const int verbose);
void freeGPUGrid(struct gpuGrid *gpu_gc, const int nl);
/* local_matrix.h:19-21 */
void getKEsubspace(MTYPE *KEarray, const MTYPE nu, const int l);

/* ---- matrix-free operators (stencil_methods.h:6-28) ---------------------- */
void applyStateOperator_stencil(const struct gpuNode *gpu_node,
                                const DTYPE *x, const CTYPE *in, CTYPE *out);
void applyStateOperatorSubspace_halo(const struct gpuNode *gpu_node,
                                     const int l, const DTYPE *x, CTYPE *in,
                                     CTYPE *out);
void getComplianceAndSensetivity_halo(const struct gpuNode *gpu_node,
                                      const DTYPE *x, STYPE *u, DTYPE *c,
                                      DTYPE *dcdx);
void projectToFinerGrid_halo(const struct gpuNode *gpu_node, const int l,
                             const CTYPE *ucoarse, CTYPE *ufine);
void projectToCoarserGrid_halo(const struct gpuNode *gpu_node, const int l,
                               const CTYPE *ufine, CTYPE *ucoarse);
void assembleInvertedMatrixDiagonalSubspace_halo(
    const struct gpuNode *gpu_node, const DTYPE *x, const int l, MTYPE *diag);

/* ---- smoothers and solver (stencil_solvers.h:8-32) ----------------------- */
void smoothDampedJacobi_halo(const struct gpuNode *gpu_node, const DTYPE *x,
                             const uint_fast32_t nswp, const CTYPE omega,
                             const MTYPE *invD, CTYPE *u, const CTYPE *b,
                             CTYPE *tmp);
void smoothDampedJacobiSubspace_halo(const struct gpuNode *gpu_node,
                                     const DTYPE *x, const int l,
                                     const uint_fast32_t nswp,
                                     const CTYPE omega, const MTYPE *invD,
                                     CTYPE *u, const CTYPE *b, CTYPE *tmp);
void smoothDampedJacobiSubspaceMatrix_halo(
    const struct gridContext *gc, const struct CSRMatrix M, const int l,
    const uint_fast32_t nswp, const CTYPE omega, const MTYPE *invD, CTYPE *u,
    const CTYPE *b, CTYPE *tmp);
void solveStateMG_halo(const struct gpuGrid *gpu_gc, DTYPE *x, const int nswp,
                       const int nl, const CTYPE tol,
                       struct SolverDataBuffer *data, int *finalIter,
                       float *finalRes, CTYPE *b, STYPE *u);
void allocateSolverData(const struct gridContext gc, const int nl,
                        struct SolverDataBuffer *data);
void freeSolverData(struct SolverDataBuffer *data, const int nl);

/* ---- assembled coarse levels + direct solve (stencil_assembly.h:29-55) ---
 * CSR arrays are host arrays in the reference's layout; assembly, SpMV,
 * factorization and solve run on the GPU.                                    */
void allocateSubspaceMatrix(const struct gridContext gc, const int l,
                            struct CSRMatrix *M);
void freeSubspaceMatrix(struct CSRMatrix *M);
void assembleSubspaceMatrix(const struct gridContext gc, const int l,
                            const DTYPE *x, struct CSRMatrix M, MTYPE *tmp);
void applyStateOperatorSubspaceMatrix(const struct gridContext gc, const int l,
                                      const struct CSRMatrix M,
                                      const CTYPE *in, CTYPE *out);
void initializeCholmod(const struct gridContext gc, const int l,
                       struct CoarseSolverData *solverData,
                       const struct CSRMatrix M);
void finishCholmod(const struct gridContext gc, const int l,
                   struct CoarseSolverData *solverData,
                   const struct CSRMatrix M, const int verbose);
void factorizeSubspaceMatrix(const struct gridContext gc, const int l,
                             struct CoarseSolverData solverData,
                             const struct CSRMatrix M);
void solveSubspaceMatrix(const struct gridContext gc, const int l,
                         struct CoarseSolverData solverData, const CTYPE *in,
                         CTYPE *out);

/* ---- design loop kernels (gpu_definitions.h:58-86,
 *      stencil_optimization.h:5-15) ---------------------------------------- */
DTYPE compute_volume(const struct gridContext gc, DTYPE *xPhys);
float compute_change(DTYPE *x, DTYPE *xnew, const int nelem);
/* OC bisection (gpu_definitions.c:127-157) on the device, three bisection
 * steps per pass over the elements; x_e sqrt(-dc_e / (dv_e lmid)) is formed
 * as x_e sqrt(-dc_e / dv_e) / sqrt(lmid), within a few ulp of the one-step
 * expression. */
void update_solution(const DTYPE *x, DTYPE *xnew, const DTYPE *dv,
                     const DTYPE *dc, const int nelem, const DTYPE g);
/* update_solution for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments and result; x, xnew, dv
 * and dc must be device pointers or mapped host arrays (unmapped host arrays
 * are rejected, TOP3D_ERR_UNMAPPED, instead of being staged).  It calls no
 * hipMalloc/hipFree: the bisection state and the per-workgroup partial sums
 * live in a per-device static array of 8 + 15 * 1024 doubles (123 KB, the
 * same for every nelem; nelem < 0 is rejected), held until the process
 * ends.  Calls on one GPU share that array, so calls that overlap on one GPU
 * (two host threads, or two streams) race; each call initialises the array
 * itself.                                                                    */
void update_solution_noalloc(const DTYPE *x, DTYPE *xnew, const DTYPE *dv,
                             const DTYPE *dc, const int nelem, const DTYPE g);
void enter_data(const struct gridContext gridContext,
                const struct SolverDataBuffer solverData,
                const struct gpuGrid gpu_gc, const DTYPE *xPhys,
                const DTYPE *x, const DTYPE *xnew, const STYPE *U,
                const CTYPE *F, const DTYPE *dc, const int nelem,
                const int nl);
void applyDensityFilter(const struct gridContext gc, const DTYPE rmin,
                        const DTYPE *rho, DTYPE *out);
void applyDensityFilterGradient(const struct gridContext gc, const DTYPE rmin,
                                DTYPE *v);
void top3dmgcg(const uint_fast32_t nelx, const uint_fast32_t nely,
               const uint_fast32_t nelz, const DTYPE volfrac,
               const DTYPE rmin, const uint_fast32_t nl, const float cgtol,
               const uint_fast32_t cgmax, const int verbose,
               const int write_result, const int max_iterations);
/* stencil_optimization.h:17-22; densityArray is a host array in this
 * library's padded element layout */
void writeDensityVtkFile(const int nelx, const int nely, const int nelz,
                         const DTYPE *densityArray, const char *filename);
void writeDensityVtkFileWithHalo(const int nelx, const int nely,
                                 const int nelz, const DTYPE *densityArray,
                                 const char *filename);

/* ======================================================================== */
/* Entry points with no counterpart in the reference.                        */

/* status codes of top3d_last_error() */
#define TOP3D_OK 0
#define TOP3D_ERR_HIP 1          /* a HIP call failed */
#define TOP3D_ERR_ARG 2          /* invalid argument / unsupported layout */
#define TOP3D_ERR_UNMAPPED 3     /* host pointer without device mirror */
#define TOP3D_ERR_NOT_SPD 4      /* coarse Cholesky met a non-positive pivot */
#define TOP3D_ERR_NO_CONVERGE 5  /* CG stopped at maxIter above tol */
#define TOP3D_ERR_NONFINITE 6    /* NaN/Inf in the CG recurrence */
int top3d_last_error(void);
const char *top3d_last_error_string(void);
void top3d_clear_error(void);

/* precision of solveStateMG_halo for a solver buffer:
 * TOP3D_FP64  every vector, operator and reduction in double (plain form);
 * TOP3D_MIXED V-cycle (level operators, smoothers, transfers, level vectors,
 *             densities) in FP32; outer CG, its matvec, dot products and the
 *             coarse Cholesky in FP64 (r.z, whose z is the V-cycle's FP32
 *             output, reads the FP32 copy of r and sums in FP64).            */
#define TOP3D_FP64 0
#define TOP3D_MIXED 1
void top3d_set_precision(struct SolverDataBuffer *data, int mode);
/* boundary-condition convention of the matrix-free levels:
 * TOP3D_BC_REFERENCE  fixed rows replaced by identity, columns kept
 *                     (stencil_methods.c:424-427, 515-518);
 * TOP3D_BC_SYMMETRIC  columns of fixed DOFs dropped as well (symmetric
 *                     V-cycle; converges on heterogeneous densities).       */
#define TOP3D_BC_REFERENCE 0
#define TOP3D_BC_SYMMETRIC 1
void top3d_set_bc_mode(struct SolverDataBuffer *data, int mode);
void top3d_set_max_iter(struct SolverDataBuffer *data, int max_iter);
/* modes top3dmgcg() applies to the solver buffer it allocates */
void top3d_set_default_precision(int mode);
void top3d_set_default_bc_mode(int mode);
/* the cantilever load of top3dmgcg (stencil_optimization.c:39-95), added
 * into F (host array in the level-0 halo layout) */
void top3d_cantilever_load(const struct gridContext *gc, CTYPE *F);
/* the line load of the MATLAB specification (top3dmgcg_reference.m:59):
 * F_z -= 1 at every node of the edge x = nelx, z = 0, added into F */
void top3d_line_load(const struct gridContext *gc, CTYPE *F);
/* the same edge with the specification's other profiles, F_z -= w_j at its
 * node j = 0..nely (top3dmgcg_reference.m:57-59):
 * TOP3D_LOAD_CONSTANT  w_j = 1 (as top3d_line_load);
 * TOP3D_LOAD_END_HALF  w_j = 1, and 1/2 at the two end nodes;
 * TOP3D_LOAD_SINE      w_j = sin(pi j / nely).                           */
#define TOP3D_LOAD_CONSTANT 0
#define TOP3D_LOAD_END_HALF 1
#define TOP3D_LOAD_SINE 2
void top3d_line_load_profile(const struct gridContext *gc, CTYPE *F, int profile);
/* supports that setFixedDof_halo (stencil_grid_utility.c:7-65), and so
 * setupGC, builds from this call on:
 * TOP3D_SUPPORT_STRIPS  two clamped strips of the x=0 face, y in
 *                       [0, nely/4] and [nely - nely/4, nely] (as shipped);
 * TOP3D_SUPPORT_FACE    the whole x=0 face clamped, the variant kept in
 *                       comments at stencil_grid_utility.c:19-32           */
#define TOP3D_SUPPORT_STRIPS 0
#define TOP3D_SUPPORT_FACE 1
void top3d_set_default_support(int mode);
int top3d_get_default_support(void);

/* Layout adapter for arrays in the reference's padding.  The reference pads
 * y so that nely_l + 1 + pad is a multiple of STENCIL_SIZE_Y = 128
 * (computePadding, gpu_definitions.c:67-92); this library's computePadding
 * does not pad y.  Node (i, j, k) and element (i, j, k) keep their halo
 * indices in both, so these copy each (i, k) row; the destination's
 * padding is zero-filled.  Host arrays; node arrays have 3 DOFs per node,
 * element arrays are level 0.  top3d_reference_padding gives the
 * reference's wrap sizes and ndof for level l.  No counterpart in the
 * reference.                                                              */
void top3d_reference_padding(const struct gridContext *gc, int l, int *wrapx,
                             int *wrapy, int *wrapz, uint64_t *ndof);
void top3d_nodes_from_reference_layout(const struct gridContext *gc, int l,
                                       const double *ref, double *lib);
void top3d_nodes_to_reference_layout(const struct gridContext *gc, int l,
                                     const double *lib, double *ref);
void top3d_elements_from_reference_layout(const struct gridContext *gc,
                                          const double *ref, double *lib);
void top3d_elements_to_reference_layout(const struct gridContext *gc,
                                        const double *lib, double *ref);

/* present table (replaces `#pragma omp target enter/exit data/update`) */
int top3d_map(const void *host, uint64_t bytes);          /* alloc + copy in */
int top3d_map_alloc(const void *host, uint64_t bytes);    /* alloc, no copy */
int top3d_unmap(const void *host);
int top3d_update_to(const void *host, uint64_t bytes);
int top3d_update_from(void *host, uint64_t bytes);
void *top3d_device_ptr(const void *host_or_device);

/* Device memory: what the library holds (every allocation it makes: solver
 * buffers, level data, coarse factor, present-table mirrors, scratch) now
 * and at most since it was loaded or the peak was last reset, and the
 * device-wide use from hipMemGetInfo (total - free: every allocator and
 * process on the device).  reset_peak != 0 restarts the peak at the current
 * holding after reading it.  No counterpart in the reference, whose device
 * memory is OpenMP's present table.                                         */
struct Top3dMemoryInfo {
  uint64_t library_bytes;
  uint64_t library_peak_bytes;
  uint64_t device_used_bytes;
  uint64_t device_total_bytes;
};
int top3d_memory_info(struct Top3dMemoryInfo *out, int reset_peak);

/* stream on which all device work is ordered (0 = the library's own) */
void top3d_set_stream(void *hip_stream);
void *top3d_get_stream(void);
int top3d_synchronize(void);

/* Statistics of the last solveStateMG_halo on a buffer. */
struct Top3dSolveStats {
  int iterations;        /* CG iterations run (finalIter + 1) */
  double relres;         /* final relative residual (FP64) */
  double setup_ms;       /* host time to queue the per-solve set-up (the
                            coarse levels and factor run on a second stream,
                            overlapping the first V-cycle) */
  double solve_ms;       /* CG loop */
  int status;            /* TOP3D_OK or an error code */
};
void top3d_solver_stats(const struct SolverDataBuffer *data,
                        struct Top3dSolveStats *out);

/* Compliance c(x) = sum_e E(x_e) u_e^T KE u_e with u from the solve and its
 * gradient with respect to the filtered densities, written for callers that
 * hold device arrays (the torch.autograd wrapper):
 *   forward : xPhys (device, padded element array) -> u (device), c (host)
 *   backward: dc/dxPhys_e = -3 (E0-Emin) xPhys_e^2 u_e^T KE u_e, scaled by gc
 *             and pulled back through the density filter: grad = H^T (dc/Hs).
 * rmin <= 0 skips the filter.                                                */
int top3d_compliance_forward(const struct gpuGrid *gpu_gc,
                             struct SolverDataBuffer *data, int nl,
                             double tol, const double *xPhys_dev,
                             const double *F_dev, double *U_dev, double *c);
int top3d_compliance_backward(const struct gpuGrid *gpu_gc, double rmin,
                              const double *xPhys_dev, const double *U_dev,
                              double grad_c, double *grad_x_dev);

/* Times `reps` back-to-back launches of one hot kernel on the buffers of the
 * last solve (HIP events on the library stream).  `which`:
 *   0 fine damped-Jacobi sweep, FP32 (mixed V-cycle)
 *   1 fine damped-Jacobi sweep, FP64
 *   2 fine operator, FP64 (the outer CG matvec q = K p)
 *   3 level-1 damped-Jacobi sweep, FP32
 *   4 level-1 damped-Jacobi sweep, FP64
 *   5 fine operator, FP64, with the per-workgroup partial sums of p.q (the
 *     form the CG iteration launches)
 * which | 16: before each launch, stream 1 GiB through a scratch buffer so
 * that the kernel finds none of its operands in the memory-side cache (as
 * inside a solve), and time each launch alone.  Without it the launches run
 * back to back on the same buffers, and an operand set smaller than that
 * cache (the FP64 operator's p, q and moduli: 212 MB at 256x128x128) is
 * served from it.
 * Outputs the mean milliseconds per launch and the algorithmic HBM bytes of
 * one launch (true nodes and elements only: vectors read/written once,
 * moduli read once).                                                         */
int top3d_time_kernel(struct SolverDataBuffer *data, int which, int reps,
                      double *avg_ms, double *bytes);

/* Live timing of the fine-level damped-Jacobi sweeps inside the V-cycles of
 * the following solves: HIP events on the library stream around each run of
 * consecutive sweeps (the 3 after the first pre-sweep, the 4 post-sweeps),
 * which execute back to back.  enable = 1 clears the accumulators and starts
 * recording, 0 stops.
 * Outputs (any may be NULL): mean milliseconds per sweep so far, number of
 * sweeps timed, algorithmic bytes of one sweep (as top3d_time_kernel 0 / 1
 * for the mixed / FP64 form).                                               */
int top3d_sweep_timing(struct SolverDataBuffer *data, int enable, double *avg_ms,
                       int *count, double *bytes);

/* The coarsest-level dense solver on its own: A (device, n x n row-major,
 * symmetric positive definite, nonzero only within |i-j| <= bw) is replaced
 * by A^{-1} (banded Cholesky on FP64 MFMA, W = L^{-1}, A^{-1} = W^T W).
 * Returns TOP3D_OK or TOP3D_ERR_NOT_SPD.                                    */
int top3d_dense_spd_inverse(double *A_dev, int n, int bw);

/* FP32 variants of the level operators, the arithmetic of the mixed
 * V-cycle, exposed for parity tests (device float arrays, same layout).    */
void applyStateOperator_stencil_f32(const struct gpuNode *gpu_node,
                                    const DTYPE *x, const float *in,
                                    float *out);
void applyStateOperatorSubspace_halo_f32(const struct gpuNode *gpu_node,
                                         const int l, const DTYPE *x,
                                         const float *in, float *out);

#ifdef __cplusplus
}
#endif
#endif /* TOP3D_HIP_H */
