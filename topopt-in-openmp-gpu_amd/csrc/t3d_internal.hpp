// Internal declarations of the gfx950 MGCG library: level geometry, the
// device-side grid state, the solver state and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/top3d_hip.h"
#include "coarse_schedule.hpp"

namespace t3d {

constexpr int kMaxLevels = 8;
constexpr double kOmega = 0.6;  // stencil_solvers.c:332

// ---------------------------------------------------------------------------
// error state
void set_error(int code, const char *fmt, ...);
bool hip_ok(hipError_t e, const char *what);
#define T3D_HIP(call) ::t3d::hip_ok((call), #call)

hipStream_t stream();
// While a StreamScope lives, stream() returns its stream on this thread.
struct StreamScope {
  explicit StreamScope(hipStream_t s);
  ~StreamScope();
  StreamScope(const StreamScope &) = delete;
  StreamScope &operator=(const StreamScope &) = delete;
  hipStream_t prev;
  bool prev_on;
};

// ---------------------------------------------------------------------------
// Geometry of one multigrid level in the halo layout: node (i,j,k) with
// i in [0,wx), j in [0,wy), k in [0,wz) at i*wy*wz + k*wy + j; real nodes are
// 1..nel+1 per direction, DOF = 3*node + comp.
struct Level {
  int l = 0, ncell = 1;
  int nelx = 0, nely = 0, nelz = 0;
  int nx = 0, ny = 0, nz = 0;
  int wx = 0, wy = 0, wz = 0;
  int64_t ndof = 0;
  int64_t nodes() const { return (int64_t)wx * wy * wz; }
  int64_t interior_nodes() const { return (int64_t)nx * ny * nz; }
  int64_t elements() const { return (int64_t)nelx * nely * nelz; }
};

// Kernel argument bundle for a level plus the fine element layout.
struct LevelArgs {
  int nelx, nely, nelz;  // level elements
  int wy, wz;            // level node strides
  int ncell;             // fine elements per level element per direction
  int few, fez;          // fine element strides: e = i*few*fez + k*few + j
};

// Device-side state of a gridContext (keyed by gc->precomputedKE).
struct DevGrid {
  int nl = 0;
  int nelx = 0, nely = 0, nelz = 0;
  Level lv[kMaxLevels];
  int ewy = 0, ewz = 0;   // fine element strides (wrapy-1, wrapz-1)
  int64_t nelem = 0;      // padded fine element count (wrapx-1)(wrapy-1)(wrapz-1)
  double E0 = 1.0, Emin = 1e-6;
  uint8_t *fixmask[kMaxLevels] = {};   // per halo node: bit c set = DOF fixed
  int *fixplane[kMaxLevels] = {};      // per node plane i: 1 if a node of it has a fixed DOF
  int64_t nfixed[kMaxLevels] = {};
  int64_t *fixed[kMaxLevels] = {};     // fixed DOF indices (device)
  double *KE[2] = {};                  // device KE_0 (576), KE_1 (8*576)
  float *KEf[2] = {};
  double *KE2 = nullptr;               // device KE_2 (64*576), nl >= 3
  double *KE3 = nullptr;               // device KE_3 (512*576), nl >= 4
  double *KEdiag[kMaxLevels] = {};     // diagonals of the level-l sub-cell matrices (24 each)
  // 3x3 blocks of KE_0 in the per-component Walsh-Hadamard basis of the
  // element's nodal values (valid if wh0)
  bool wh0 = false;
  double M0[8][3][3] = {};
  static constexpr int kScratchSlots = 9;
  void *scratch[kScratchSlots] = {};   // lazily allocated work arrays of the API
  size_t scratch_bytes[kScratchSlots] = {};
  LevelArgs args(int l) const;
  void *work(int slot, size_t bytes);  // scratch[slot], grown to >= bytes
};

DevGrid *dev_grid(const struct gridContext *gc);  // nullptr + error if absent
DevGrid *dev_grid_node(const struct gpuNode *node);
void release_dev_grid(const struct gridContext *gc);
void compute_level(const struct gridContext *gc, int l, Level *out);

// present table
void *dptr(const void *p, const char *what);  // mapped host or device pointer
bool is_device_ptr(const void *p);

// device allocation helpers (zero-filled)
void *dalloc(size_t bytes);          // zero-filled on the library stream
void *dalloc_uninit(size_t bytes);   // contents undefined
void dfree(void *p);

// ---------------------------------------------------------------------------
// Kernel launchers.  T is double or float; all run on stream().
template <typename T>
void launch_young(const DevGrid &g, const double *x, T *E);
template <typename T>
void launch_young_float_input(const DevGrid &g, const float *x, T *E);

// fine level: mode 0 out = K in; 1 Jacobi sweep out = in + w invD (b - K in)
// (interior nodes; other nodes copied from in); 2 residual out = b - K in;
// 3 as 1 with invD the inverse diagonal of this operator, which the kernel
// may form from E instead of reading it
template <typename T>
void launch_fine(const DevGrid &g, int mode, const T *E, const T *in, T *out,
                 const T *b, const T *invD, T omega, int bcsym);
template <typename T>
void launch_fine_tile(const DevGrid &g, int mode, const T *E, const T *in, T *out,
                      const T *b, const T *invD, T omega, int bcsym);
// Eh: the transformed moduli of launch_level1_moduli
template <typename T>
void launch_level1_tile(const DevGrid &g, int mode, const T *Eh, const T *in, T *out,
                        const T *b, const T *invD, T omega, int bcsym);
// per level-1 element slot the 8-point transform of its fine moduli (the
// operand of the level-1 tile sweep); level1_moduli_count values
template <typename T>
void launch_level1_moduli(const DevGrid &g, const T *E, T *Eh);
int64_t level1_moduli_count(const DevGrid &g);
// DevGrid::work slot of the transformed moduli formed by launch_level1
// when the caller passes none
constexpr int kSlotLevel1Moduli = 8;
bool tile_level1_available();  // the level-1 element matrices have the mirror structure
bool tile_fine_diag_available();  // the fine tile sweep can form its diagonal from E
// launch_fine mode 3 forms the diagonal from E (no invD read)
bool fine_sweep_forms_diagonal();
// zero the padded y rows (j > nely+1) of a level-0 vector
void launch_zero_pad_rows(const DevGrid &g, double *out);
// level-1 matrix-free operator with fine moduli E, same modes; Eh: their
// transform by launch_level1_moduli (formed here into a DevGrid work slot
// when null)
template <typename T>
void launch_level1(const DevGrid &g, int mode, const T *E, const T *in, T *out,
                   const T *b, const T *invD, T omega, int bcsym, const T *Eh = nullptr);
// assembled 27-point node stencil (levels >= 2), same modes
template <typename T>
void launch_stencil(const DevGrid &g, int l, int mode, const T *S, const T *in,
                    T *out, const T *b, const T *invD, T omega);
// the same on the interior nodes only (one thread per node and row); the
// other nodes of `out` are left as they are
template <typename T>
void launch_stencil_rows(const DevGrid &g, int l, int mode, const T *S, const T *in,
                         T *out, const T *b, const T *invD, T omega);
// fine-level sweep from z = 0 with the diagonal formed from E; false if the
// element matrix does not allow it
template <typename TB, typename T>
bool launch_fine_jacobi_zero(const DevGrid &g, const TB *b, const T *E, T omega, T *out);
// first pre-smoothing sweep from z = 0:  out = w invD b on interior, else 0
template <typename TB, typename T>
void launch_jacobi_zero(const DevGrid &g, int l, const TB *b, const T *invD,
                        T omega, T *out);
template <typename T>
void launch_restrict(const DevGrid &g, int l, const T *fine, T *coarse,
                     const T *invDc = nullptr, T omega = T(0), T *zc = nullptr);
template <typename T>
void launch_prolong_add(const DevGrid &g, int l, const T *coarse, T *fine);
template <typename T>
void launch_prolong(const DevGrid &g, int l, const T *coarse, T *fine);
// inverse diagonal of level l (0, 1: from moduli; >= 2: from a stencil)
template <typename T>
void launch_invdiag_mf(const DevGrid &g, int l, const double *E, T *invD);
template <typename T>
void launch_invdiag_stencil(const DevGrid &g, int l, const double *S, T *invD);

// coarse element matrices and stencils (FP64)
void launch_elem_level2(const DevGrid &g, const double *E, double *Ke2);
// Level-3 element matrices Ke3[e] = sum over the element's 512 fine elements
// f of E_f KE_3[f], as a library FP64 GEMM on the moduli gathered into Eg
// (level-3 elements x 512); false if it cannot run (no KE_3, library error).
bool launch_elem_level3(const DevGrid &g, const double *E, double *Eg, double *Ke3);
void launch_elem_galerkin(const DevGrid &g, int l, const double *Kchild,
                          double *Kparent);
// S32: FP32 copy of S written alongside (optional).  half: the symmetric
// half layout (kHalfPlanes planes, symmetric elimination of fixed DOFs):
// neighbours nb = 14..26 at planes (nb-14)*9 + 3r + c, the centre's entries
// r <= c at planes 117 + half_centre_index(r, c)
constexpr int kHalfPlanes = 123;
__host__ __device__ constexpr int half_centre_index(int r, int c) {
  return r == 0 ? c : (r == 1 ? 2 + c : 5);  // (0,0..2) -> 0..2, (1,1..2) -> 3..4, (2,2) -> 5
}
void launch_node_stencil(const DevGrid &g, int l, const double *Ke, double *S,
                         int bc_symmetric, float *S32 = nullptr, bool half = false);
// smoother sweep with a half-layout stencil (interior nodes only, as
// launch_stencil_rows)
template <typename T>
void launch_stencil_half(const DevGrid &g, int l, int mode, const T *Sh, const T *in, T *out,
                         const T *b, const T *invD, T omega);
// inverse diagonal from a half-layout stencil
template <typename T>
void launch_invdiag_half(const DevGrid &g, int l, const double *Sh, T *invD);
struct CoarseFactor;
// assemble level l's stencil into cf.A (storage order, after begin_assembly)
void launch_stencil_to_dense(const DevGrid &g, int l, const double *S, CoarseFactor &cf);

// dense FP64 coarse solver: P A P^T = L L^T over 64x64 tiles in an
// elimination order (nested dissection over the grid's x-planes, or banded),
// W = L^{-1}, WT = W^T, so that A^{-1} = P^T WT W P; FP32 copies of W and WT
// for the mixed V-cycle.  Storage is np x np (np = 64 nblk), the permuted
// DOFs plus identity padding.
struct CoarseFactor {
  int n = 0, bw = 0, np = 0, nblk = 0;
  double *A = nullptr;     // np x np matrix in storage order (overwritten)
  double *L = nullptr;     // np x np Cholesky factor
  double *Dinv = nullptr;  // nblk 64x64 inverses of L's diagonal blocks
  double *W = nullptr, *WT = nullptr;
  float *W32 = nullptr, *WT32 = nullptr;
  int *info = nullptr;
  double *cw64 = nullptr;  // 2 np work (gathered input, intermediate)
  float *cw32 = nullptr;
  int *pos = nullptr;      // logical DOF -> storage index (device, n)
  int *ipos = nullptr;     // storage index -> logical DOF or -1 (device, np)
  // schedule: launch q runs tasks [launch_first[q], +launch_count[q]);
  // W rows per tile column (colinfo: offset, count into rows; rows: I,
  // offset, count into klist)
  CholTask *dtasks = nullptr;
  int ntask = 0;
  std::vector<int> launch_first, launch_count;
  int2 *dcolinfo = nullptr;
  int4 *drows = nullptr;
  int *dklist = nullptr;
  // the tiles I * nblk + J the factorization reads (targets and sources of
  // its tasks): the only part of A that assembly has to clear
  int *dztiles = nullptr;
  int nztiles = 0;
  // planes > 1 with n divisible by planes: nested dissection over the planes
  // (n / planes consecutive DOFs each, coupled to adjacent planes only);
  // otherwise banded with half bandwidth bw
  bool allocate(int n, int bw, bool f32, int planes);
  void release();
  void begin_assembly();  // A = 0 on the tiles read by the factorization, identity padding
  // TOP3D_OK or TOP3D_ERR_NOT_SPD / TOP3D_ERR_HIP; f32out: W32 and WT32 for
  // the FP32 apply instead of the FP64 WT
  int factor(bool f32out = false);
  void factor_async(bool f32out);  // the same work, queued without waiting
  int check_info();                // after factor_async has completed
  template <typename T> const T *w() const {
    if constexpr (sizeof(T) == 8) return (const T *)W; else return (const T *)W32;
  }
  template <typename T> const T *wt() const {
    if constexpr (sizeof(T) == 8) return (const T *)WT; else return (const T *)WT32;
  }
  template <typename T> T *cw() const {
    if constexpr (sizeof(T) == 8) return (T *)cw64; else return (T *)cw32;
  }
};
// out (level-l halo layout) = A^{-1} in (interior DOFs)
template <typename T>
void launch_coarse_apply(const DevGrid &g, int l, const CoarseFactor &cf, const T *in, T *out);

// BLAS-1 (FP64 accumulation in the reductions)
// *result = a.b; also into the pinned host word `host` when given
template <typename T>
void launch_dot(const T *a, const T *b, int64_t n, double *partials, double *result,
                double *host = nullptr);
template <typename TO, typename TI>
void launch_convert(int64_t n, const TI *in, TO *out);
template <typename T>
void launch_fill(int64_t n, T *out, T v);
template <typename T>
void launch_sub(int64_t n, const T *a, const T *b, T *out);  // out = a - b
template <typename T>
void launch_add(int64_t n, const T *a, T *out);               // out += a
// fused CG steps (partial sums consumed by the next kernel, kRedGrid = 1024
// partials for launch_dot_rz and the r.r partials)
// (r and z in the same precision: the FP32 copy of r in the mixed form)
template <typename TZ>
void launch_dot_rz(const TZ *r, const TZ *z, int64_t n, double *partials);
// p update; from the second iteration on also the previous u += alpha p
template <typename TZ>
void launch_pupdate_fused(int64_t n, const double *partials, int npart, double *scal, double *p,
                          const TZ *z, int first, double *u);
// u += alpha p of the last CG update (alpha from scal)
void launch_u_update(int64_t n, const double *scal, double *u, const double *p);
// r -= alpha q; r.r goes to *rr and, when given, to the pinned host word host_rr
void launch_axpy_fused(int64_t n, const double *pq_part, int npart, double *scal, double *r,
                       const double *q, float *r32, double *rr_part, double *rr,
                       double *host_rr);
// FP64 fine apply q = K p with per-workgroup partials of p.q; returns their
// count (0 if `cap` is too small)
// fine Jacobi sweep (diagonal from the moduli) with per-workgroup partial
// sums of b . out; returns their count (0: none written)
template <typename T>
int launch_fine_tile_sweep_dot(const DevGrid &g, const T *E, const T *in, T *out, const T *b,
                               const T *invD, T omega, int bcsym, double *partials, int cap);
template <typename T>
int launch_fine_sweep_dot(const DevGrid &g, const T *E, const T *in, T *out, const T *b,
                          const T *invD, T omega, int bcsym, double *partials, int cap);
// fine inverse diagonal from the moduli (false: the element matrix lacks the
// per-component diagonal this needs)
bool launch_fine_invdiag(const DevGrid &g, const double *E, double *invD);
int launch_fine_apply_dot(const DevGrid &g, const double *E, const double *p, double *q,
                          int bcsym, double *partials, int cap);
void launch_copy_fixed(const DevGrid &g, int l, const double *in, double *out);

// design-loop kernels
void launch_sensitivity(const DevGrid &g, const double *x, const double *u,
                        double *dc, double *partials, double *c);
void launch_filter(const DevGrid &g, double rmin, const double *rho,
                   double *out);
void launch_filter_grad(const DevGrid &g, double rmin, const double *v,
                        double *tmp, double *out);
void launch_volume(const DevGrid &g, const double *x, double *partials,
                   double *result);
void launch_change(int64_t n, double *x, const double *xnew, double *partials,
                   double *result);
// scratch: kOcScratchDoubles doubles of device memory.
constexpr int kOcScratchDoubles = 8 + 15 * 1024;
int oc_update(const double *x, double *xnew, const double *dv,
              const double *dc, int64_t nelem, double gval, double *scratch);

// ---------------------------------------------------------------------------
// Host-side KE generation (same quadrature results as local_matrix.c).
void element_matrices(double nu, int l, double *KE);
// M[s][c][d]: the 3x3 blocks of T KE T^T / 64 (T: per-component 8-point
// Walsh-Hadamard transform); false if KE lacks the reflection structure.
bool hadamard_blocks(const double *KE, double M[8][3][3]);

}  // namespace t3d
