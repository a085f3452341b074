// Device-side state of one SolverDataBuffer (kept in its bottomSolver slot).
#pragma once

#include "t3d_internal.hpp"

namespace t3d {

struct SolverState {
  static constexpr uint32_t kMagic = 0x74336453u;
  enum { kR = 0, kZ = 1, kD = 2 };
  uint32_t magic = kMagic;
  DevGrid *g = nullptr;
  int nl = 0;
  int nswp = 4;
  int mixed = 0;
  int bcsym = 0;
  int max_iter = 1000;
  // outer CG vectors (FP64, level-0 halo layout)
  double *r = nullptr, *p = nullptr, *q = nullptr, *z = nullptr;
  // Young's moduli of the fine elements
  double *E64 = nullptr;
  float *E32 = nullptr;
  // their per-level-1-element transform for the level-1 sweeps, in the
  // precision of the last set-up's V-cycle (launch_level1_moduli; sized for
  // FP64)
  void *Eh1 = nullptr;
  bool eh1_f32 = false, eh1_set = false;
  template <typename T> const T *eh1_t() const {
    return (eh1_set && eh1_f32 == (sizeof(T) == 4)) ? (const T *)Eh1 : nullptr;
  }
  // per level: V-cycle vectors r, z, d (FP64 and FP32), ping-pong buffer
  double *v64[kMaxLevels][3] = {};
  float *v32[kMaxLevels][3] = {};
  double *t64[kMaxLevels] = {};
  float *t32[kMaxLevels] = {};
  double *invD64[kMaxLevels] = {};
  float *invD32[kMaxLevels] = {};
  // assembled levels: element matrices and 27-point stencils
  double *Ke[kMaxLevels] = {};
  double *Eg3 = nullptr;  // moduli gathered per level-3 element (library GEMM operand)
  double *S64[kMaxLevels] = {};
  float *S32[kMaxLevels] = {};
  // coarsest level: Cholesky factor and W = L^{-1}
  int ncoarse = 0, bandwidth = 0;
  CoarseFactor coarse;
  // device scalars, reduction partials, pinned host copy
  static constexpr int kPartRZ = 1024, kPartPQ = 1 << 16, kPartRR = 1024;
  double *scal = nullptr, *partials = nullptr, *hscal = nullptr;
  double *cgpart = nullptr;  // partial sums of r.z, p.q, r.r
  double *hscal_dev = nullptr;  // device address of hscal
  int rz_parts = 0;  // r.z partials written by the last preconditioner application
  hipEvent_t ev = nullptr;
  // The assembled levels and the coarse factorization are built on `aux`,
  // concurrently with the start of the solve on the library stream, which
  // waits for ev_stencil before its first sweep on level 2 and for ev_coarse
  // before its first coarse solve.
  // The element matrices of the assembled levels and the coarse factor run
  // on `aux`; the smoother levels' stencils run on `aux2` as their element
  // matrices become ready (ev_ke[l]), off the coarse factor's path.
  hipStream_t aux = nullptr, aux2 = nullptr;
  hipEvent_t ev_base = nullptr, ev_stencil = nullptr, ev_coarse = nullptr, ev_done = nullptr;
  hipEvent_t ev_done2 = nullptr;
  hipEvent_t ev_ke[kMaxLevels] = {};
  bool wait_stencil = false, wait_coarse = false;
  void join_stencil();
  void join_coarse();
  Top3dSolveStats stats{};
  // live timing of the fine sweeps (top3d_sweep_timing): one event pair
  // around each run of consecutive sweeps, sw_n[k] sweeps in pair k
  bool sweep_timing = false;
  std::vector<hipEvent_t> sw_ev;
  std::vector<int> sw_n;
  size_t sw_used = 0;
  double sw_ms = 0.0;
  int sw_count = 0;
  bool sweep_timing_begin();
  void sweep_timing_end(bool rec, int nsweeps);
  void collect_sweep_timing();

  bool allocate(DevGrid *grid, int nlev);
  void release();
  bool setup_main(const double *x);
  bool setup_aux();
  void precondition();
  int solve(const double *x, const double *b, double *u, double tol, int *final_iter,
            float *final_res);

  template <typename T> void op(int l, int mode, const T *in, T *out, const T *b);
  // zero_done: the first sweep from zero of level l is already in its buffer
  template <typename T> void vcycle(int l, bool zero_done = false);
  template <typename T> T *smooth_from_zero(int l, const T *b, T *zb, T *tb,
                                            bool zero_done = false);
  template <typename T> T *smooth(int l, const T *b, T *cur, T *oth);

  template <typename T> T *vec(int l, int which) {
    if constexpr (sizeof(T) == 8) return (T *)v64[l][which]; else return (T *)v32[l][which];
  }
  template <typename T> T *tmp(int l) {
    if constexpr (sizeof(T) == 8) return (T *)t64[l]; else return (T *)t32[l];
  }
  template <typename T> const T *invd(int l) {
    if constexpr (sizeof(T) == 8) return (const T *)invD64[l]; else return (const T *)invD32[l];
  }
  template <typename T> const T *young_t() {
    if constexpr (sizeof(T) == 8) return (const T *)E64; else return (const T *)E32;
  }
  template <typename T> const T *stencil_t(int l) {
    if constexpr (sizeof(T) == 8) return (const T *)S64[l]; else return (const T *)S32[l];
  }
};

SolverState *solver_state(const struct SolverDataBuffer *data);

}  // namespace t3d
