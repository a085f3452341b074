// MGCG state solve on one GPU: per-solve set-up of every level, the
// V-cycle preconditioner and the preconditioned CG recurrence
// (solveStateMG_halo, stencil_solvers.c:129-397).
//
// Levels: 0 fine matrix-free, 1 matrix-free with the 8 sub-cell matrices,
// 2..nl-2 assembled 27-point stencils (symmetric BC elimination), nl-1 the
// FP64 Cholesky-factored coarsest matrix, applied as W^T W with W = L^{-1}.  Nothing
// leaves the device inside the CG loop except the residual norm read each
// iteration for the stopping test.
#include <chrono>
#include <cmath>
#include <cstring>

#include "solver_state.hpp"

namespace t3d {

template <typename T>
static T *alloc_t(int64_t n) {
  return (T *)dalloc(sizeof(T) * (size_t)n);
}

SolverState *solver_state(const struct SolverDataBuffer *data) {
  if (!data) {
    set_error(TOP3D_ERR_ARG, "null SolverDataBuffer");
    return nullptr;
  }
  SolverState *s = (SolverState *)data->bottomSolver.cholmodCommon;
  if (!s || s->magic != SolverState::kMagic) {
    set_error(TOP3D_ERR_ARG, "SolverDataBuffer was not allocated by this library");
    return nullptr;
  }
  return s;
}

// half bandwidth of the coarsest matrix in the x-slowest DOF order
static int coarse_bandwidth(const Level &L) {
  return 3 * (L.ny * L.nz + L.ny + 1) + 2;
}

bool SolverState::allocate(DevGrid *grid, int nlev) {
  g = grid;
  nl = nlev;
  const int64_t n0 = g->lv[0].ndof;
  r = alloc_t<double>(n0);
  p = alloc_t<double>(n0);
  q = alloc_t<double>(n0);
  z = alloc_t<double>(n0);
  E64 = alloc_t<double>(g->nelem);
  E32 = alloc_t<float>(g->nelem);
  if (nl >= 2) Eh1 = alloc_t<double>(level1_moduli_count(*g));
  for (int l = 0; l < nl; l++) {
    const int64_t nd = g->lv[l].ndof;
    invD64[l] = alloc_t<double>(nd);
    invD32[l] = alloc_t<float>(nd);
    for (int b = 0; b < 3; b++) {
      v64[l][b] = (l == 0 && b == 0) ? nullptr : alloc_t<double>(nd);
      v32[l][b] = alloc_t<float>(nd);
    }
    t64[l] = alloc_t<double>(nd);
    t32[l] = alloc_t<float>(nd);
    if (l >= 2) {
      Ke[l] = alloc_t<double>(576 * g->lv[l].elements());
      const int64_t sn = 243 * g->lv[l].interior_nodes();
      S64[l] = alloc_t<double>(sn);
      S32[l] = alloc_t<float>(sn);
    }
  }
  // level-0 V-cycle vectors in FP64 alias the CG vectors r (input) and z
  v64[0][kR] = r;
  dfree(v64[0][kZ]);
  v64[0][kZ] = z;
  const Level &C = g->lv[nl - 1];
  if (nl >= 4) Eg3 = alloc_t<double>(g->lv[3].elements() * 512);
  ncoarse = (int)(3 * C.interior_nodes());
  bandwidth = coarse_bandwidth(C);
  if (!coarse.allocate(ncoarse, bandwidth, true, C.nx)) return false;
  scal = alloc_t<double>(16);
  partials = alloc_t<double>(1024 + 16);
  cgpart = alloc_t<double>(kPartRZ + kPartPQ + kPartRR);
  if (!T3D_HIP(hipHostMalloc((void **)&hscal, sizeof(double) * 16, hipHostMallocDefault)) ||
      !T3D_HIP(hipHostGetDevicePointer((void **)&hscal_dev, hscal, 0)))
    return false;
  if (!T3D_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming))) return false;
  if (!T3D_HIP(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking)) ||
      !T3D_HIP(hipStreamCreateWithFlags(&aux2, hipStreamNonBlocking)) ||
      !T3D_HIP(hipEventCreateWithFlags(&ev_base, hipEventDisableTiming)) ||
      !T3D_HIP(hipEventCreateWithFlags(&ev_stencil, hipEventDisableTiming)) ||
      !T3D_HIP(hipEventCreateWithFlags(&ev_coarse, hipEventDisableTiming)) ||
      !T3D_HIP(hipEventCreateWithFlags(&ev_done, hipEventDisableTiming)) ||
      !T3D_HIP(hipEventCreateWithFlags(&ev_done2, hipEventDisableTiming)))
    return false;
  for (int l = 2; l < nl - 1; l++)
    if (!T3D_HIP(hipEventCreateWithFlags(&ev_ke[l], hipEventDisableTiming))) return false;
  return T3D_HIP(hipStreamSynchronize(stream())) && r && p && q && z;
}

void SolverState::release() {
  hipStreamSynchronize(stream());
  if (aux) hipStreamSynchronize(aux);
  if (aux2) hipStreamSynchronize(aux2);
  auto f = [](void *ptr) { dfree(ptr); };
  for (int l = 0; l < nl; l++) {
    f(invD64[l]);
    f(invD32[l]);
    for (int b = 0; b < 3; b++) {
      if (!(l == 0 && (b == kR || b == kZ))) f(v64[l][b]);
      f(v32[l][b]);
    }
    f(t64[l]);
    f(t32[l]);
    f(Ke[l]);
    f(S64[l]);
    f(S32[l]);
  }
  f(r); f(p); f(q); f(z); f(E64); f(E32); f(Eh1); f(Eg3);
  coarse.release();
  f(scal); f(partials); f(cgpart);
  if (hscal) hipHostFree(hscal);
  if (ev) hipEventDestroy(ev);
  for (hipEvent_t e : {ev_base, ev_stencil, ev_coarse, ev_done, ev_done2})
    if (e) hipEventDestroy(e);
  for (hipEvent_t &e : ev_ke) {
    if (e) hipEventDestroy(e);
    e = nullptr;
  }
  if (aux) hipStreamDestroy(aux);
  if (aux2) hipStreamDestroy(aux2);
  aux = aux2 = nullptr;
  for (hipEvent_t e : sw_ev) hipEventDestroy(e);
  sw_ev.clear();
  magic = 0;
}

// ---------------------------------------------------------------------------
// per-solve set-up (stencil_solvers.c:283-302): the fine levels on the
// library stream
bool SolverState::setup_main(const double *x) {
  launch_young<double>(*g, x, E64);
  // the second stream needs only E64
  if (!T3D_HIP(hipEventRecord(ev_base, stream()))) return false;
  if (mixed) launch_convert<float, double>(g->nelem, E64, E32);
  eh1_set = false;
  if (nl >= 2 && Eh1 && tile_level1_available()) {
    if (mixed) launch_level1_moduli<float>(*g, E32, (float *)Eh1);
    else launch_level1_moduli<double>(*g, E64, (double *)Eh1);
    eh1_f32 = mixed != 0;
    eh1_set = true;
  }
  // the fine Jacobi sweeps form their diagonal from E when they can
  const bool fdiag = fine_sweep_forms_diagonal();
  if (!fdiag) launch_invdiag_mf<double>(*g, 0, E64, invD64[0]);
  launch_invdiag_mf<double>(*g, 1, E64, invD64[1]);
  if (mixed)
    for (int l = fdiag ? 1 : 0; l < 2; l++)
      launch_convert<float, double>(g->lv[l].ndof, invD64[l], invD32[l]);
  return true;
}

// The assembled levels and the coarse factorization after the moduli
// (ev_base).  On `aux`: the element matrices level by level from level 3,
// whose matrices come straight from the fine moduli (one library GEMM), so
// that the coarse chain does not wait for level 2, then the coarsest level's
// stencil, dense matrix and factor.  On `aux2`: the level-2 element matrices
// and the smoother levels' stencils and inverse diagonals, each after its
// level's element matrices (ev_ke[l]).  Without the level-3 GEMM (fewer than
// 4 levels, or the library unavailable) `aux` starts from level 2 and builds
// level 3 by the Galerkin product.
bool SolverState::setup_aux() {
  bool direct3 = false;
  {
    StreamScope on_aux(aux);
    if (!T3D_HIP(hipStreamWaitEvent(aux, ev_base, 0))) return false;
    if (nl >= 4) direct3 = launch_elem_level3(*g, E64, Eg3, Ke[3]);
    for (int l = direct3 ? 3 : 2; l < nl; l++) {
      if (l == 2)
        launch_elem_level2(*g, E64, Ke[2]);
      else if (!(direct3 && l == 3))
        launch_elem_galerkin(*g, l, Ke[l - 1], Ke[l]);
      if (l < nl - 1) {
        if (!T3D_HIP(hipEventRecord(ev_ke[l], aux))) return false;
      } else {
        // the coarsest level: the full stencil for its dense assembly
        launch_node_stencil(*g, l, Ke[l], S64[l], 1, nullptr, false);
        launch_invdiag_stencil<double>(*g, l, S64[l], invD64[l]);
      }
    }
    launch_stencil_to_dense(*g, nl - 1, S64[nl - 1], coarse);
    coarse.factor_async(mixed);
    if (!T3D_HIP(hipEventRecord(ev_coarse, aux))) return false;
    // the fine inverse diagonal is not read by the fine sweeps when they form
    // it from E; it is still filled in for callers of the SolverDataBuffer
    if (fine_sweep_forms_diagonal() && !launch_fine_invdiag(*g, E64, invD64[0]))
      launch_invdiag_mf<double>(*g, 0, E64, invD64[0]);
    if (!T3D_HIP(hipEventRecord(ev_done, aux))) return false;
  }
  {
    StreamScope on_aux2(aux2);
    if (direct3) {
      if (!T3D_HIP(hipStreamWaitEvent(aux2, ev_base, 0))) return false;
      launch_elem_level2(*g, E64, Ke[2]);
    }
    // smoother levels: the symmetric half layout (and its FP32 copy for the
    // mixed V-cycle)
    for (int l = 2; l < nl - 1; l++) {
      if (!(direct3 && l == 2) && !T3D_HIP(hipStreamWaitEvent(aux2, ev_ke[l], 0))) return false;
      launch_node_stencil(*g, l, Ke[l], S64[l], 1, mixed ? S32[l] : nullptr, true);
      launch_invdiag_half<double>(*g, l, S64[l], invD64[l]);
      if (mixed) launch_convert<float, double>(g->lv[l].ndof, invD64[l], invD32[l]);
    }
    if (!T3D_HIP(hipEventRecord(ev_stencil, aux2))) return false;
    if (!T3D_HIP(hipEventRecord(ev_done2, aux2))) return false;
  }
  wait_stencil = wait_coarse = true;
  return true;
}

// the library stream waits for the assembled smoother levels / the coarse
// factor built on `aux` (once per solve)
void SolverState::join_stencil() {
  if (!wait_stencil) return;
  hipStreamWaitEvent(stream(), ev_stencil, 0);
  wait_stencil = false;
}
void SolverState::join_coarse() {
  if (!wait_coarse) return;
  join_stencil();
  hipStreamWaitEvent(stream(), ev_coarse, 0);
  wait_coarse = false;
}

// ---------------------------------------------------------------------------
// level operator dispatch
template <typename T>
void SolverState::op(int l, int mode, const T *in, T *out, const T *b) {
  const T omega = (T)kOmega;
  const T *invD = invd<T>(l);
  if (l == 0)  // Jacobi sweeps form the diagonal from the moduli
    launch_fine<T>(*g, mode == 1 ? 3 : mode, young_t<T>(), in, out, b, invD, omega, bcsym);
  else if (l == 1)
    launch_level1<T>(*g, mode, young_t<T>(), in, out, b, invD, omega, bcsym, eh1_t<T>());
  else
    launch_stencil_half<T>(*g, l, mode, stencil_t<T>(l), in, out, b, invD, omega);
}

// Live sweep timing: an event pair brackets each run of consecutive fine
// sweeps, which the stream executes back to back.  (A pair around every
// sweep delays the next launch by 4-8 us per record.)
bool SolverState::sweep_timing_begin() {
  if (!sweep_timing) return false;
  while (sw_ev.size() < sw_used + 2) {
    hipEvent_t e = nullptr;
    // timing only: no system-scope fence (cache write-back) at the record
    if (!T3D_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence))) return false;
    sw_ev.push_back(e);
  }
  return hipEventRecord(sw_ev[sw_used], stream()) == hipSuccess;
}

void SolverState::sweep_timing_end(bool rec, int nsweeps) {
  if (!rec) return;
  if (hipEventRecord(sw_ev[sw_used + 1], stream()) != hipSuccess) return;
  if (sw_n.size() < sw_used / 2 + 1) sw_n.resize(sw_used / 2 + 1);
  sw_n[sw_used / 2] = nsweeps;
  sw_used += 2;
}

// nswp damped-Jacobi sweeps on level l starting from z = 0 (first sweep
// without an operator application); ping-pong between z and t.  Returns the
// buffer holding the result.
template <typename T>
T *SolverState::smooth_from_zero(int l, const T *b, T *zb, T *tb, bool zero_done) {
  T *cur = (nswp % 2 == 1) ? zb : tb;
  T *oth = (cur == zb) ? tb : zb;
  if (zero_done)
    ;  // written by the restriction that produced b
  else if (l == 0 && fine_sweep_forms_diagonal())
    launch_fine_jacobi_zero<T, T>(*g, b, young_t<T>(), (T)kOmega, cur);
  else
    launch_jacobi_zero<T, T>(*g, l, b, invd<T>(l), (T)kOmega, cur);
  const bool rec = l == 0 && nswp > 1 && sweep_timing_begin();
  for (int s = 1; s < nswp; s++) {
    op<T>(l, 1, cur, oth, b);
    std::swap(cur, oth);
  }
  sweep_timing_end(rec, nswp - 1);
  return cur;  // == zb
}

// The last fine sweep of a preconditioner application also writes the
// partial sums of r . z for the CG (rz_parts of them, 0 if it could not).
template <typename T>
T *SolverState::smooth(int l, const T *b, T *cur, T *oth) {
  const bool rec = l == 0 && nswp > 0 && sweep_timing_begin();
  for (int s = 0; s < nswp; s++) {
    if (l == 0 && s == nswp - 1)
      rz_parts = launch_fine_sweep_dot<T>(*g, young_t<T>(), cur, oth, b, invd<T>(0), (T)kOmega,
                                          bcsym, cgpart, kPartRZ);
    else
      op<T>(l, 1, cur, oth, b);
    std::swap(cur, oth);
  }
  sweep_timing_end(rec, nswp);
  return cur;
}

// V-cycle at level l: reads vec<T>(l, kR), leaves the result in vec<T>(l, kZ)
// (stencil_solvers.c:129-252).
template <typename T>
void SolverState::vcycle(int l, bool zero_done) {
  T *r_ = vec<T>(l, kR), *z_ = vec<T>(l, kZ), *d_ = vec<T>(l, kD), *t_ = tmp<T>(l);
  if (l >= 2) join_stencil();
  T *zc = smooth_from_zero<T>(l, r_, z_, t_, zero_done);
  T *other = (zc == z_) ? t_ : z_;
  op<T>(l, 2, zc, d_, r_);  // d = r - K z
  if (nl == l + 2) {
    launch_restrict<T>(*g, l, d_, vec<T>(l + 1, kR));
    join_coarse();
    launch_coarse_apply<T>(*g, l + 1, coarse, vec<T>(l + 1, kR), vec<T>(l + 1, kZ));
  } else {
    // the restriction also writes level l+1's first sweep from zero into the
    // buffer smooth_from_zero starts in; level l+1's inverse diagonal is
    // ready (library stream for l+1 = 1, the stencil event for l+1 >= 2)
    if (l + 1 >= 2) join_stencil();
    T *first = (nswp % 2 == 1) ? vec<T>(l + 1, kZ) : tmp<T>(l + 1);
    launch_restrict<T>(*g, l, d_, vec<T>(l + 1, kR), invd<T>(l + 1), (T)kOmega, first);
    vcycle<T>(l + 1, true);
  }
  launch_prolong_add<T>(*g, l, vec<T>(l + 1, kZ), zc);  // z += P zc
  T *res = smooth<T>(l, r_, zc, other);
  if (res != z_)
    hipMemcpyAsync(z_, res, sizeof(T) * g->lv[l].ndof, hipMemcpyDeviceToDevice, stream());
}

// z = M r.  Mixed form: the V-cycle reads the FP32 copy of r written by the
// CG update and leaves z in FP32, which the CG kernels read directly.
void SolverState::precondition() {
  rz_parts = 0;
  if (!mixed) {
    vcycle<double>(0);  // r -> z (aliases)
    return;
  }
  vcycle<float>(0);
}

void SolverState::collect_sweep_timing() {
  for (size_t k = 0; k + 1 < sw_used; k += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, sw_ev[k], sw_ev[k + 1]) == hipSuccess) {
      sw_ms += ms;
      sw_count += sw_n[k / 2];
    }
  }
  sw_used = 0;
}

// ---------------------------------------------------------------------------
// PCG (stencil_solvers.c:259-397) with scalars on the device:
// scal[0] rho, [1] rho_old, [2] p.q, [3] r.r, [4] b.b
int SolverState::solve(const double *x, const double *b, double *u, double tol, int *final_iter,
                       float *final_res) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  stats = Top3dSolveStats{};
  const int64_t n0 = g->lv[0].ndof;
  hipMemcpyAsync(z, u, sizeof(double) * n0, hipMemcpyDeviceToDevice, stream());
  if (!setup_main(x)) {
    stats.status = top3d_last_error();
    return stats.status;
  }
  launch_fine<double>(*g, 2, E64, z, r, b, nullptr, 0.0, bcsym);  // r = b - K z
  hscal[4] = 0.0;
  launch_dot<double>(b, b, n0, partials, scal + 4, hscal_dev + 4);  // b.b, also into hscal[4]
  if (mixed) launch_convert<float, double>(n0, r, v32[0][kR]);
  // the coarse set-up is queued while the library stream computes r
  if (!setup_aux()) {
    hipStreamSynchronize(aux);
    hipStreamSynchronize(aux2);
    hipStreamSynchronize(stream());
    wait_stencil = wait_coarse = false;
    stats.status = top3d_last_error();
    return stats.status;
  }
  const auto t1 = clk::now();
  stats.setup_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  double bnorm = 0.0;  // read once the first iteration's event has completed
  int status = TOP3D_OK;
  int iter = 0;
  double relres = 0.0;
  bool pending = false;  // a V-cycle for iteration `iter` is already queued
  // relative residuals of the two previous iterations (< 0: not yet known)
  double res1 = -1.0, res2 = -1.0;
  double *part_rz = cgpart, *part_pq = cgpart + kPartRZ, *part_rr = part_pq + kPartPQ;
  for (iter = 0; iter < max_iter; iter++) {
    if (!pending) precondition();
    pending = false;
    // rho = r.z (partials from the V-cycle's last fine sweep, else a dot
    // kernel); p = z + (rho / rho_old) p
    int nrz = rz_parts;
    if (mixed) {
      if (nrz == 0) launch_dot_rz<float>(v32[0][kR], v32[0][kZ], n0, part_rz);
      launch_pupdate_fused<float>(n0, part_rz, nrz ? nrz : kPartRZ, scal, p, v32[0][kZ],
                                  iter == 0, u);
    } else {
      if (nrz == 0) launch_dot_rz<double>(r, z, n0, part_rz);
      launch_pupdate_fused<double>(n0, part_rz, nrz ? nrz : kPartRZ, scal, p, z, iter == 0, u);
    }
    // q = K p with the partial sums of p.q
    int npq = launch_fine_apply_dot(*g, E64, p, q, bcsym, part_pq, kPartPQ);
    if (npq == 0) {
      launch_dot_rz<double>(p, q, n0, part_pq);
      npq = kPartRZ;
    }
    // r -= alpha q (and its FP32 copy), r.r (also into hscal[3]); u += alpha p
    // follows with the next p update
    launch_axpy_fused(n0, part_pq, npq, scal, r, q, mixed ? v32[0][kR] : nullptr, part_rr,
                      scal + 3, hscal_dev + 3);
    hipEventRecord(ev, stream());
    // queue the next V-cycle while the residual norm travels to the host
    // (if this iteration converges, its result is not used), unless the last
    // contraction predicts convergence: then the host waits for the norm
    // (a few tens of microseconds with the device idle when the prediction
    // is wrong) instead of a whole V-cycle running after the last iteration
    bool queue_next = iter + 1 < max_iter;
    if (queue_next && res1 > 0.0 && res2 > 0.0 && res1 * (res1 / res2) < 2.0 * tol)
      queue_next = false;
    if (queue_next) {
      precondition();
      pending = true;
    }
    if (!T3D_HIP(hipEventSynchronize(ev))) {
      status = TOP3D_ERR_HIP;
      break;
    }
    if (iter == 0) bnorm = std::sqrt(hscal[4]);
    relres = std::sqrt(hscal[3]) / bnorm;
    res2 = res1;
    res1 = relres;
    if (final_iter) *final_iter = iter;
    if (final_res) *final_res = (float)relres;
    if (!std::isfinite(relres)) {
      status = TOP3D_ERR_NONFINITE;
      set_error(status, "CG: non-finite residual at iteration %d", iter);
      break;
    }
    if (relres < tol) break;
  }
  if (max_iter > 0) launch_u_update(n0, scal, u, p);  // the last iteration's u += alpha p
  join_coarse();  // the work on `aux` and `aux2` has finished before the solve returns
  hipStreamWaitEvent(stream(), ev_done, 0);
  hipStreamWaitEvent(stream(), ev_done2, 0);
  if (!T3D_HIP(hipStreamSynchronize(stream()))) status = TOP3D_ERR_HIP;
  if (status == TOP3D_OK) {
    const int rc = coarse.check_info();
    if (rc != TOP3D_OK) status = rc;
  }
  if (status == TOP3D_OK && iter >= max_iter) {
    status = TOP3D_ERR_NO_CONVERGE;
    set_error(status, "CG: relres %.3e above tol after %d iterations", relres, max_iter);
    iter = max_iter - 1;
  }
  stats.iterations = iter + 1;
  stats.relres = relres;
  stats.status = status;
  stats.solve_ms = std::chrono::duration<double, std::milli>(clk::now() - t1).count();
  if (sweep_timing) collect_sweep_timing();
  return status;
}

template void SolverState::vcycle<double>(int, bool);
template void SolverState::vcycle<float>(int, bool);
template void SolverState::op<double>(int, int, const double *, double *, const double *);
template void SolverState::op<float>(int, int, const float *, float *, const float *);

}  // namespace t3d

using namespace t3d;

extern "C" {

// stencil_solvers.c:399-441
void allocateSolverData(const struct gridContext gc, const int nl,
                        struct SolverDataBuffer *data) {
  std::memset(data, 0, sizeof(*data));
  DevGrid *g = dev_grid(&gc);
  if (!g) return;
  if (nl != g->nl) {
    set_error(TOP3D_ERR_ARG, "allocateSolverData: nl=%d but the grid has %d levels", nl, g->nl);
    return;
  }
  SolverState *s = new SolverState();
  if (!s->allocate(g, nl)) {
    s->release();
    delete s;
    return;
  }
  data->r = s->r;
  data->p = s->p;
  data->q = s->q;
  data->z = s->z;
  data->invD = (MTYPE **)std::calloc(nl, sizeof(MTYPE *));
  data->dmg = (CTYPE **)std::calloc(nl, sizeof(CTYPE *));
  data->rmg = (CTYPE **)std::calloc(nl, sizeof(CTYPE *));
  data->zmg = (CTYPE **)std::calloc(nl, sizeof(CTYPE *));
  for (int l = 0; l < nl; l++) {
    data->invD[l] = s->invD64[l];
    data->dmg[l] = s->v64[l][SolverState::kD];
    data->rmg[l] = s->v64[l][SolverState::kR];
    data->zmg[l] = s->v64[l][SolverState::kZ];
  }
  data->coarseMatrices = (struct CSRMatrix *)std::calloc(nl, sizeof(struct CSRMatrix));
  data->bottomSolver.cholmodCommon = s;
}

void freeSolverData(struct SolverDataBuffer *data, const int nl) {
  SolverState *s = solver_state(data);
  if (!s) return;
  s->release();
  delete s;
  std::free(data->invD);
  std::free(data->dmg);
  std::free(data->rmg);
  std::free(data->zmg);
  std::free(data->coarseMatrices);
  std::memset(data, 0, sizeof(*data));
  (void)nl;
}

void top3d_set_precision(struct SolverDataBuffer *data, int mode) {
  SolverState *s = solver_state(data);
  if (s) s->mixed = (mode == TOP3D_MIXED);
}
void top3d_set_bc_mode(struct SolverDataBuffer *data, int mode) {
  SolverState *s = solver_state(data);
  if (s) s->bcsym = (mode == TOP3D_BC_SYMMETRIC);
}
void top3d_set_max_iter(struct SolverDataBuffer *data, int max_iter) {
  SolverState *s = solver_state(data);
  if (s) s->max_iter = max_iter > 0 ? max_iter : 1000;
}
void top3d_solver_stats(const struct SolverDataBuffer *data, struct Top3dSolveStats *out) {
  SolverState *s = solver_state(data);
  if (s && out) *out = s->stats;
}

int top3d_sweep_timing(struct SolverDataBuffer *data, int enable, double *avg_ms, int *count,
                       double *bytes) {
  SolverState *s = solver_state(data);
  if (!s) return top3d_last_error();
  if (enable) {
    s->sw_ms = 0.0;
    s->sw_count = 0;
    s->sw_used = 0;
  }
  s->sweep_timing = enable != 0;
  if (avg_ms) *avg_ms = s->sw_count ? s->sw_ms / s->sw_count : 0.0;
  if (count) *count = s->sw_count;
  if (bytes) {  // u, b (and invD unless formed from E) in; u out; E
    const DevGrid &g = *s->g;
    const double nn0 = (double)g.lv[0].interior_nodes(), ne0 = (double)g.lv[0].elements();
    const double vec = fine_sweep_forms_diagonal() ? 3.0 : 4.0;
    *bytes = s->mixed ? 12.0 * vec * nn0 + 4.0 * ne0 : 24.0 * vec * nn0 + 8.0 * ne0;
  }
  return TOP3D_OK;
}

// `which` flag: time each launch alone, with the memory-side cache flushed
static constexpr int kTimeCold = 16;

int top3d_time_kernel(struct SolverDataBuffer *data, int which, int reps, double *avg_ms,
                      double *bytes) {
  SolverState *s = solver_state(data);
  if (!s) return top3d_last_error();
  const bool cold = (which & kTimeCold) != 0;
  which &= ~kTimeCold;
  if (reps < 1 || which < 0 || which > 5) {
    set_error(TOP3D_ERR_ARG, "top3d_time_kernel: which=%d reps=%d", which, reps);
    return TOP3D_ERR_ARG;
  }
  const DevGrid &g = *s->g;
  const Level &L0 = g.lv[0], &L1 = g.lv[1];
  const double nn0 = (double)L0.interior_nodes(), ne0 = (double)L0.elements();
  const double nn1 = (double)L1.interior_nodes();
  // the FP32 buffers hold data only after a mixed solve; derive them here
  launch_convert<float, double>(g.nelem, s->E64, s->E32);
  for (int l = 0; l < 2; l++) {
    launch_convert<float, double>(g.lv[l].ndof, s->invD64[l], s->invD32[l]);
    launch_convert<float, double>(g.lv[l].ndof, s->v64[l][SolverState::kD],
                                  s->v32[l][SolverState::kD]);
  }
  // the level-1 sweeps' transformed moduli, formed before the timed launches
  const void *eh1 = nullptr;
  if (which == 3 || which == 4) {
    void *w = s->g->work(kSlotLevel1Moduli, (which == 3 ? 4 : 8) * level1_moduli_count(g));
    if (!w) return top3d_last_error();
    if (which == 3) launch_level1_moduli<float>(g, s->E32, (float *)w);
    else launch_level1_moduli<double>(g, s->E64, (double *)w);
    eh1 = w;
  }
  auto launch = [&](int rep) {
    const bool odd = rep & 1;
    switch (which) {
      case 0: {
        float *a = s->v32[0][SolverState::kZ], *b = s->t32[0];
        launch_fine<float>(g, 3, s->E32, odd ? b : a, odd ? a : b, s->v32[0][SolverState::kD],
                           s->invD32[0], (float)kOmega, s->bcsym);
        break;
      }
      case 1: {
        double *a = s->z, *b = s->t64[0];
        launch_fine<double>(g, 3, s->E64, odd ? b : a, odd ? a : b, s->v64[0][SolverState::kD],
                            s->invD64[0], kOmega, s->bcsym);
        break;
      }
      case 2:
        launch_fine<double>(g, 0, s->E64, s->p, s->q, nullptr, nullptr, 0.0, s->bcsym);
        break;
      case 5:  // as the CG runs it: with the per-workgroup partial sums of p.q
        launch_fine_apply_dot(g, s->E64, s->p, s->q, s->bcsym, s->cgpart + SolverState::kPartRZ,
                              SolverState::kPartPQ);
        break;
      case 3: {
        float *a = s->v32[1][SolverState::kZ], *b = s->t32[1];
        launch_level1<float>(g, 1, s->E32, odd ? b : a, odd ? a : b, s->v32[1][SolverState::kD],
                             s->invD32[1], (float)kOmega, s->bcsym, (const float *)eh1);
        break;
      }
      default: {
        double *a = s->v64[1][SolverState::kZ], *b = s->t64[1];
        launch_level1<double>(g, 1, s->E64, odd ? b : a, odd ? a : b,
                              s->v64[1][SolverState::kD], s->invD64[1], kOmega, s->bcsym,
                              (const double *)eh1);
      }
    }
  };
  launch(0);  // warm-up
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  float ms = 0.f;
  if (!cold) {
    hipEventRecord(e0, stream());
    for (int r = 0; r < reps; r++) launch(r);
    hipEventRecord(e1, stream());
    if (!T3D_HIP(hipEventSynchronize(e1))) return TOP3D_ERR_HIP;
    hipEventElapsedTime(&ms, e0, e1);
  } else {
    // before each launch, a device-to-device copy of kFlushBytes moves
    // 2 x kFlushBytes through the memory-side cache (256 MiB on MI355X), so
    // the kernel finds none of its operands there, as in a solve whose
    // other vectors pass between two launches; each launch is timed alone
    constexpr size_t kFlushBytes = size_t(512) << 20;
    char *flush = nullptr;
    bool ok = T3D_HIP(hipMalloc(&flush, 2 * kFlushBytes));
    for (int r = 0; ok && r < reps; r++) {
      hipMemcpyAsync(flush + (r & 1) * kFlushBytes, flush + (~r & 1) * kFlushBytes, kFlushBytes,
                     hipMemcpyDeviceToDevice, stream());
      hipEventRecord(e0, stream());
      launch(r);
      hipEventRecord(e1, stream());
      ok = T3D_HIP(hipEventSynchronize(e1));
      float m = 0.f;
      if (ok) hipEventElapsedTime(&m, e0, e1);
      ms += m;
    }
    if (flush) hipFree(flush);
    if (!ok) {
      hipEventDestroy(e0);
      hipEventDestroy(e1);
      return TOP3D_ERR_HIP;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (avg_ms) *avg_ms = ms / reps;
  if (bytes) {
    // fine sweeps: u, b (and invD unless formed from E) in; u out; E
    const double vec = fine_sweep_forms_diagonal() ? 3.0 : 4.0;
    switch (which) {
      case 0: *bytes = 12.0 * vec * nn0 + 4.0 * ne0; break;
      case 1: *bytes = 24.0 * vec * nn0 + 8.0 * ne0; break;
      case 2:
      case 5: *bytes = 48.0 * nn0 + 8.0 * ne0; break;   // p in, q out, E
      case 3: *bytes = 48.0 * nn1 + 4.0 * ne0; break;   // level-1 vectors, fine E
      default: *bytes = 96.0 * nn1 + 8.0 * ne0;
    }
  }
  return TOP3D_OK;
}

// stencil_solvers.c:259-397
void solveStateMG_halo(const struct gpuGrid *gpu_gc, DTYPE *x, const int nswp, const int nl,
                       const CTYPE tol, struct SolverDataBuffer *data, int *finalIter,
                       float *finalRes, CTYPE *b, STYPE *u) {
  if (!gpu_gc || gpu_gc->num_targets != 1 || !gpu_gc->target) {
    set_error(TOP3D_ERR_ARG, "solveStateMG_halo: one GPU target expected");
    return;
  }
  SolverState *s = solver_state(data);
  if (!s) return;
  if (nl != s->nl || nswp < 1) {
    set_error(TOP3D_ERR_ARG, "solveStateMG_halo: nl=%d nswp=%d (solver has %d levels)", nl,
              nswp, s->nl);
    return;
  }
  const double *dx = (const double *)dptr(x, "solveStateMG_halo x");
  const double *db = (const double *)dptr(b, "solveStateMG_halo b");
  double *du = (double *)dptr(u, "solveStateMG_halo u");
  if (!dx || !db || !du) return;
  s->nswp = nswp;
  s->solve(dx, db, du, tol, finalIter, finalRes);
}

}  // extern "C"
