// Design-loop kernels: compliance and sensitivity, the density filter and
// its transpose, volume, change and the optimality-criteria update.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

#include "t3d_internal.hpp"

namespace t3d {

constexpr int kBS = 256;
constexpr int kPartials = 1024;

__host__ __device__ constexpr int dn_x(int n) { return (n == 1 || n == 2 || n == 5 || n == 6) ? 1 : 0; }
__host__ __device__ constexpr int dn_y(int n) { return (n == 0 || n == 1 || n == 4 || n == 5) ? 1 : 0; }
__host__ __device__ constexpr int dn_z(int n) { return n >= 4 ? 1 : 0; }

__device__ __forceinline__ double blk_sum(double v) {
  __shared__ double sm[kBS / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBS / 64; i++) s += sm[i];
  __syncthreads();
  return s;
}

__device__ __forceinline__ double blk_max(double v) {
  __shared__ double sm[kBS / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBS / 64; i++) s = fmax(s, sm[i]);
  __syncthreads();
  return s;
}

__global__ void k_reduce_final(const double *__restrict__ partials, int n, int op,
                               double *__restrict__ out) {
  double s = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    s = op ? fmax(s, partials[i]) : s + partials[i];
  s = op ? blk_max(s) : blk_sum(s);
  if (threadIdx.x == 0) *out = s;
}

// ---- compliance and sensitivity (stencil_methods.c:808-862) ---------------
// ce = u_e^T KE u_e; c = sum E(x_e) ce; dc_e = -3 (E0-Emin) x_e^2 ce.
__global__ __launch_bounds__(kBS) void k_sensitivity(
    int nelx, int nely, int nelz, int wy, int wz, int few, int fez, double E0, double Emin,
    const double *__restrict__ KE, const double *__restrict__ x,
    const double *__restrict__ u, double *__restrict__ dc, double *__restrict__ partials) {
  const int64_t ne = (int64_t)nelx * nely * nelz;
  double csum = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ne;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(t % nely) + 1, k = (int)((t / nely) % nelz) + 1,
              i = (int)(t / ((int64_t)nely * nelz)) + 1;
    double ue[24];
#pragma unroll
    for (int n = 0; n < 8; n++) {
      const int64_t nd = 3 * (((int64_t)(i + dn_x(n)) * wz + (k + dn_z(n))) * wy + (j + dn_y(n)));
      ue[3 * n] = u[nd];
      ue[3 * n + 1] = u[nd + 1];
      ue[3 * n + 2] = u[nd + 2];
    }
    double ce = 0.0;
#pragma unroll 4
    for (int r = 0; r < 24; r++) {
      double tr = 0.0;
#pragma unroll
      for (int c = 0; c < 24; c++) tr += KE[24 * r + c] * ue[c];
      ce += ue[r] * tr;
    }
    const int64_t e = ((int64_t)i * fez + k) * few + j;
    const double xe = x[e];
    csum += ce * (Emin + xe * xe * xe * (E0 - Emin));
    dc[e] = ce * (-3.0 * (E0 - Emin) * xe * xe);
  }
  csum = blk_sum(csum);
  if (threadIdx.x == 0) partials[blockIdx.x] = csum;
}

// The same quantities with u_e^T KE u_e = h^T (T KE T^T / 64) h, h = T u_e the
// per-component 8-point Walsh-Hadamard transform of the element's nodal
// values: 72 additions and 8 3x3 blocks instead of 576 multiply-adds with KE
// read per element.  Element coordinates in 32-bit arithmetic (ne < 2^31).
struct WhBlocks {
  double m[8][3][3];
};

__global__ __launch_bounds__(kBS) void k_sensitivity_wh(
    int nelx, int nely, int nelz, int wy, int wz, int few, int fez, double E0, double Emin,
    const WhBlocks B, const double *__restrict__ x, const double *__restrict__ u,
    double *__restrict__ dc, double *__restrict__ partials) {
  const uint32_t ne = (uint32_t)nelx * nely * nelz;
  double csum = 0.0;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ne; t += gridDim.x * blockDim.x) {
    const uint32_t ty = t / (uint32_t)nely, tz = ty / (uint32_t)nelz;
    const int j = (int)(t - ty * nely) + 1, k = (int)(ty - tz * nelz) + 1, i = (int)tz + 1;
    double h[8][3];  // [p][c], p = px | py << 1 | pz << 2
#pragma unroll
    for (int p = 0; p < 8; p++) {
      const int64_t nd =
          3 * (((int64_t)(i + (p & 1)) * wz + (k + (p >> 2))) * wy + (j + ((p >> 1) & 1)));
      h[p][0] = u[nd];
      h[p][1] = u[nd + 1];
      h[p][2] = u[nd + 2];
    }
#pragma unroll
    for (int bit = 1; bit < 8; bit <<= 1)
#pragma unroll
      for (int p = 0; p < 8; p++)
        if (!(p & bit))
#pragma unroll
          for (int c = 0; c < 3; c++) {
            const double a = h[p][c], b = h[p | bit][c];
            h[p][c] = a + b;
            h[p | bit][c] = a - b;
          }
    double ce = 0.0;
#pragma unroll
    for (int s = 0; s < 8; s++) {
      const double hs[3] = {h[s ^ 1][0], h[s ^ 2][1], h[s ^ 4][2]};
#pragma unroll
      for (int c = 0; c < 3; c++) {
        double tr = 0.0;
#pragma unroll
        for (int d = 0; d < 3; d++) tr += B.m[s][c][d] * hs[d];
        ce += hs[c] * tr;
      }
    }
    const int64_t e = ((int64_t)i * fez + k) * few + j;
    const double xe = x[e];
    csum += ce * (Emin + xe * xe * xe * (E0 - Emin));
    dc[e] = ce * (-3.0 * (E0 - Emin) * xe * xe);
  }
  csum = blk_sum(csum);
  if (threadIdx.x == 0) partials[blockIdx.x] = csum;
}

void launch_sensitivity(const DevGrid &g, const double *x, const double *u, double *dc,
                        double *partials, double *c) {
  const Level &L = g.lv[0];
  if (g.wh0 && (int64_t)L.nelx * L.nely * L.nelz < (int64_t)INT32_MAX) {
    WhBlocks B;
    std::memcpy(B.m, g.M0, sizeof(B.m));
    hipLaunchKernelGGL(k_sensitivity_wh, dim3(kPartials), dim3(kBS), 0, stream(), L.nelx,
                       L.nely, L.nelz, L.wy, L.wz, g.ewy, g.ewz, g.E0, g.Emin, B, x, u, dc,
                       partials);
  } else {
    hipLaunchKernelGGL(k_sensitivity, dim3(kPartials), dim3(kBS), 0, stream(), L.nelx, L.nely,
                       L.nelz, L.wy, L.wz, g.ewy, g.ewz, g.E0, g.Emin, g.KE[0], x, u, dc,
                       partials);
  }
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBS), 0, stream(), partials, kPartials, 0,
                     c);
}

// ---- density filter (stencil_optimization.c:228-395) ----------------------
// neighbourhood i2 in [i1-(ceil(r)-1), i1+ceil(r)] clipped to [1, nel],
// weight max(0, r - dist).
struct FilterWin {
  int lo, hi;
};
__device__ __forceinline__ FilterWin fwin(int i1, int cr, int nel) {
  return FilterWin{max(i1 - (cr - 1), 1), min(i1 + cr + 1, nel + 1)};
}

// MODE 0: out = (H rho)/Hs; MODE 1: out = v / Hs; MODE 2: out = H v
template <int MODE>
__global__ __launch_bounds__(kBS) void k_filter(int nelx, int nely, int nelz, int few,
                                                int fez, double rmin, int cr,
                                                const double *__restrict__ in,
                                                double *__restrict__ out) {
  const int64_t ne = (int64_t)nelx * nely * nelz;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ne) return;
  const int j1 = (int)(t % nely) + 1, k1 = (int)((t / nely) % nelz) + 1,
            i1 = (int)(t / ((int64_t)nely * nelz)) + 1;
  const FilterWin wi = fwin(i1, cr, nelx), wk = fwin(k1, cr, nelz), wj = fwin(j1, cr, nely);
  double acc = 0.0, ws = 0.0;
  for (int i2 = wi.lo; i2 < wi.hi; i2++)
    for (int k2 = wk.lo; k2 < wk.hi; k2++)
      for (int j2 = wj.lo; j2 < wj.hi; j2++) {
        const int d2 = (i1 - i2) * (i1 - i2) + (j1 - j2) * (j1 - j2) + (k1 - k2) * (k1 - k2);
        const double w = fmax(0.0, rmin - sqrt((double)d2));
        if (MODE != 1) acc += w * in[((int64_t)i2 * fez + k2) * few + j2];
        ws += w;
      }
  const int64_t e1 = ((int64_t)i1 * fez + k1) * few + j1;
  if (MODE == 0) out[e1] = acc / ws;
  if (MODE == 1) out[e1] = in[e1] / ws;
  if (MODE == 2) out[e1] = acc;
}

// The non-zero weights of the window, in the loop order of the kernel above
// (i2, then k2, then j2 ascending).  The window's zero weights add nothing to
// either sum, so leaving them out keeps both sums bit for bit; the weights use
// the same expression as the loop (rmin 2.5: 81 of 216 terms).  The list is
// padded to the kernel's unrolled length with zero weights at offset 0, which
// add +0 to both sums.
constexpr int kFilterMaxTaps = 240;
struct FilterTaps {
  int8_t di[kFilterMaxTaps], dk[kFilterMaxTaps], dj[kFilterMaxTaps];
  double w[kFilterMaxTaps];
};

static int filter_taps(double rmin, FilterTaps &t) {
  const int cr = (int)std::ceil(rmin);
  int n = 0;
  for (int di = -(cr - 1); di <= cr; di++)
    for (int dk = -(cr - 1); dk <= cr; dk++)
      for (int dj = -(cr - 1); dj <= cr; dj++) {
        const double w = std::fmax(0.0, rmin - std::sqrt((double)(di * di + dj * dj + dk * dk)));
        if (w == 0.0) continue;
        if (n == kFilterMaxTaps) return -1;
        t.di[n] = (int8_t)di;
        t.dk[n] = (int8_t)dk;
        t.dj[n] = (int8_t)dj;
        t.w[n++] = w;
      }
  for (int q = n; q < kFilterMaxTaps; q++) {
    t.di[q] = t.dk[q] = t.dj[q] = 0;
    t.w[q] = 0.0;
  }
  return n;
}

template <int MODE, int NT>
__global__ __launch_bounds__(kBS) void k_filter_taps(int nelx, int nely, int nelz, int few,
                                                     int fez, const FilterTaps taps,
                                                     const double *__restrict__ in,
                                                     double *__restrict__ out) {
  const uint32_t ne = (uint32_t)nelx * nely * nelz;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ne) return;
  const uint32_t ty = t / (uint32_t)nely, tz = ty / (uint32_t)nelz;
  const int j1 = (int)(t - ty * nely) + 1, k1 = (int)(ty - tz * nelz) + 1, i1 = (int)tz + 1;
  const int64_t e1 = ((int64_t)i1 * fez + k1) * few + j1;
  double acc = 0.0, ws = 0.0;
  // branch-free: a tap outside the grid reads the element itself with weight
  // +0, so every load of the window can be in flight at once
#pragma unroll
  for (int q = 0; q < NT; q++) {
    const int i2 = i1 + taps.di[q], k2 = k1 + taps.dk[q], j2 = j1 + taps.dj[q];
    const bool in_grid = i2 >= 1 && i2 <= nelx && k2 >= 1 && k2 <= nelz && j2 >= 1 && j2 <= nely;
    const double w = in_grid ? taps.w[q] : 0.0;
    const int64_t off = in_grid ? ((int64_t)taps.di[q] * fez + taps.dk[q]) * few + taps.dj[q] : 0;
    if (MODE != 1) acc += w * in[e1 + off];
    ws += w;
  }
  if (MODE == 0) out[e1] = acc / ws;
  if (MODE == 1) out[e1] = in[e1] / ws;
  if (MODE == 2) out[e1] = acc;
}

// Radius at most 2 (window offsets -1..2 per axis): the offsets with
// di^2 + dk^2 + dj^2 <= D2MAX are compile-time constants, the weight depends
// only on that squared distance (w2[d2], same expression as the loop), and
// the per-axis range checks are formed once per element.
template <int MODE, int D2MAX>
__global__ __launch_bounds__(kBS) void k_filter_r2(int nelx, int nely, int nelz, int few,
                                                   int fez, double w0, double w1, double w2,
                                                   double w3, const double *__restrict__ in,
                                                   double *__restrict__ out) {
  const uint32_t ne = (uint32_t)nelx * nely * nelz;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ne) return;
  const uint32_t ty = t / (uint32_t)nely, tz = ty / (uint32_t)nelz;
  const int j1 = (int)(t - ty * nely) + 1, k1 = (int)(ty - tz * nelz) + 1, i1 = (int)tz + 1;
  const int64_t e1 = ((int64_t)i1 * fez + k1) * few + j1;
  const int64_t plane = (int64_t)fez * few;
  const double wd[4] = {w0, w1, w2, w3};
  bool oi[4], ok[4], oj[4];  // offset d = index - 1 inside the grid
#pragma unroll
  for (int d = 0; d < 4; d++) {
    oi[d] = i1 + d - 1 >= 1 && i1 + d - 1 <= nelx;
    ok[d] = k1 + d - 1 >= 1 && k1 + d - 1 <= nelz;
    oj[d] = j1 + d - 1 >= 1 && j1 + d - 1 <= nely;
  }
  double acc = 0.0, ws = 0.0;
#pragma unroll
  for (int di = -1; di <= 2; di++)
#pragma unroll
    for (int dk = -1; dk <= 2; dk++)
#pragma unroll
      for (int dj = -1; dj <= 2; dj++) {
        const int d2 = di * di + dk * dk + dj * dj;
        if (d2 > D2MAX) continue;
        const bool in_grid = oi[di + 1] && ok[dk + 1] && oj[dj + 1];
        const double w = in_grid ? wd[d2] : 0.0;
        if (MODE != 1) acc += w * in[in_grid ? e1 + di * plane + dk * few + dj : e1];
        ws += w;
      }
  if (MODE == 0) out[e1] = acc / ws;
  if (MODE == 1) out[e1] = in[e1] / ws;
  if (MODE == 2) out[e1] = acc;
}

template <int MODE>
static void filter_pass(const DevGrid &g, double rmin, const double *in, double *out) {
  const Level &L = g.lv[0];
  const int64_t ne = L.elements();
  const dim3 grid((ne + kBS - 1) / kBS);
  if (ne < (int64_t)INT32_MAX && rmin > 0.0 && rmin <= 2.0) {
    double w[4];
    int d2max = -1;
    for (int d2 = 0; d2 < 4; d2++) {
      w[d2] = std::fmax(0.0, rmin - std::sqrt((double)d2));
      if (w[d2] != 0.0) d2max = d2;
    }
#define T3D_FILTER_R2(D)                                                                    \
  hipLaunchKernelGGL((k_filter_r2<MODE, D>), grid, dim3(kBS), 0, stream(), L.nelx, L.nely, \
                     L.nelz, g.ewy, g.ewz, w[0], w[1], w[2], w[3], in, out)
    if (d2max == 0) T3D_FILTER_R2(0);
    if (d2max == 1) T3D_FILTER_R2(1);
    if (d2max == 2) T3D_FILTER_R2(2);
    if (d2max == 3) T3D_FILTER_R2(3);
#undef T3D_FILTER_R2
    return;
  }
  FilterTaps taps;
  const int n = ne < (int64_t)INT32_MAX ? filter_taps(rmin, taps) : -1;
#define T3D_FILTER_TAPS(NT)                                                                  \
  hipLaunchKernelGGL((k_filter_taps<MODE, NT>), grid, dim3(kBS), 0, stream(), L.nelx, L.nely, \
                     L.nelz, g.ewy, g.ewz, taps, in, out)
  if (n >= 0 && n <= 64) T3D_FILTER_TAPS(64);
  else if (n >= 0) T3D_FILTER_TAPS(kFilterMaxTaps);
  else
    hipLaunchKernelGGL((k_filter<MODE>), grid, dim3(kBS), 0, stream(), L.nelx, L.nely, L.nelz,
                       g.ewy, g.ewz, rmin, (int)std::ceil(rmin), in, out);
#undef T3D_FILTER_TAPS
}

void launch_filter(const DevGrid &g, double rmin, const double *rho, double *out) {
  filter_pass<0>(g, rmin, rho, out);
}

void launch_filter_grad(const DevGrid &g, double rmin, const double *v, double *tmp,
                        double *out) {
  launch_fill<double>(g.nelem, tmp, 0.0);
  filter_pass<1>(g, rmin, v, tmp);
  filter_pass<2>(g, rmin, tmp, out);
}

// ---- volume (gpu_definitions.c:94-110) ------------------------------------
__global__ __launch_bounds__(kBS) void k_volume(int nelx, int nely, int nelz, int few, int fez,
                                                const double *__restrict__ x,
                                                double *__restrict__ partials) {
  const int64_t ne = (int64_t)nelx * nely * nelz;
  double s = 0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ne;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(t % nely) + 1, k = (int)((t / nely) % nelz) + 1,
              i = (int)(t / ((int64_t)nely * nelz)) + 1;
    s += x[((int64_t)i * fez + k) * few + j];
  }
  s = blk_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

void launch_volume(const DevGrid &g, const double *x, double *partials, double *result) {
  const Level &L = g.lv[0];
  hipLaunchKernelGGL(k_volume, dim3(kPartials), dim3(kBS), 0, stream(), L.nelx, L.nely,
                     L.nelz, g.ewy, g.ewz, x, partials);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBS), 0, stream(), partials, kPartials, 0,
                     result);
}

// ---- change = max |x - xnew|, then x = xnew (gpu_definitions.c:112-125) ----
__global__ __launch_bounds__(kBS) void k_change(int64_t n, double *__restrict__ x,
                                                const double *__restrict__ xnew,
                                                double *__restrict__ partials) {
  double m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    m = fmax(m, fabs(x[i] - xnew[i]));
    x[i] = xnew[i];
  }
  m = blk_max(m);
  if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

void launch_change(int64_t n, double *x, const double *xnew, double *partials, double *result) {
  hipLaunchKernelGGL(k_change, dim3(kPartials), dim3(kBS), 0, stream(), n, x, xnew, partials);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBS), 0, stream(), partials, kPartials, 1,
                     result);
}

// ---- optimality criteria (gpu_definitions.c:127-157) ----------------------
// The bisection of gpu_definitions.c:127-157 evaluates one multiplier per
// pass over the elements.  Here one pass evaluates the kOcDepth levels of the
// bisection tree below the current interval (2^kOcDepth - 1 multipliers), and
// a one-block kernel walks the tree with the sums, so the multipliers visited
// are those of the step-by-step loop.  The element arrays are read once per
// kOcDepth steps.  The candidate x_e sqrt(-dc_e / (dv_e lmid)) is formed as
// x_e sqrt(-dc_e / dv_e) * (1 / sqrt(lmid)), one division and square root per
// element instead of per multiplier (a few ulp from the one-step form).
constexpr int kOcDepth = 3;
constexpr int kOcCand = (1 << kOcDepth) - 1;
constexpr int kOcChunk = 24;  // passes between host checks (72 bisection steps)
static_assert(kOcScratchDoubles >= 8 + kOcCand * kPartials, "OC scratch size");

// Multipliers of the tree below [l1, l2], in heap order: node n has the
// lower half as child 2n+1 and the upper half as child 2n+2.
__device__ __forceinline__ void oc_tree(double l1, double l2, double (&mid)[kOcCand]) {
  double lo[kOcCand], hi[kOcCand];
  lo[0] = l1;
  hi[0] = l2;
#pragma unroll
  for (int n = 0; n < kOcCand; n++) {
    mid[n] = 0.5 * (hi[n] + lo[n]);
    if (2 * n + 2 < kOcCand) {
      lo[2 * n + 1] = lo[n];
      hi[2 * n + 1] = mid[n];
      lo[2 * n + 2] = mid[n];
      hi[2 * n + 2] = hi[n];
    }
  }
}

// xnew_e for a multiplier with rl = 1 / sqrt(lmid), a = x_e sqrt(-dc_e / dv_e):
// move limit 0.2, bounds [0, 1].
__device__ __forceinline__ double oc_x(double xi, double a, double rl) {
  const double move = 0.2;
  return fmax(0.0, fmax(xi - move, fmin(1.0, fmin(xi + move, a * rl))));
}

// state: [0] l1, [1] l2, [3] converged flag, [4] gval, [5] last multiplier,
// [6] steps taken.  partials: kOcCand rows of gridDim.x block sums.
__global__ __launch_bounds__(kBS) void k_oc_pass(int64_t n, const double *__restrict__ x,
                                                 const double *__restrict__ dv,
                                                 const double *__restrict__ dc,
                                                 const double *__restrict__ state,
                                                 double *__restrict__ partials) {
  if (state[3] != 0.0) return;
  double rl[kOcCand];
  oc_tree(state[0], state[1], rl);
  double s[kOcCand];
#pragma unroll
  for (int c = 0; c < kOcCand; c++) {
    rl[c] = 1.0 / sqrt(rl[c]);
    s[c] = 0.0;
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double xi = x[i], dvi = dv[i];
    const double a = xi * sqrt(-dc[i] / dvi);
#pragma unroll
    for (int c = 0; c < kOcCand; c++) s[c] += dvi * (oc_x(xi, a, rl[c]) - xi);
  }
#pragma unroll
  for (int c = 0; c < kOcCand; c++) {
    const double t = blk_sum(s[c]);
    if (threadIdx.x == 0) partials[(int64_t)c * gridDim.x + blockIdx.x] = t;
  }
}

// One workgroup of kOcCand wavefronts, wave c summing multiplier c's
// partials; then up to kOcDepth bisection steps down the tree, stopping where
// the interval has converged.
__global__ __launch_bounds__(64 * kOcCand) void k_oc_walk(const double *__restrict__ partials,
                                                          int np, double *__restrict__ state) {
  if (state[3] != 0.0) return;
  __shared__ double S[kOcCand];
  const int c = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double s = 0;
  for (int i = lane; i < np; i += 64) s += partials[c * np + i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) S[c] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double l1 = state[0], l2 = state[1];
    int node = 0;
    for (int d = 0; d < kOcDepth; d++) {
      const double lmid = 0.5 * (l2 + l1);
      state[5] = lmid;
      state[6] += 1.0;
      if (S[node] + state[4] > 0) {
        l1 = lmid;
        node = 2 * node + 2;
      } else {
        l2 = lmid;
        node = 2 * node + 1;
      }
      if (!((l2 - l1) / (l1 + l2) > 1e-6)) {
        state[3] = 1.0;
        break;
      }
    }
    state[0] = l1;
    state[1] = l2;
  }
}

// xnew for the last multiplier the walk visited.
__global__ __launch_bounds__(kBS) void k_oc_write(int64_t n, const double *__restrict__ x,
                                                  double *__restrict__ xnew,
                                                  const double *__restrict__ dv,
                                                  const double *__restrict__ dc,
                                                  const double *__restrict__ state) {
  const double rl = 1.0 / sqrt(state[5]);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double xi = x[i];
    xnew[i] = oc_x(xi, xi * sqrt(-dc[i] / dv[i]), rl);
  }
}

// the bisection state of one OC update: [l1, l2] = [0, 1e9], not converged
__global__ void k_oc_init(double gval, double *__restrict__ state) {
  if (threadIdx.x < 8) {
    const double init[8] = {0.0, 1e9, 0.0, 0.0, gval, 0.0, 0.0, 0.0};
    state[threadIdx.x] = init[threadIdx.x];
  }
}

// Runs the bisection, checking on the host every kOcChunk passes (passes
// after convergence return at once).  xnew is written after each chunk for
// the last multiplier, so a converged chunk needs no further launch.
// scratch (kOcScratchDoubles doubles of device memory) is initialised here:
// its previous contents do not matter.  Returns the number of bisection
// steps or -1.
int oc_update(const double *x, double *xnew, const double *dv, const double *dc,
              int64_t nelem, double gval, double *scratch) {
  double *state = scratch;          // 8 doubles
  double *partials = scratch + 8;   // kOcCand * kPartials doubles
  hipLaunchKernelGGL(k_oc_init, dim3(1), dim3(64), 0, stream(), gval, state);
  if (!T3D_HIP(hipGetLastError())) return -1;
  double h[8];
  for (int it = 0; it < 4096; it += kOcChunk * kOcDepth) {
    for (int q = 0; q < kOcChunk; q++) {
      hipLaunchKernelGGL(k_oc_pass, dim3(kPartials), dim3(kBS), 0, stream(), nelem, x, dv, dc,
                         state, partials);
      hipLaunchKernelGGL(k_oc_walk, dim3(1), dim3(64 * kOcCand), 0, stream(), partials,
                         kPartials, state);
    }
    hipLaunchKernelGGL(k_oc_write, dim3(kPartials), dim3(kBS), 0, stream(), nelem, x, xnew, dv,
                       dc, state);
    if (!T3D_HIP(hipMemcpyAsync(h, state, sizeof(h), hipMemcpyDeviceToHost, stream())) ||
        !T3D_HIP(hipStreamSynchronize(stream())))
      return -1;
    if (h[3] != 0.0) return (int)h[6];
  }
  set_error(TOP3D_ERR_NO_CONVERGE, "update_solution: OC bisection did not converge in 4096 steps");
  return -1;
}

}  // namespace t3d
