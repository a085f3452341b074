// Matrix-free level operators, smoother sweeps, transfers and inverse
// diagonals for gfx950.
//
// Layout (shared with the reference): node (i,j,k) of a level lives at
// i*wy*wz + k*wy + j, DOF 3*node + comp; fine element (i,j,k) at
// i*few*fez + k*few + j.  Real nodes are 1..nel+1 and real elements 1..nel in
// each direction; everything else is halo and stays zero.
//
// Operator modes (one kernel per level, the update fused into the operator):
//   kApply    out = K in                     (fixed DOFs: out = in)
//   kJacobi   out = in + w invD (b - K in)   (interior nodes; else out = in)
//   kResidual out = b - K in
#include <cstdlib>
#include <string>

#include "t3d_internal.hpp"

namespace t3d {

enum { kApply = 0, kJacobi = 1, kResidual = 2 };

// local node positions (x, y, z), stencil_utility.h:19-26
__host__ __device__ constexpr int lnode_x(int n) { return (n == 1 || n == 2 || n == 5 || n == 6) ? 1 : 0; }
__host__ __device__ constexpr int lnode_y(int n) { return (n == 0 || n == 1 || n == 4 || n == 5) ? 1 : 0; }
__host__ __device__ constexpr int lnode_z(int n) { return n >= 4 ? 1 : 0; }
__host__ __device__ constexpr int lnode_of(int ox, int oy, int oz) {
  return oz * 4 + (oy ? (ox ? 1 : 0) : (ox ? 2 : 3));
}

template <typename T>
__device__ __forceinline__ T finish(int mode, T kin, T in, T b, T invD, T omega,
                                    bool fixed) {
  if (fixed) kin = in;
  if (mode == kApply) return kin;
  if (mode == kJacobi) return in + omega * invD * (b - kin);
  return b - kin;
}

// ---------------------------------------------------------------------------
// Young's modulus per fine element: Emin + x^3 (E0 - Emin) inside, 0 in the
// halo (stencil_methods.h:105-114).
template <typename T, typename TX>
__global__ void k_young(int nelx, int nely, int nelz, int few, int fez,
                        int64_t nelem, double E0, double Emin,
                        const TX *__restrict__ x, T *__restrict__ E) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  const int j = (int)(e % few);
  const int k = (int)((e / few) % fez);
  const int i = (int)(e / ((int64_t)few * fez));
  T v = 0;
  if (i >= 1 && i <= nelx && j >= 1 && j <= nely && k >= 1 && k <= nelz) {
    const TX xe = x[e];
    const TX x3 = xe * xe * xe;
    v = (T)(Emin + (double)x3 * (E0 - Emin));
  }
  E[e] = v;
}

template <typename T>
void launch_young(const DevGrid &g, const double *x, T *E) {
  const int bs = 256;
  const int64_t nb = (g.nelem + bs - 1) / bs;
  hipLaunchKernelGGL((k_young<T, double>), dim3(nb), dim3(bs), 0, stream(), g.nelx,
                     g.nely, g.nelz, g.ewy, g.ewz, g.nelem, g.E0, g.Emin, x, E);
}
template <typename T>
void launch_young_float_input(const DevGrid &g, const float *x, T *E) {
  const int bs = 256;
  const int64_t nb = (g.nelem + bs - 1) / bs;
  hipLaunchKernelGGL((k_young<T, float>), dim3(nb), dim3(bs), 0, stream(), g.nelx,
                     g.nely, g.nelz, g.ewy, g.ewz, g.nelem, g.E0, g.Emin, x, E);
}

// ---------------------------------------------------------------------------
// Fine operator, one thread per node: sum over the 27 neighbour nodes and
// the (up to 8) elements each shares with the node (stencil_methods.c:7-428).
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_fine(
    LevelArgs a, const T *__restrict__ KE, const T *__restrict__ E,
    const T *__restrict__ in, T *__restrict__ out, const T *__restrict__ b,
    const T *__restrict__ invD, T omega, const uint8_t *__restrict__ fix,
    int bcsym) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y * blockDim.y + threadIdx.y;
  const int i = blockIdx.z;
  if (j >= a.wy || k >= a.wz) return;
  const int64_t node = ((int64_t)i * a.wz + k) * a.wy + j;
  const int64_t o = 3 * node;
  const bool inside = i >= 1 && i <= a.nelx + 1 && j >= 1 && j <= a.nely + 1 &&
                      k >= 1 && k <= a.nelz + 1;
  if (!inside) {
    // the reference zeroes only the padded y rows of `out` and leaves the
    // other halo nodes untouched (stencil_methods.c:407-418)
    if (MODE == kApply && j <= a.nely + 1) return;
    for (int c = 0; c < 3; c++) {
      const T v = MODE == kApply ? T(0) : (MODE == kJacobi ? in[o + c] : b[o + c]);
      out[o + c] = v;
    }
    return;
  }
  T Ee[8];
#pragma unroll
  for (int e = 0; e < 8; e++) {
    const int ex = (e & 1) ? 0 : -1, ey = (e & 2) ? 0 : -1, ez = (e & 4) ? 0 : -1;
    Ee[e] = E[((int64_t)(i + ex) * a.fez + (k + ez)) * a.few + (j + ey)];
  }
  T t[8][3];
#pragma unroll
  for (int e = 0; e < 8; e++) t[e][0] = t[e][1] = t[e][2] = 0;
  const int64_t sx = (int64_t)a.wy * a.wz, sz = a.wy;
#pragma unroll
  for (int dx = -1; dx <= 1; dx++)
#pragma unroll
    for (int dz = -1; dz <= 1; dz++)
#pragma unroll
      for (int dy = -1; dy <= 1; dy++) {
        const int64_t nb = node + dx * sx + dz * sz + dy;
        T v[3] = {in[3 * nb], in[3 * nb + 1], in[3 * nb + 2]};
        if (bcsym) {
          const uint8_t m = fix[nb];
          if (m & 1) v[0] = 0;
          if (m & 2) v[1] = 0;
          if (m & 4) v[2] = 0;
        }
#pragma unroll
        for (int e = 0; e < 8; e++) {
          const int ex = (e & 1) ? 0 : -1, ey = (e & 2) ? 0 : -1, ez = (e & 4) ? 0 : -1;
          const int sxo = dx - ex, syo = dy - ey, szo = dz - ez;
          if (sxo < 0 || sxo > 1 || syo < 0 || syo > 1 || szo < 0 || szo > 1) continue;
          const int recv = lnode_of(-ex, -ey, -ez);
          const int send = lnode_of(sxo, syo, szo);
#pragma unroll
          for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++)
              t[e][r] += KE[(3 * recv + r) * 24 + 3 * send + c] * v[c];
        }
      }
  const uint8_t m = fix[node];
#pragma unroll
  for (int r = 0; r < 3; r++) {
    T acc = 0;
#pragma unroll
    for (int e = 0; e < 8; e++) acc += Ee[e] * t[e][r];
    const T bb = MODE == kApply ? T(0) : b[o + r];
    const T dd = MODE == kJacobi ? invD[o + r] : T(0);
    out[o + r] = finish<T>(MODE, acc, in[o + r], bb, dd, omega, (m >> r) & 1);
  }
}

// node-centric variant kept for cross-checking (TOP3D_FINE_KERNEL=node)
static bool use_node_kernel() {
  static int v = -1;
  if (v < 0) {
    const char *e = std::getenv("TOP3D_FINE_KERNEL");
    v = (e && std::string(e) == "node") ? 1 : 0;
  }
  return v == 1;
}

bool fine_sweep_forms_diagonal() { return !use_node_kernel() && tile_fine_diag_available(); }

template <typename T>
void launch_fine(const DevGrid &g, int mode, const T *E, const T *in, T *out,
                 const T *b, const T *invD, T omega, int bcsym) {
  if (!use_node_kernel()) {
    launch_fine_tile<T>(g, mode, E, in, out, b, invD, omega, bcsym);
    return;
  }
  if (mode == 3) mode = kJacobi;  // the node kernel reads invD
  const Level &L = g.lv[0];
  const LevelArgs a = g.args(0);
  const T *KE = nullptr;
  if constexpr (sizeof(T) == 8) KE = (const T *)g.KE[0]; else KE = (const T *)g.KEf[0];
  dim3 bs(64, 4, 1), gs((L.wy + 63) / 64, (L.wz + 3) / 4, L.wx);
  if (mode == kApply)
    hipLaunchKernelGGL((k_fine<T, kApply>), gs, bs, 0, stream(), a, KE, E, in, out, b,
                       invD, omega, g.fixmask[0], bcsym);
  else if (mode == kJacobi)
    hipLaunchKernelGGL((k_fine<T, kJacobi>), gs, bs, 0, stream(), a, KE, E, in, out, b,
                       invD, omega, g.fixmask[0], bcsym);
  else
    hipLaunchKernelGGL((k_fine<T, kResidual>), gs, bs, 0, stream(), a, KE, E, in, out,
                       b, invD, omega, g.fixmask[0], bcsym);
}

// A fine Jacobi sweep that also writes per-workgroup partial sums of b . out
// (at most cap of them); returns their count, or 0 when the sweep ran
// without them (node-centric kernel, or more workgroups than cap).
template <typename T>
int launch_fine_sweep_dot(const DevGrid &g, const T *E, const T *in, T *out, const T *b,
                          const T *invD, T omega, int bcsym, double *partials, int cap) {
  if (use_node_kernel()) {
    launch_fine<T>(g, 3, E, in, out, b, invD, omega, bcsym);
    return 0;
  }
  return launch_fine_tile_sweep_dot<T>(g, E, in, out, b, invD, omega, bcsym, partials, cap);
}
template int launch_fine_sweep_dot<float>(const DevGrid &, const float *, const float *, float *,
                                          const float *, const float *, float, int, double *,
                                          int);
template int launch_fine_sweep_dot<double>(const DevGrid &, const double *, const double *,
                                           double *, const double *, const double *, double, int,
                                           double *, int);

// ---------------------------------------------------------------------------
// Level-1 operator with the 8 sub-cell matrices and fine moduli
// (stencil_methods.c:436-519), one thread per coarse node.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_level1(
    LevelArgs a, const T *__restrict__ KE1, const T *__restrict__ E,
    const T *__restrict__ in, T *__restrict__ out, const T *__restrict__ b,
    const T *__restrict__ invD, T omega, const uint8_t *__restrict__ fix,
    int bcsym) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y * blockDim.y + threadIdx.y;
  const int i = blockIdx.z;
  if (j >= a.wy || k >= a.wz) return;
  const int64_t node = ((int64_t)i * a.wz + k) * a.wy + j;
  const int64_t o = 3 * node;
  const bool inside = i >= 1 && i <= a.nelx + 1 && j >= 1 && j <= a.nely + 1 &&
                      k >= 1 && k <= a.nelz + 1;
  if (!inside) {
    for (int c = 0; c < 3; c++) {
      const T v = MODE == kApply ? T(0) : (MODE == kJacobi ? in[o + c] : b[o + c]);
      out[o + c] = v;
    }
    return;
  }
  T acc[3] = {0, 0, 0};
#pragma unroll
  for (int e = 0; e < 8; e++) {
    const int ex = (e & 1) ? 0 : -1, ey = (e & 2) ? 0 : -1, ez = (e & 4) ? 0 : -1;
    const int ie = i + ex, je = j + ey, ke = k + ez;  // coarse element
    if (ie < 1 || ie > a.nelx || je < 1 || je > a.nely || ke < 1 || ke > a.nelz) continue;
    const int recv = lnode_of(-ex, -ey, -ez);
    T v[8][3];
#pragma unroll
    for (int s = 0; s < 8; s++) {
      const int64_t nb = ((int64_t)(ie + lnode_x(s)) * a.wz + (ke + lnode_z(s))) * a.wy +
                         (je + lnode_y(s));
      v[s][0] = in[3 * nb];
      v[s][1] = in[3 * nb + 1];
      v[s][2] = in[3 * nb + 2];
      if (bcsym) {
        const uint8_t m = fix[nb];
        if (m & 1) v[s][0] = 0;
        if (m & 2) v[s][1] = 0;
        if (m & 4) v[s][2] = 0;
      }
    }
    for (int ii = 0; ii < 2; ii++)
      for (int kk = 0; kk < 2; kk++)
#pragma unroll
        for (int jj = 0; jj < 2; jj++) {
          const int64_t fe = ((int64_t)(2 * (ie - 1) + ii + 1) * a.fez + (2 * (ke - 1) + kk + 1)) *
                                 a.few + (2 * (je - 1) + jj + 1);
          const T Ef = E[fe];
          const T *K = KE1 + 576 * (4 * ii + 2 * kk + jj) + 72 * recv;
          T t[3] = {0, 0, 0};
#pragma unroll
          for (int s = 0; s < 8; s++)
#pragma unroll
            for (int r = 0; r < 3; r++)
#pragma unroll
              for (int c = 0; c < 3; c++) t[r] += K[24 * r + 3 * s + c] * v[s][c];
          acc[0] += Ef * t[0];
          acc[1] += Ef * t[1];
          acc[2] += Ef * t[2];
        }
  }
  const uint8_t m = fix[node];
#pragma unroll
  for (int r = 0; r < 3; r++) {
    const T bb = MODE == kApply ? T(0) : b[o + r];
    const T dd = MODE == kJacobi ? invD[o + r] : T(0);
    out[o + r] = finish<T>(MODE, acc[r], in[o + r], bb, dd, omega, (m >> r) & 1);
  }
}

// node-centric level-1 variant kept for cross-checking (TOP3D_LEVEL1_KERNEL=node)
static bool use_node_level1() {
  static int v = -1;
  if (v < 0) {
    const char *e = std::getenv("TOP3D_LEVEL1_KERNEL");
    v = (e && std::string(e) == "node") ? 1 : 0;
  }
  return v == 1;
}

template <typename T>
void launch_level1(const DevGrid &g, int mode, const T *E, const T *in, T *out,
                   const T *b, const T *invD, T omega, int bcsym, const T *Eh) {
  if (!use_node_level1() && tile_level1_available()) {
    if (!Eh) {
      T *w = (T *)const_cast<DevGrid &>(g).work(kSlotLevel1Moduli,
                                                sizeof(T) * level1_moduli_count(g));
      if (!w) return;
      launch_level1_moduli<T>(g, E, w);
      Eh = w;
    }
    launch_level1_tile<T>(g, mode, Eh, in, out, b, invD, omega, bcsym);
    return;
  }
  const Level &L = g.lv[1];
  const LevelArgs a = g.args(1);
  const T *KE = nullptr;
  if constexpr (sizeof(T) == 8) KE = (const T *)g.KE[1]; else KE = (const T *)g.KEf[1];
  dim3 bs(64, 4, 1), gs((L.wy + 63) / 64, (L.wz + 3) / 4, L.wx);
  if (mode == kApply)
    hipLaunchKernelGGL((k_level1<T, kApply>), gs, bs, 0, stream(), a, KE, E, in, out,
                       b, invD, omega, g.fixmask[1], bcsym);
  else if (mode == kJacobi)
    hipLaunchKernelGGL((k_level1<T, kJacobi>), gs, bs, 0, stream(), a, KE, E, in, out,
                       b, invD, omega, g.fixmask[1], bcsym);
  else
    hipLaunchKernelGGL((k_level1<T, kResidual>), gs, bs, 0, stream(), a, KE, E, in,
                       out, b, invD, omega, g.fixmask[1], bcsym);
}

// ---------------------------------------------------------------------------
// Assembled levels (l >= 2): 27-point node stencil, 3x3 block per neighbour.
// S[(nb*9 + r*3 + c) * nint + n] with n = ((i-1)*nz + (k-1))*ny + (j-1) and
// nb = (dx+1)*9 + (dz+1)*3 + (dy+1).  Fixed DOFs are eliminated symmetrically
// inside S (stencil_assembly.c:226-259), so no BC pass is needed.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_stencil(
    int nx, int ny, int nz, int wy, int wz, const T *__restrict__ S,
    const T *__restrict__ in, T *__restrict__ out, const T *__restrict__ b,
    const T *__restrict__ invD, T omega) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y * blockDim.y + threadIdx.y;
  const int i = blockIdx.z;
  if (j >= wy || k >= wz) return;
  const int64_t node = ((int64_t)i * wz + k) * wy + j;
  const int64_t o = 3 * node;
  const bool inside = i >= 1 && i <= nx && j >= 1 && j <= ny && k >= 1 && k <= nz;
  if (!inside) {
    for (int c = 0; c < 3; c++) {
      const T v = MODE == kApply ? T(0) : (MODE == kJacobi ? in[o + c] : b[o + c]);
      out[o + c] = v;
    }
    return;
  }
  const int64_t nint = (int64_t)nx * ny * nz;
  const int64_t n = ((int64_t)(i - 1) * nz + (k - 1)) * ny + (j - 1);
  T acc[3] = {0, 0, 0};
#pragma unroll
  for (int dx = -1; dx <= 1; dx++)
#pragma unroll
    for (int dz = -1; dz <= 1; dz++)
#pragma unroll
      for (int dy = -1; dy <= 1; dy++) {
        const int nbi = (dx + 1) * 9 + (dz + 1) * 3 + (dy + 1);
        const int64_t q = node + ((int64_t)dx * wz + dz) * wy + dy;
        const T v0 = in[3 * q], v1 = in[3 * q + 1], v2 = in[3 * q + 2];
        const T *s = S + (int64_t)nbi * 9 * nint + n;
#pragma unroll
        for (int r = 0; r < 3; r++)
          acc[r] += s[(3 * r) * nint] * v0 + s[(3 * r + 1) * nint] * v1 +
                    s[(3 * r + 2) * nint] * v2;
      }
#pragma unroll
  for (int r = 0; r < 3; r++) {
    const T bb = MODE == kApply ? T(0) : b[o + r];
    const T dd = MODE == kJacobi ? invD[o + r] : T(0);
    out[o + r] = finish<T>(MODE, acc[r], in[o + r], bb, dd, omega, false);
  }
}

// Solver variant: one thread per (interior node, row r) with blockIdx.y = r,
// the same arithmetic per row as k_stencil; nodes outside the interior are
// not written (the solver's vectors of these levels stay zero there).
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_stencil_rows(
    int nx, int ny, int nz, int wy, int wz, const T *__restrict__ S,
    const T *__restrict__ in, T *__restrict__ out, const T *__restrict__ b,
    const T *__restrict__ invD, T omega) {
  const int64_t nint = (int64_t)nx * ny * nz;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (n >= nint) return;
  const int j = (int)(n % ny), k = (int)((n / ny) % nz), i = (int)(n / ((int64_t)ny * nz));
  const int64_t node = ((int64_t)(i + 1) * wz + (k + 1)) * wy + (j + 1);
  T acc = 0;
#pragma unroll
  for (int dx = -1; dx <= 1; dx++)
#pragma unroll
    for (int dz = -1; dz <= 1; dz++)
#pragma unroll
      for (int dy = -1; dy <= 1; dy++) {
        const int nbi = (dx + 1) * 9 + (dz + 1) * 3 + (dy + 1);
        const int64_t q = node + ((int64_t)dx * wz + dz) * wy + dy;
        const T *s = S + (int64_t)(nbi * 9 + 3 * r) * nint + n;
        acc += s[0] * in[3 * q] + s[nint] * in[3 * q + 1] + s[2 * nint] * in[3 * q + 2];
      }
  const int64_t o = 3 * node + r;
  const T bb = MODE == kApply ? T(0) : b[o];
  const T dd = MODE == kJacobi ? invD[o] : T(0);
  out[o] = finish<T>(MODE, acc, in[o], bb, dd, omega, false);
}

template <typename T>
void launch_stencil_rows(const DevGrid &g, int l, int mode, const T *S, const T *in, T *out,
                         const T *b, const T *invD, T omega) {
  const Level &L = g.lv[l];
  const int64_t nint = L.interior_nodes();
  dim3 bs(256), gs((unsigned)((nint + 255) / 256), 3);
  if (mode == kApply)
    hipLaunchKernelGGL((k_stencil_rows<T, kApply>), gs, bs, 0, stream(), L.nx, L.ny, L.nz,
                       L.wy, L.wz, S, in, out, b, invD, omega);
  else if (mode == kJacobi)
    hipLaunchKernelGGL((k_stencil_rows<T, kJacobi>), gs, bs, 0, stream(), L.nx, L.ny, L.nz,
                       L.wy, L.wz, S, in, out, b, invD, omega);
  else
    hipLaunchKernelGGL((k_stencil_rows<T, kResidual>), gs, bs, 0, stream(), L.nx, L.ny,
                       L.nz, L.wy, L.wz, S, in, out, b, invD, omega);
}

// Solver variant on the symmetric half layout (launch_node_stencil, half),
// per (interior node n, row r).  Row r of the block towards an
// upper neighbour nb > 13 is stored at n; towards a lower neighbour
// nb < 13 it is column r of the block that the neighbour n' stores towards
// n (its upper neighbour 26 - nb), read there when n' is an interior node
// (otherwise the block is zero).
// Addresses are a wave-uniform base (plane, neighbour offset) plus the
// thread's 32-bit byte offset, which maps onto scalar-base + vector-offset
// loads.  The host checks the byte sizes of Sh and `in` < 2^31.
template <typename T>
__device__ __forceinline__ T ld_at(const T *base, uint32_t byte_off) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off);
}

// Part of row r of (K in) at node n from the 9 neighbours in node plane
// i + DX: sum over (dz, dy) in the row-major order of the 27-point stencil.
template <typename T, int DX>
__device__ __forceinline__ T half_row_part(const T *__restrict__ Sh, const T *__restrict__ in,
                                           uint32_t nint, int nx, int ny, int nz, int wy, int wz,
                                           int i, int j, int k, int r, uint32_t sb,
                                           uint32_t ub) {
  T acc = 0;
#pragma unroll
  for (int dz = -1; dz <= 1; dz++)
#pragma unroll
    for (int dy = -1; dy <= 1; dy++) {
      const int nbi = (DX + 1) * 9 + (dz + 1) * 3 + (dy + 1);
      const T *uq = in + 3 * (int64_t)((DX * wz + dz) * wy + dy);  // neighbour q of the node
      T s0, s1, s2;
      if (nbi > 13) {
        const T *s = Sh + (int64_t)((nbi - 14) * 9 + 3 * r) * nint;
        s0 = ld_at(s, sb);
        s1 = ld_at(s + nint, sb);
        s2 = ld_at(s + 2 * (int64_t)nint, sb);
      } else if (nbi == 13) {  // symmetric centre: C[r][c] = C[min][max]
        const T *c = Sh + (int64_t)117 * nint;
        s0 = ld_at(c + (int64_t)half_centre_index(0, r) * nint, sb);
        s1 = ld_at(c + (int64_t)half_centre_index(r < 1 ? r : 1, r < 1 ? 1 : r) * nint, sb);
        s2 = ld_at(c + (int64_t)half_centre_index(r, 2) * nint, sb);
      } else {
        const bool inside = i + DX >= 0 && i + DX < nx && j + dy >= 0 && j + dy < ny &&
                            k + dz >= 0 && k + dz < nz;
        s0 = s1 = s2 = T(0);
        if (inside) {  // column r of the block neighbour n + dn stores towards n
          const T *s = Sh + (int64_t)((12 - nbi) * 9 + r) * nint + ((DX * nz + dz) * ny + dy);
          s0 = ld_at(s, sb);
          s1 = ld_at(s + 3 * (int64_t)nint, sb);
          s2 = ld_at(s + 6 * (int64_t)nint, sb);
        }
      }
      acc += s0 * ld_at(uq, ub) + s1 * ld_at(uq + 1, ub) + s2 * ld_at(uq + 2, ub);
    }
  return acc;
}

// Workgroups of 3 waves on 64 consecutive nodes and one row r: wave w sums
// the neighbours in node plane i + w - 1, and the partial sums meet in LDS
// (row = (P_-1 + P_0) + P_+1), which the middle wave finishes.  Three waves
// per (node, row) instead of one thread hold three times the loads in
// flight on these small, latency-bound levels.
// The device deals workgroups to its 8 XCDs in turn.  Workgroup w works on
// (node block, row) pair v = (w % 8) * per + w / 8 of the 3 * nblk pairs,
// so each XCD sweeps one contiguous eighth of the stencil, the same eighth
// in every sweep, and can keep it in its own L2 across sweeps.
template <typename T, int MODE>
__global__ __launch_bounds__(192) void k_stencil_half(
    int nx, int ny, int nz, int wy, int wz, const T *__restrict__ Sh,
    const T *__restrict__ in, T *__restrict__ out, const T *__restrict__ b,
    const T *__restrict__ invD, T omega, uint32_t nblk, uint32_t per) {
  __shared__ T part[2][64];
  const uint32_t v = (blockIdx.x % 8u) * per + blockIdx.x / 8u;
  if (v >= 3u * nblk) return;  // (uniform in the workgroup)
  const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  const uint32_t nint = (uint32_t)nx * ny * nz;
  const uint32_t n = (v / 3u) * 64u + (uint32_t)lane;
  const int r = (int)(v % 3u);
  const bool ok = n < nint;
  const uint32_t nyz = (uint32_t)ny * nz;
  const int i = ok ? (int)(n / nyz) : 0, k = ok ? (int)((n - i * nyz) / ny) : 0;
  const int j = ok ? (int)(n - i * nyz - k * ny) : 0;
  const uint32_t node = ((uint32_t)(i + 1) * wz + (k + 1)) * wy + (j + 1);
  const uint32_t sb = (uint32_t)sizeof(T) * n;         // byte offset in a stencil plane
  const uint32_t ub = (uint32_t)sizeof(T) * 3 * node;  // byte offset of the node's values
  T acc = 0;
  if (ok) {  // (w is uniform in the wave)
    if (w == 0)
      acc = half_row_part<T, -1>(Sh, in, nint, nx, ny, nz, wy, wz, i, j, k, r, sb, ub);
    else if (w == 1)
      acc = half_row_part<T, 0>(Sh, in, nint, nx, ny, nz, wy, wz, i, j, k, r, sb, ub);
    else
      acc = half_row_part<T, 1>(Sh, in, nint, nx, ny, nz, wy, wz, i, j, k, r, sb, ub);
  }
  if (w != 1) part[w >> 1][lane] = acc;
  __syncthreads();
  if (w != 1 || !ok) return;
  acc = (part[0][lane] + acc) + part[1][lane];
  const uint32_t o = 3 * node + r;
  const T bb = MODE == kApply ? T(0) : b[o];
  const T dd = MODE == kJacobi ? invD[o] : T(0);
  out[o] = finish<T>(MODE, acc, in[o], bb, dd, omega, false);
}

template <typename T>
void launch_stencil_half(const DevGrid &g, int l, int mode, const T *Sh, const T *in, T *out,
                         const T *b, const T *invD, T omega) {
  const Level &L = g.lv[l];
  const int64_t nint = L.interior_nodes();
  if ((int64_t)sizeof(T) * kHalfPlanes * nint >= (int64_t(1) << 31) ||
      (int64_t)sizeof(T) * 3 * L.nodes() >= (int64_t(1) << 31)) {
    set_error(TOP3D_ERR_ARG, "level %d too large for 32-bit stencil offsets", l);
    return;
  }
  const uint32_t nblk = (uint32_t)((nint + 63) / 64), per = (3 * nblk + 7) / 8;
  dim3 bs(192), gs(8 * per);
  if (mode == kApply)
    hipLaunchKernelGGL((k_stencil_half<T, kApply>), gs, bs, 0, stream(), L.nx, L.ny, L.nz,
                       L.wy, L.wz, Sh, in, out, b, invD, omega, nblk, per);
  else if (mode == kJacobi)
    hipLaunchKernelGGL((k_stencil_half<T, kJacobi>), gs, bs, 0, stream(), L.nx, L.ny, L.nz,
                       L.wy, L.wz, Sh, in, out, b, invD, omega, nblk, per);
  else
    hipLaunchKernelGGL((k_stencil_half<T, kResidual>), gs, bs, 0, stream(), L.nx, L.ny,
                       L.nz, L.wy, L.wz, Sh, in, out, b, invD, omega, nblk, per);
}

template <typename T>
void launch_stencil(const DevGrid &g, int l, int mode, const T *S, const T *in,
                    T *out, const T *b, const T *invD, T omega) {
  const Level &L = g.lv[l];
  dim3 bs(64, 4, 1), gs((L.wy + 63) / 64, (L.wz + 3) / 4, L.wx);
  if (mode == kApply)
    hipLaunchKernelGGL((k_stencil<T, kApply>), gs, bs, 0, stream(), L.nx, L.ny, L.nz,
                       L.wy, L.wz, S, in, out, b, invD, omega);
  else if (mode == kJacobi)
    hipLaunchKernelGGL((k_stencil<T, kJacobi>), gs, bs, 0, stream(), L.nx, L.ny, L.nz,
                       L.wy, L.wz, S, in, out, b, invD, omega);
  else
    hipLaunchKernelGGL((k_stencil<T, kResidual>), gs, bs, 0, stream(), L.nx, L.ny,
                       L.nz, L.wy, L.wz, S, in, out, b, invD, omega);
}

// First pre-smoothing sweep from z = 0 (K 0 = 0): out = w invD b inside.
template <typename TB, typename T>
__global__ __launch_bounds__(256) void k_jacobi_zero(int64_t ndof, const TB *__restrict__ b,
                                                   const T *__restrict__ invD, T omega,
                                                   T *__restrict__ out) {
  // invD is zero on every non-real node of a level (k_invdiag_*), so the
  // sweep from zero needs no node classification
#pragma unroll 4
  for (int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; d < ndof;
       d += (int64_t)gridDim.x * blockDim.x)
    out[d] = (T)(omega * invD[d] * ((T)b[d] - T(0)));
}

template <typename TB, typename T>
void launch_jacobi_zero(const DevGrid &g, int l, const TB *b, const T *invD, T omega,
                        T *out) {
  const Level &L = g.lv[l];
  const int bs = 256;
  const int64_t blocks = std::min<int64_t>((L.ndof + bs - 1) / bs, 4096);
  hipLaunchKernelGGL((k_jacobi_zero<TB, T>), dim3(blocks), dim3(bs), 0, stream(), L.ndof, b,
                     invD, omega, out);
}

// ---------------------------------------------------------------------------
// Full-weighting restriction level l -> l+1 (stencil_methods.c:595-688):
// weights 1, 1/2, 1/4, 1/8 by squared distance over the 3x3x3 fine patch.
template <typename T>
__global__ void k_restrict(int cnx, int cny, int cnz, int cwy, int cwz, int fwy, int fwz,
                           const T *__restrict__ fine, T *__restrict__ coarse,
                           const T *__restrict__ invDc, T omega, T *__restrict__ zc) {
  // one thread per interior coarse node; only those are written
  // (stencil_methods.c:629-686).  With zc, also the coarse level's first
  // pre-smoothing sweep from zero, zc = omega invDc rc (as k_jacobi_zero; the
  // other nodes of zc stay zero)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)cnx * cny * cnz) return;
  const int j = (int)(t % cny) + 1, k = (int)((t / cny) % cnz) + 1,
            i = (int)(t / ((int64_t)cny * cnz)) + 1;
  const int64_t cn = ((int64_t)i * cwz + k) * cwy + j;
  const int fi = 2 * i - 1, fj = 2 * j - 1, fk = 2 * k - 1;
  T acc[3] = {0, 0, 0};
#pragma unroll
  for (int dx = -1; dx <= 1; dx++)
#pragma unroll
    for (int dz = -1; dz <= 1; dz++)
#pragma unroll
      for (int dy = -1; dy <= 1; dy++) {
        const int d2 = dx * dx + dy * dy + dz * dz;
        const T w = d2 == 0 ? T(1) : (d2 == 1 ? T(0.5) : (d2 == 2 ? T(0.25) : T(0.125)));
        const int64_t fnode = ((int64_t)(fi + dx) * fwz + (fk + dz)) * fwy + (fj + dy);
#pragma unroll
        for (int c = 0; c < 3; c++) acc[c] += w * fine[3 * fnode + c];
      }
  coarse[3 * cn] = acc[0];
  coarse[3 * cn + 1] = acc[1];
  coarse[3 * cn + 2] = acc[2];
  if (zc) {
#pragma unroll
    for (int c = 0; c < 3; c++) zc[3 * cn + c] = omega * invDc[3 * cn + c] * (acc[c] - T(0));
  }
}

template <typename T>
void launch_restrict(const DevGrid &g, int l, const T *fine, T *coarse, const T *invDc,
                     T omega, T *zc) {
  const Level &C = g.lv[l + 1], &F = g.lv[l];
  const int64_t nth = C.interior_nodes();
  hipLaunchKernelGGL((k_restrict<T>), dim3((nth + 255) / 256), dim3(256), 0, stream(), C.nx,
                     C.ny, C.nz, C.wy, C.wz, F.wy, F.wz, fine, coarse, invDc, omega, zc);
}

// Trilinear prolongation level l+1 -> l (stencil_methods.c:521-590);
// ADD = true fuses the correction z += P zc of the V-cycle.
template <typename T, bool ADD>
__global__ void k_prolong(int cnx, int cny, int cnz, int cwy, int cwz, int fnx, int fny, int fnz,
                          int fwy, int fwz, const T *__restrict__ coarse, T *__restrict__ fine) {
  // thread per coarse node (I, J, K): the fine nodes 2I-1, 2I (and the same
  // in y, z) take their trilinear weights from the 8 coarse nodes I..I+1,
  // J..J+1, K..K+1, loaded once; per fine node the sum runs over the same 8
  // terms in the same order as stencil_methods.c:546-589
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)cnx * cny * cnz) return;
  const int J = (int)(t % cny) + 1, K = (int)((t / cny) % cnz) + 1,
            I = (int)(t / ((int64_t)cny * cnz)) + 1;
  const int64_t csx = (int64_t)cwy * cwz;
  T cv[2][2][2][3];  // [di][dk][dj][comp]
#pragma unroll
  for (int di = 0; di < 2; di++)
#pragma unroll
    for (int dk = 0; dk < 2; dk++)
#pragma unroll
      for (int dj = 0; dj < 2; dj++) {
        const T *pc = coarse + 3 * ((I + di) * csx + (int64_t)(K + dk) * cwy + (J + dj));
#pragma unroll
        for (int c = 0; c < 3; c++) cv[di][dk][dj][c] = pc[c];
      }
  const int64_t fsx = (int64_t)fwy * fwz;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int cc = 0; cc < 2; cc++)
#pragma unroll
      for (int bb = 0; bb < 2; bb++) {
        const int i = 2 * I - 1 + a, k = 2 * K - 1 + cc, j = 2 * J - 1 + bb;
        if (i > fnx || j > fny || k > fnz) continue;
        // i1 = (i-1)/2+1, i2 = i/2+1 relative to I: offsets 0 or a
        const int i1 = 0, i2 = a, j1 = 0, j2 = bb, k1 = 0, k2 = cc;
        const int q[8][3] = {{i1, k1, j2}, {i2, k1, j2}, {i2, k1, j1}, {i1, k1, j1},
                             {i1, k2, j2}, {i2, k2, j2}, {i2, k2, j1}, {i1, k2, j1}};
        T *pf = fine + 3 * (i * fsx + (int64_t)k * fwy + j);
#pragma unroll
        for (int d = 0; d < 3; d++) {
          T s = 0;
#pragma unroll
          for (int t = 0; t < 8; t++) s += T(0.125) * cv[q[t][0]][q[t][1]][q[t][2]][d];
          if (ADD) pf[d] += s; else pf[d] = s;
        }
      }
}

template <typename T, bool ADD>
static void launch_prolong_t(const DevGrid &g, int l, const T *coarse, T *fine) {
  const Level &C = g.lv[l + 1], &F = g.lv[l];
  const int64_t nth = C.interior_nodes();
  hipLaunchKernelGGL((k_prolong<T, ADD>), dim3((nth + 255) / 256), dim3(256), 0, stream(), C.nx,
                     C.ny, C.nz, C.wy, C.wz, F.nx, F.ny, F.nz, F.wy, F.wz, coarse, fine);
}
template <typename T>
void launch_prolong_add(const DevGrid &g, int l, const T *coarse, T *fine) {
  launch_prolong_t<T, true>(g, l, coarse, fine);
}
template <typename T>
void launch_prolong(const DevGrid &g, int l, const T *coarse, T *fine) {
  launch_prolong_t<T, false>(g, l, coarse, fine);
}

// ---------------------------------------------------------------------------
// Inverse diagonal of a matrix-free level (stencil_methods.c:693-773) as a
// node-centric gather: sum over the level elements around the node and their
// sub-cells of E_f * KE_l[cell](d,d); fixed DOFs -> 1; interior inverted; the
// halo is 0.  KEdiag holds the 24 diagonal entries of each sub-cell matrix.
template <typename T>
__global__ void k_invdiag_mf(LevelArgs a, int ncell, const double *__restrict__ KEdiag,
                             const double *__restrict__ E, const uint8_t *__restrict__ fix,
                             T *__restrict__ invD) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y * blockDim.y + threadIdx.y;
  const int i = blockIdx.z;
  if (j >= a.wy || k >= a.wz) return;
  const int64_t node = ((int64_t)i * a.wz + k) * a.wy + j;
  if (!(i >= 1 && i <= a.nelx + 1 && j >= 1 && j <= a.nely + 1 && k >= 1 &&
        k <= a.nelz + 1)) {
    invD[3 * node] = invD[3 * node + 1] = invD[3 * node + 2] = T(0);
    return;
  }
  double d[3] = {0, 0, 0};
  for (int e = 0; e < 8; e++) {
    const int ex = (e & 1) ? 0 : -1, ey = (e & 2) ? 0 : -1, ez = (e & 4) ? 0 : -1;
    const int ie = i + ex, je = j + ey, ke = k + ez;
    if (ie < 1 || ie > a.nelx || je < 1 || je > a.nely || ke < 1 || ke > a.nelz) continue;
    const int loc = lnode_of(-ex, -ey, -ez);
    for (int ii = 0; ii < ncell; ii++)
      for (int kk = 0; kk < ncell; kk++)
        for (int jj = 0; jj < ncell; jj++) {
          const int64_t fe =
              ((int64_t)((ie - 1) * ncell + ii + 1) * a.fez + ((ke - 1) * ncell + kk + 1)) *
                  a.few + ((je - 1) * ncell + jj + 1);
          const double Ef = E[fe];
          const double *kd = KEdiag + 24 * ((ii * ncell + kk) * ncell + jj) + 3 * loc;
          d[0] += Ef * kd[0];
          d[1] += Ef * kd[1];
          d[2] += Ef * kd[2];
        }
  }
  const uint8_t m = fix[node];
  for (int c = 0; c < 3; c++) invD[3 * node + c] = (T)(1.0 / (((m >> c) & 1) ? 1.0 : d[c]));
}

// diagonal entries of the sub-cell matrices of level l (24 per sub-cell) at
// the reference's nu = 0.3, for levels whose table the grid does not keep;
// built once per level and kept for the life of the process
static double *kediag_for(int l) {
  static double *cache[kMaxLevels] = {};
  if (cache[l]) return cache[l];
  const int n = 1 << l;
  std::vector<double> KE(576 * (size_t)n * n * n), dg(24 * (size_t)n * n * n);
  element_matrices(0.3, l, KE.data());
  for (int c = 0; c < n * n * n; c++)
    for (int r = 0; r < 24; r++) dg[24 * c + r] = KE[576 * (size_t)c + 25 * r];
  // dalloc clears the block on the library stream, so the copy goes on that
  // stream too (after the clear) and completes before dg goes away
  double *d = (double *)dalloc(sizeof(double) * dg.size());
  if (d && T3D_HIP(hipMemcpyAsync(d, dg.data(), sizeof(double) * dg.size(),
                                  hipMemcpyHostToDevice, stream())) &&
      T3D_HIP(hipStreamSynchronize(stream())))
    cache[l] = d;
  return cache[l];
}

template <typename T>
void launch_invdiag_mf(const DevGrid &g, int l, const double *E, T *invD) {
  const Level &L = g.lv[l];
  const double *kd = g.KEdiag[l] ? g.KEdiag[l] : kediag_for(l);
  dim3 bs(64, 4, 1), gs((L.wy + 63) / 64, (L.wz + 3) / 4, L.wx);
  hipLaunchKernelGGL((k_invdiag_mf<T>), gs, bs, 0, stream(), g.args(l), L.ncell, kd, E,
                     g.fixmask[l], invD);
}

// inverse diagonal of an assembled level from its stencil (centre block)
template <typename T>
__global__ void k_invdiag_stencil(int nx, int ny, int nz, int wy, int wz,
                                  const double *__restrict__ S, T *__restrict__ invD) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y * blockDim.y + threadIdx.y;
  const int i = blockIdx.z;
  if (j >= wy || k >= wz) return;
  const int64_t node = ((int64_t)i * wz + k) * wy + j;
  if (!(i >= 1 && i <= nx && j >= 1 && j <= ny && k >= 1 && k <= nz)) {
    invD[3 * node] = invD[3 * node + 1] = invD[3 * node + 2] = T(0);
    return;
  }
  const int64_t nint = (int64_t)nx * ny * nz;
  const int64_t n = ((int64_t)(i - 1) * nz + (k - 1)) * ny + (j - 1);
  const int centre = 13;
  for (int c = 0; c < 3; c++)
    invD[3 * node + c] = (T)(1.0 / S[(int64_t)(centre * 9 + 4 * c) * nint + n]);
}

// inverse diagonal from a half-layout stencil (centre entries (c, c))
template <typename T>
__global__ void k_invdiag_half(int nx, int ny, int nz, int wy, int wz,
                               const double *__restrict__ Sh, T *__restrict__ invD) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y * blockDim.y + threadIdx.y;
  const int i = blockIdx.z;
  if (j >= wy || k >= wz) return;
  const int64_t node = ((int64_t)i * wz + k) * wy + j;
  if (!(i >= 1 && i <= nx && j >= 1 && j <= ny && k >= 1 && k <= nz)) {
    invD[3 * node] = invD[3 * node + 1] = invD[3 * node + 2] = T(0);
    return;
  }
  const int64_t nint = (int64_t)nx * ny * nz;
  const int64_t n = ((int64_t)(i - 1) * nz + (k - 1)) * ny + (j - 1);
  for (int c = 0; c < 3; c++)
    invD[3 * node + c] = (T)(1.0 / Sh[(int64_t)(117 + half_centre_index(c, c)) * nint + n]);
}

template <typename T>
void launch_invdiag_half(const DevGrid &g, int l, const double *Sh, T *invD) {
  const Level &L = g.lv[l];
  dim3 bs(64, 4, 1), gs((L.wy + 63) / 64, (L.wz + 3) / 4, L.wx);
  hipLaunchKernelGGL((k_invdiag_half<T>), gs, bs, 0, stream(), L.nx, L.ny, L.nz, L.wy, L.wz,
                     Sh, invD);
}

template <typename T>
void launch_invdiag_stencil(const DevGrid &g, int l, const double *S, T *invD) {
  const Level &L = g.lv[l];
  dim3 bs(64, 4, 1), gs((L.wy + 63) / 64, (L.wz + 3) / 4, L.wx);
  hipLaunchKernelGGL((k_invdiag_stencil<T>), gs, bs, 0, stream(), L.nx, L.ny, L.nz, L.wy,
                     L.wz, S, invD);
}

// ---------------------------------------------------------------------------
// explicit instantiations
#define T3D_INST(T)                                                                      \
  template void launch_young<T>(const DevGrid &, const double *, T *);                   \
  template void launch_young_float_input<T>(const DevGrid &, const float *, T *);        \
  template void launch_fine<T>(const DevGrid &, int, const T *, const T *, T *,          \
                               const T *, const T *, T, int);                            \
  template void launch_level1<T>(const DevGrid &, int, const T *, const T *, T *,        \
                                 const T *, const T *, T, int, const T *);               \
  template void launch_stencil<T>(const DevGrid &, int, int, const T *, const T *, T *,  \
                                  const T *, const T *, T);                              \
  template void launch_stencil_rows<T>(const DevGrid &, int, int, const T *, const T *,  \
                                       T *, const T *, const T *, T);                    \
  template void launch_stencil_half<T>(const DevGrid &, int, int, const T *, const T *,  \
                                       T *, const T *, const T *, T);                    \
  template void launch_invdiag_half<T>(const DevGrid &, int, const double *, T *);       \
  template void launch_restrict<T>(const DevGrid &, int, const T *, T *, const T *, T, T *); \
  template void launch_prolong_add<T>(const DevGrid &, int, const T *, T *);             \
  template void launch_prolong<T>(const DevGrid &, int, const T *, T *);                 \
  template void launch_invdiag_mf<T>(const DevGrid &, int, const double *, T *);         \
  template void launch_invdiag_stencil<T>(const DevGrid &, int, const double *, T *);
T3D_INST(double)
T3D_INST(float)
template void launch_jacobi_zero<double, double>(const DevGrid &, int, const double *,
                                                 const double *, double, double *);
template void launch_jacobi_zero<float, float>(const DevGrid &, int, const float *,
                                               const float *, float, float *);
template void launch_jacobi_zero<double, float>(const DevGrid &, int, const double *,
                                                const float *, float, float *);

}  // namespace t3d
