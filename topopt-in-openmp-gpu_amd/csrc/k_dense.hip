// Dense FP64 direct solver of the coarsest level (replaces CHOLMOD,
// stencil_assembly.c:366-503).
//
// The coarsest matrix (n = 3 * interior nodes, 4131 at 16x8x8 elements) is
// block tridiagonal over the grid's x-planes (3 ny nz DOFs each, x slowest).
// Per design iteration:
//   1. Cholesky P A P^T = L L^T over 64x64 tiles, with the planes in
//      nested-dissection order (odd-even reduction of the plane chain, each
//      plane padded to whole tiles): every launch eliminates one tile column
//      of every plane of the current independent set, so a 17-plane chain
//      takes 5 levels x 4 columns (+1) launches instead of the 65 of the
//      banded order.  Each workgroup owns one target tile of the update,
//      forms the panel tiles Q = A_XK L_KK^{-T} it needs on FP64 MFMA
//      (v_mfma_f64_16x16x4f64) and applies C_IJ -= sum_K Q_IK Q_JK^T; the
//      workgroup of a diagonal tile eliminated next factors it in LDS right
//      after its update (Cholesky, then the inverse of the 64x64 factor);
//   2. W = L^{-1} (lower triangle, zero outside the reach of each column),
//      one workgroup per 16-column strip visiting only the tile rows where
//      its column of W is non-zero; W^T is written alongside.
// A V-cycle applies A^{-1} = P^T W^T W P as two triangular GEMVs.
#include <algorithm>

#include "t3d_internal.hpp"

namespace t3d {

constexpr int kNB = kTile;      // tile size
constexpr int kLd = kNB + 1;    // LDS row stride of a 64x64 tile
constexpr int kStrip = 16;      // columns per workgroup of the W = L^{-1} kernel

typedef double double4_t __attribute__((ext_vector_type(4)));

// MFMA operand layout (v_mfma_f64_16x16x4f64, measured on gfx950): lane l
// feeds A[l%16][l/16] and B[l/16][l%16]; accumulator register v of lane l
// holds D[l/16 + 4v][l%16].
__device__ inline double4_t mfma4(double a, double b, double4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// acc += X Y^T on the wave's 32x32 quadrant (rows qr.., cols qc..) of a
// 64x64 product; X, Y are 64x64 tiles in LDS with row stride kLd.
__device__ inline void mm_nt_tile(const double *X, const double *Y, int qr, int qc, int lane,
                                  double4_t acc[2][2]) {
  const int li = lane & 15, lk = lane >> 4;
  for (int k = 0; k < kNB; k += 4) {
    double av[2], bv[2];
#pragma unroll
    for (int x = 0; x < 2; x++) av[x] = X[(qr + 16 * x + li) * kLd + k + lk];
#pragma unroll
    for (int y = 0; y < 2; y++) bv[y] = Y[(qc + 16 * y + li) * kLd + k + lk];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 2; y++) acc[x][y] = mfma4(av[x], bv[y], acc[x][y]);
  }
}

__device__ inline void zero_acc(double4_t acc[2][2]) {
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) acc[x][y] = double4_t{0.0, 0.0, 0.0, 0.0};
}

// Store the wave's quadrant of an accumulator into a 64x64 LDS tile.
__device__ inline void store_acc(double *X, const double4_t acc[2][2], int qr, int qc, int lane) {
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) X[(qr + 16 * x + lk + 4 * v) * kLd + qc + 16 * y + li] = acc[x][y][v];
}

// ---------------------------------------------------------------------------
// Cholesky and inverse of one 64x64 SPD tile in LDS, 256 threads (4 waves),
// blocked by 16: per diagonal block kb
//   (a) wave 0, lanes 0..15 (one row each; each pivot column broadcast
//       through LDS): the 16x16 Cholesky L_D (reciprocal square roots
//       refined by Newton steps) and its inverse X_D by forward substitution;
//   (b) the panel below: P <- P X_D^T on MFMA;
//   (c) the trailing lower triangle: C -= P P^T as 16x16 MFMA tiles;
// then the off-diagonal 16x16 blocks of the inverse by distance on MFMA.
// In: T (lower triangle; rows/cols beyond the matrix padded with identity).
// Out: T = L (zeros above the diagonal), X = L^{-1}.  A non-positive pivot
// sets *info (and is replaced by 1 so that the factorization finishes).
// Value of v in lane `lane` (a wave-uniform index), for every lane.
__device__ inline double readlane_d(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 1/sqrt(x) for x > 0: the hardware estimate and two Newton steps
// y <- y + y (1 - x y^2) / 2
__device__ inline double rsqrt_refined(double x) {
  double y = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int it = 0; it < 2; it++) {
    const double h = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, h, y);
  }
  return y;
}

__device__ __forceinline__ void factor_tile(double *T, double *X, int *info) {
  __shared__ double col[16];  // the current pivot column of a 16x16 block
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int li = lane & 15, lk = lane >> 4;
  for (int e = t; e < kNB * kNB; e += 256) X[(e >> 6) * kLd + (e & 63)] = 0.0;
  __syncthreads();
  for (int kb = 0; kb < 4; kb++) {
    const int d0 = 16 * kb;
    if (t < 16) {  // (a): one row per lane
      const int r = t;
      // the row's entries right of the diagonal are updated as well and
      // dropped when the row is stored: a lane only ever reads the column
      // values of lanes below the current pivot, which are exact
      double v[16], inv[16];
#pragma unroll
      for (int c = 0; c < 16; c++) v[c] = T[(d0 + r) * kLd + d0 + c];
#pragma unroll
      for (int j = 0; j < 16; j++) {
        double piv = readlane_d(v[j], j);
        if (!(piv > 0.0)) {
          if (r == 0) info[0] = 1;
          piv = 1.0;
        }
        const double ij = rsqrt_refined(piv);
        inv[j] = ij;  // 1 / L_jj
        const double lrj = (r == j) ? piv * ij : v[j] * ij;
        v[j] = lrj;
        if (j < 15) {
          // column j to every lane through LDS (the wave's LDS accesses
          // complete in order: each read sees this write, and the next
          // pivot's write follows these reads)
          col[r] = lrj;
#pragma unroll
          for (int c = j + 1; c < 16; c++) v[c] -= lrj * col[c];
        }
      }
#pragma unroll
      for (int c = 0; c < 16; c++) T[(d0 + r) * kLd + d0 + c] = (c <= r) ? v[c] : 0.0;
      // column r of X_D = L_D^{-1}: forward substitution with the rows of
      // L_D read back from LDS (the same address for every lane), two
      // partial sums to halve the dependent chain
      double x[16];
#pragma unroll
      for (int i = 0; i < 16; i++) {
        double a0 = (i == r) ? 1.0 : 0.0, a1 = 0.0;
#pragma unroll
        for (int m = 0; m < i; m += 2) {
          a0 -= T[(d0 + i) * kLd + d0 + m] * x[m];
          if (m + 1 < i) a1 -= T[(d0 + i) * kLd + d0 + m + 1] * x[m + 1];
        }
        x[i] = (a0 + a1) * inv[i];
        X[(d0 + i) * kLd + d0 + r] = x[i];
      }
    }
    __syncthreads();
    const int npan = kNB - d0 - 16;  // panel rows below the diagonal block
    if (npan > 0) {
      // (b) P[p][c] <- sum_m P[p][m] X_D[c][m], one 16-row block per wave
      const int p0 = d0 + 16 + 16 * w;
      double4_t pacc = {0.0, 0.0, 0.0, 0.0};
      if (16 * w < npan) {
#pragma unroll
        for (int k = 0; k < 16; k += 4)
          pacc = mfma4(T[(p0 + li) * kLd + d0 + k + lk], X[(d0 + li) * kLd + d0 + k + lk], pacc);
      }
      __syncthreads();
      if (16 * w < npan) {
#pragma unroll
        for (int v = 0; v < 4; v++) T[(p0 + lk + 4 * v) * kLd + d0 + li] = pacc[v];
      }
      __syncthreads();
      // (c) trailing tiles (ti, tj), kb < tj <= ti < 4, round-robin over waves
      const int nt = 3 - kb, ntiles = nt * (nt + 1) / 2;
      for (int q = w; q < ntiles; q += 4) {
        int ti = 0, tj = 0, cnt = 0;
        for (int a = 0; a < nt; a++)
          for (int b2 = 0; b2 <= a; b2++) {
            if (cnt == q) {
              ti = kb + 1 + a;
              tj = kb + 1 + b2;
            }
            cnt++;
          }
        double4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < 16; k += 4)
          acc = mfma4(T[(16 * ti + li) * kLd + d0 + k + lk], T[(16 * tj + li) * kLd + d0 + k + lk],
                      acc);
#pragma unroll
        for (int v = 0; v < 4; v++) T[(16 * ti + lk + 4 * v) * kLd + 16 * tj + li] -= acc[v];
      }
      __syncthreads();
    }
  }
  for (int e = t; e < kNB * kNB; e += 256) {
    const int r = e >> 6, c = e & 63;
    if (c > r) T[r * kLd + c] = 0.0;
  }
  __syncthreads();
  // off-diagonal blocks of the inverse by distance d, one block per wave:
  //   Y = sum_{K=J}^{I-1} L_IK X_KJ,  X_IJ = -X_II Y
  // (the accumulator of Y is the B operand of the second product as is:
  // register v of lane l holds Y[l/16 + 4v][l%16], the B element of k-step v)
  for (int d = 1; d < 4; d++) {
    if (w < 4 - d) {
      const int J = w, I = d + w;
      double4_t y = {0.0, 0.0, 0.0, 0.0};
      for (int K = J; K < I; K++)
#pragma unroll
        for (int k = 0; k < 16; k += 4)
          y = mfma4(T[(16 * I + li) * kLd + 16 * K + k + lk], X[(16 * K + k + lk) * kLd + 16 * J + li],
                    y);
      double4_t o = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int v = 0; v < 4; v++)
        o = mfma4(X[(16 * I + li) * kLd + 16 * I + 4 * v + lk], y[v], o);
#pragma unroll
      for (int v = 0; v < 4; v++) X[(16 * I + lk + 4 * v) * kLd + 16 * J + li] = -o[v];
    }
    __syncthreads();
  }
}

// Write the factored diagonal tile (blk, blk): L to Lf, L^{-1} to Dinv.
__device__ inline void store_diag(const double *T, const double *X, double *__restrict__ Lf,
                                  double *__restrict__ Dinv, int np, int blk) {
  const int64_t k0 = (int64_t)blk * kNB;
  for (int e = threadIdx.x; e < kNB * kNB; e += 256) {
    const int r = e >> 6, c = e & 63;
    Lf[(k0 + r) * np + k0 + c] = T[r * kLd + c];
    Dinv[(int64_t)blk * kNB * kNB + e] = X[r * kLd + c];
  }
}

// One task of an elimination step (one workgroup): the trailing tile (I, J),
// I >= J, receives the updates of the tile columns K[0..nsrc) eliminated in
// this step,
//   C_IJ = A_IJ - sum_K Q_IK Q_JK^T,   Q_XK = A_XK L_KK^{-T}
// (each workgroup forms the two panel tiles it needs on MFMA).  The task of a
// diagonal tile (I, I) also writes L_IK = Q_IK.  With kTaskFactor the tile is
// complete after this update and is factored in LDS right away (Cholesky,
// then the inverse of the 64x64 factor); nsrc = 0 factors A_II as it is.
// Matrices are np x np row-major, np a multiple of 64.
__global__ __launch_bounds__(256) void k_chol_task(double *__restrict__ A, double *__restrict__ Lf,
                                                   double *__restrict__ Dinv, int np,
                                                   const CholTask *__restrict__ tasks,
                                                   int *__restrict__ info) {
  __shared__ double PI[kNB * kLd];
  __shared__ double PJ[kNB * kLd];
  __shared__ double DV[kNB * kLd];
  __shared__ double CT[kNB * kLd];
  const CholTask tk = tasks[blockIdx.x];
  const int tid = threadIdx.x;
  const bool diag = tk.I == tk.J;
  const int64_t gi = (int64_t)tk.I * kNB, gj = (int64_t)tk.J * kNB;
  if (tk.nsrc == 0) {
    double rv[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int e = tid + 256 * q;
      rv[q] = A[(gi + (e >> 6)) * np + gi + (e & 63)];
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int e = tid + 256 * q;
      DV[(e >> 6) * kLd + (e & 63)] = rv[q];
    }
  } else {
    const int w = tid >> 6, lane = tid & 63;
    const int qr = (w >> 1) * 32, qc = (w & 1) * 32;
    double4_t cacc[2][2];
    zero_acc(cacc);
    for (int s = 0; s < tk.nsrc; s++) {
      const int64_t gk = (int64_t)tk.K[s] * kNB;
      // stage A_IK, A_JK, L_KK^{-1} (and C_IJ with the first source); all
      // loads of a thread are issued before its LDS stores
      double ri[16], rj[16], rd[16], rc[16];
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int e = tid + 256 * q, r = e >> 6, c = e & 63;
        ri[q] = A[(gi + r) * np + gk + c];
        rj[q] = diag ? 0.0 : A[(gj + r) * np + gk + c];
        rd[q] = Dinv[(int64_t)tk.K[s] * kNB * kNB + e];
        rc[q] = s == 0 ? A[(gi + r) * np + gj + c] : 0.0;
      }
      __syncthreads();  // the previous source's products are done with the tiles
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int e = tid + 256 * q, o = (e >> 6) * kLd + (e & 63);
        PI[o] = ri[q];
        if (!diag) PJ[o] = rj[q];
        DV[o] = rd[q];
        if (s == 0) CT[o] = rc[q];
      }
      __syncthreads();
      double4_t qi[2][2], qj[2][2];
      zero_acc(qi);
      zero_acc(qj);
      mm_nt_tile(PI, DV, qr, qc, lane, qi);  // Q_I = A_IK L_KK^{-T}
      if (!diag) mm_nt_tile(PJ, DV, qr, qc, lane, qj);
      __syncthreads();
      store_acc(PI, qi, qr, qc, lane);
      if (!diag) store_acc(PJ, qj, qr, qc, lane);
      __syncthreads();
      if (diag) {
#pragma unroll 4
        for (int e = tid; e < kNB * kNB; e += 256) {
          const int r = e >> 6, c = e & 63;
          Lf[(gi + r) * np + gk + c] = PI[r * kLd + c];
        }
      }
      mm_nt_tile(PI, diag ? PI : PJ, qr, qc, lane, cacc);  // += Q_I Q_J^T
    }
    const bool fac = (tk.flags & kTaskFactor) != 0;
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 2; y++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int r = qr + 16 * x + lk + 4 * v, c = qc + 16 * y + li;
          const double a = CT[r * kLd + c] - cacc[x][y][v];
          if (fac) DV[r * kLd + c] = a;
          else A[(gi + r) * np + gj + c] = a;
        }
    if (!fac) return;
  }
  __syncthreads();
  factor_tile(DV, PJ, info);
  store_diag(DV, PJ, Lf, Dinv, np, tk.I);
}

// ---------------------------------------------------------------------------
// W = L^{-1}, WT = W^T by 16-column strips.  The strip [16s, 16s+16) of tile
// column J starts with W_JJ = L_JJ^{-1} and visits the tile rows where column
// J of W is non-zero, in increasing order:
//   W_I = -L_II^{-1} sum_{K in ks(I)} L_IK W_K,
// ks(I) the strip's earlier rows with L_IK != 0 (row and K lists from the
// elimination schedule).  Two groups of 4 waves take alternate K of a row,
// each with its own double-buffered LDS stage for the L_IK tiles (the next
// tile in flight during the current product); their 64x16 partial sums meet
// in LDS.  W_K is read back from W (written by this workgroup
// before a barrier).  W is always written; WT, W32 and WT32 (FP32 copies)
// when given.
constexpr int kLinvGroups = 2;

__global__ __launch_bounds__(256 * kLinvGroups) void k_linv_seq(
    const double *__restrict__ Lf, const double *__restrict__ Dinv, int np,
    const int2 *__restrict__ colinfo, const int4 *__restrict__ rows,
    const int *__restrict__ klist, double *__restrict__ W, double *__restrict__ WT,
    float *__restrict__ W32, float *__restrict__ WT32) {
  __shared__ double Bx[kLinvGroups][kNB * 17];
  __shared__ double Lt[kLinvGroups][2 * kNB * kLd];
  const int tid = threadIdx.x, grp = tid >> 8, gtid = tid & 255;
  const int w = gtid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int c0 = blockIdx.x * kStrip;
  const int J = c0 / kNB, cofs = c0 - J * kNB;
  const int gc = c0 + li;
  for (int e = tid; e < kNB * kStrip; e += 256 * kLinvGroups) {
    const int r = e >> 4, j = e & 15;
    const double v = Dinv[(int64_t)J * kNB * kNB + r * kNB + cofs + j];
    const int64_t gr = (int64_t)J * kNB + r;
    W[gr * np + c0 + j] = v;
    if (WT) WT[(int64_t)(c0 + j) * np + gr] = v;
    if (W32) {
      W32[gr * np + c0 + j] = (float)v;
      WT32[(int64_t)(c0 + j) * np + gr] = (float)v;
    }
  }
  auto load_tile = [&](int I, int K, double (&reg)[16]) {
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int e = gtid + 256 * q;
      reg[q] = Lf[((int64_t)I * kNB + (e >> 6)) * np + (int64_t)K * kNB + (e & 63)];
    }
  };
  auto store_tile = [&](double *dst, const double (&reg)[16]) {
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int e = gtid + 256 * q;
      dst[(e >> 6) * kLd + (e & 63)] = reg[q];
    }
  };
  // this lane's B elements of W_K: rows k + lk (k = 0, 4, .., 60), column gc
  auto load_b = [&](int K, double (&breg)[16]) {
    const double *wk = W + (int64_t)K * kNB * np + gc;
#pragma unroll
    for (int q = 0; q < 16; q++) breg[q] = wk[(int64_t)(4 * q + lk) * np];
  };
  double *Lg = Lt[grp];
  __syncthreads();
  const int2 ci = colinfo[J];
  for (int q = 0; q < ci.y; q++) {
    const int4 rw = rows[ci.x + q];  // (I, offset into klist, count, -)
    const int I = rw.x, nk = rw.z;
    double dreg[16];  // group 0: rows 16w.. of L_II^{-1}, columns k < 16(w+1)
    if (grp == 0) {
      const double *D = Dinv + (int64_t)I * kNB * kNB;
#pragma unroll
      for (int t = 0; t < 16; t++)
        dreg[t] = (4 * t < 16 * (w + 1)) ? D[(16 * w + li) * kNB + 4 * t + lk] : 0.0;
    }
    double treg[16], bcur[16];
    if (grp < nk) {  // this group's first K: t = grp
      load_tile(I, klist[rw.y + grp], treg);
      store_tile(Lg, treg);
    }
    __syncthreads();
    double4_t acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    const int iters = (nk + kLinvGroups - 1) / kLinvGroups;
    for (int it = 0; it < iters; it++) {
      const int t = kLinvGroups * it + grp;  // this group's K-term
      const bool has = t < nk, more = t + kLinvGroups < nk;
      const double *cur = Lg + (size_t)(it & 1) * kNB * kLd;
      double *nxt = Lg + (size_t)((it & 1) ^ 1) * kNB * kLd;
      if (more) load_tile(I, klist[rw.y + t + kLinvGroups], treg);  // in flight meanwhile
      if (has) {
        load_b(klist[rw.y + t], bcur);
#pragma unroll
        for (int qq = 0; qq < 16; qq += 2) {
          acc0 = mfma4(cur[(16 * w + li) * kLd + 4 * qq + lk], bcur[qq], acc0);
          acc1 = mfma4(cur[(16 * w + li) * kLd + 4 * qq + 4 + lk], bcur[qq + 1], acc1);
        }
      }
      if (more) store_tile(nxt, treg);
      __syncthreads();
    }
#pragma unroll
    for (int v = 0; v < 4; v++) Bx[grp][(16 * w + lk + 4 * v) * 17 + li] = acc0[v] + acc1[v];
    __syncthreads();
    if (grp == 0) {
      double4_t o = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < 16; t++)
        if (4 * t < 16 * (w + 1))
          o = mfma4(dreg[t], Bx[0][(4 * t + lk) * 17 + li] + Bx[1][(4 * t + lk) * 17 + li], o);
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int64_t gr = (int64_t)I * kNB + 16 * w + lk + 4 * v;
        const double val = -o[v];
        W[gr * np + gc] = val;
        if (WT) WT[(int64_t)gc * np + gr] = val;
        if (W32) {
          W32[gr * np + gc] = (float)val;
          WT32[(int64_t)gc * np + gr] = (float)val;
        }
      }
    }
    __syncthreads();
  }
}

// C = WT W restricted to m >= max(i, j) (the explicit inverse A^{-1}),
// 64x64 tiles, k staged through LDS in chunks of 16.  C is n x n; W and WT
// have row stride ld.
__global__ __launch_bounds__(256) void k_gemm_wtw(const double *__restrict__ WT,
                                                  const double *__restrict__ W, int n, int ld,
                                                  double *__restrict__ C) {
  __shared__ double As[kNB][17];
  __shared__ double Bs[16][kNB + 1];
  const int nblk = (n + kNB - 1) / kNB;
  const int bi = blockIdx.x / nblk, bj = blockIdx.x % nblk;
  const int i0 = bi * kNB, j0 = bj * kNB;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int qr = (w >> 1) * 32, qc = (w & 1) * 32;
  double4_t acc[2][2];
  zero_acc(acc);
  for (int m0 = std::max(i0, j0); m0 < n; m0 += 16) {
    for (int e = tid; e < kNB * 16; e += 256) {
      const int r = e >> 4, k = e & 15;
      As[r][k] = (i0 + r < n && m0 + k < n) ? WT[(int64_t)(i0 + r) * ld + m0 + k] : 0.0;
      const int kk = e >> 6, c = e & 63;
      Bs[kk][c] = (m0 + kk < n && j0 + c < n) ? W[(int64_t)(m0 + kk) * ld + j0 + c] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k += 4) {
      double av[2], bv[2];
#pragma unroll
      for (int x = 0; x < 2; x++) av[x] = As[qr + 16 * x + li][k + lk];
#pragma unroll
      for (int y = 0; y < 2; y++) bv[y] = Bs[k + lk][qc + 16 * y + li];
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) acc[x][y] = mfma4(av[x], bv[y], acc[x][y]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = i0 + qr + 16 * x + lk + 4 * v, c = j0 + qc + 16 * y + li;
        if (r < n && c < n) C[(int64_t)r * n + c] = acc[x][y][v];
      }
}

// ---------------------------------------------------------------------------
// host side
// A[s][s] = 1 on the padding indices (ipos[s] < 0)
__global__ void k_pad_identity(int np, const int *__restrict__ ipos, double *__restrict__ A) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < np && ipos[s] < 0) A[(int64_t)s * np + s] = 1.0;
}

bool CoarseFactor::allocate(int n_, int bw_, bool f32, int planes) {
  CoarseSchedule cs;
  if (!build_coarse_schedule(n_, bw_, planes, cs)) {
    set_error(TOP3D_ERR_ARG, "coarse schedule: %s", cs.error.c_str());
    return false;
  }
  n = n_;
  bw = std::max(0, std::min(bw_, n_ - 1));
  np = cs.np;
  nblk = cs.nblk;
  std::vector<int> hipos(np, -1);
  for (int d = 0; d < n; d++) hipos[cs.pos[d]] = d;
  const size_t nn = (size_t)np * np;
  A = (double *)dalloc(sizeof(double) * nn);
  L = (double *)dalloc(sizeof(double) * nn);
  W = (double *)dalloc(sizeof(double) * nn);
  WT = (double *)dalloc(sizeof(double) * nn);
  Dinv = (double *)dalloc(sizeof(double) * (size_t)nblk * kNB * kNB);
  info = (int *)dalloc(sizeof(int) * 4);
  cw64 = (double *)dalloc(sizeof(double) * 2 * (size_t)np);
  pos = (int *)dalloc(sizeof(int) * (size_t)std::max(n, 1));
  ipos = (int *)dalloc(sizeof(int) * (size_t)np);
  std::vector<int> zt;
  for (const CholTask &tk : cs.tasks) {
    zt.push_back(tk.I * cs.nblk + tk.J);
    for (int q = 0; q < tk.nsrc; q++) {
      zt.push_back(tk.I * cs.nblk + tk.K[q]);
      zt.push_back(tk.J * cs.nblk + tk.K[q]);
    }
  }
  std::sort(zt.begin(), zt.end());
  zt.erase(std::unique(zt.begin(), zt.end()), zt.end());
  nztiles = (int)zt.size();
  dztiles = (int *)dalloc(sizeof(int) * std::max<size_t>(1, zt.size()));
  ntask = (int)cs.tasks.size();
  dtasks = (CholTask *)dalloc(sizeof(CholTask) * std::max<size_t>(1, cs.tasks.size()));
  dcolinfo = (int2 *)dalloc(sizeof(int) * cs.colinfo.size());
  drows = (int4 *)dalloc(sizeof(int) * std::max<size_t>(4, cs.rows.size()));
  dklist = (int *)dalloc(sizeof(int) * std::max<size_t>(1, cs.klist.size()));
  if (f32) {
    W32 = (float *)dalloc(sizeof(float) * nn);
    WT32 = (float *)dalloc(sizeof(float) * nn);
    cw32 = (float *)dalloc(sizeof(float) * 2 * (size_t)np);
  }
  if (!(A && L && W && WT && Dinv && info && cw64 && pos && ipos && dtasks && dcolinfo &&
        drows && dklist && dztiles && (!f32 || (W32 && WT32 && cw32))))
    return false;
  launch_first = cs.launch_first;
  launch_count = cs.launch_count;
  hipStream_t s = stream();
  auto up = [&](void *dst, const void *src, size_t bytes) {
    return bytes == 0 || T3D_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  };
  return up(pos, cs.pos.data(), sizeof(int) * n) && up(ipos, hipos.data(), sizeof(int) * np) &&
         up(dztiles, zt.data(), sizeof(int) * zt.size()) &&
         up(dtasks, cs.tasks.data(), sizeof(CholTask) * cs.tasks.size()) &&
         up(dcolinfo, cs.colinfo.data(), sizeof(int) * cs.colinfo.size()) &&
         up(drows, cs.rows.data(), sizeof(int) * cs.rows.size()) &&
         up(dklist, cs.klist.data(), sizeof(int) * cs.klist.size()) &&
         T3D_HIP(hipStreamSynchronize(s));
}

void CoarseFactor::release() {
  for (void *p : {(void *)A, (void *)L, (void *)W, (void *)WT, (void *)Dinv, (void *)info,
                  (void *)cw64, (void *)W32, (void *)WT32, (void *)cw32, (void *)pos,
                  (void *)ipos, (void *)dtasks, (void *)dcolinfo, (void *)drows,
                  (void *)dklist, (void *)dztiles})
    if (p) dfree(p);
  dztiles = nullptr;
  nztiles = 0;
  A = L = W = WT = Dinv = cw64 = nullptr;
  W32 = WT32 = cw32 = nullptr;
  info = pos = ipos = dklist = nullptr;
  dtasks = nullptr;
  dcolinfo = nullptr;
  drows = nullptr;
}

// zero the listed 64x64 tiles of an np x np matrix, one workgroup per tile
__global__ __launch_bounds__(256) void k_zero_tiles(double *__restrict__ A, int np, int nblk,
                                                    const int *__restrict__ tiles) {
  const int tile = tiles[blockIdx.x];
  const int64_t r0 = (int64_t)(tile / nblk) * kNB, c0 = (int64_t)(tile % nblk) * kNB;
  for (int e = threadIdx.x; e < kNB * kNB; e += 256) A[(r0 + (e >> 6)) * np + c0 + (e & 63)] = 0.0;
}

// A = 0 on the tiles the factorization reads, identity on the padding
void CoarseFactor::begin_assembly() {
  if (nztiles > 0)
    hipLaunchKernelGGL(k_zero_tiles, dim3(nztiles), dim3(256), 0, stream(), A, np, nblk,
                       dztiles);
  hipLaunchKernelGGL(k_pad_identity, dim3((np + 255) / 256), dim3(256), 0, stream(), np, ipos,
                     A);
}

int CoarseFactor::factor(bool f32out) {
  factor_async(f32out);
  return check_info();
}

int CoarseFactor::check_info() {
  int hinfo = 0;
  hipStream_t s = stream();
  if (!T3D_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s)) ||
      !T3D_HIP(hipStreamSynchronize(s)))
    return TOP3D_ERR_HIP;
  if (hinfo) {
    set_error(TOP3D_ERR_NOT_SPD, "coarse Cholesky: non-positive pivot");
    return TOP3D_ERR_NOT_SPD;
  }
  return TOP3D_OK;
}

void CoarseFactor::factor_async(bool f32out) {
  hipStream_t s = stream();
  if (!T3D_HIP(hipMemsetAsync(info, 0, sizeof(int), s))) return;
  for (size_t q = 0; q < launch_first.size(); q++)
    hipLaunchKernelGGL(k_chol_task, dim3(launch_count[q]), dim3(256), 0, s, A, L, Dinv, np,
                       dtasks + launch_first[q], info);
  const bool f32 = f32out && W32;  // FP32 W, W^T for the mixed V-cycle, else FP64 W^T
  hipLaunchKernelGGL(k_linv_seq, dim3(np / kStrip), dim3(256 * kLinvGroups), 0, s, L, Dinv, np,
                     dcolinfo,
                     drows, dklist, W, f32 ? nullptr : WT, f32 ? W32 : nullptr,
                     f32 ? WT32 : nullptr);
}

// ---------------------------------------------------------------------------
// Coarsest solve z = W^T (W r) in the factor's storage order: gather the
// interior DOFs to their storage positions, lower-triangular GEMV,
// upper-triangular GEMV scattered back to the halo layout (the halo of z
// stays zero).  One workgroup per row, FP64 accumulation.
template <typename T>
__global__ void k_gather_interior(int nx, int ny, int nz, int wy, int wz,
                                  const int *__restrict__ pos, const T *__restrict__ in,
                                  T *__restrict__ out) {
  const int64_t nint = (int64_t)nx * ny * nz;
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= 3 * nint) return;
  const int64_t p = d / 3;
  const int j = (int)(p % ny), k = (int)((p / ny) % nz), i = (int)(p / ((int64_t)ny * nz));
  out[pos[d]] = in[3 * (((int64_t)(i + 1) * wz + (k + 1)) * wy + (j + 1)) + d % 3];
}

// One 256-thread workgroup per row: the row (up to n values) is read by all
// four waves at once, FP64 partial sums, fixed-order block reduction.
__device__ __forceinline__ double row_block_sum(double s) {
  __shared__ double sm[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
  return (sm[0] + sm[1]) + (sm[2] + sm[3]);  // meaningful in thread 0
}

template <typename T>
__global__ __launch_bounds__(256) void k_trmv_lower(const T *__restrict__ Wm, int n,
                                                    const T *__restrict__ x, T *__restrict__ y) {
  const int row = blockIdx.x;
  const T *a = Wm + (int64_t)row * n;
  double s = 0.0;
  // (unrolled so that several loads are in flight; the sum order is unchanged)
#pragma unroll 8
  for (int c = threadIdx.x; c <= row; c += 256) s += (double)a[c] * (double)x[c];
  s = row_block_sum(s);
  if (threadIdx.x == 0) y[row] = (T)s;
}

template <typename T>
__global__ __launch_bounds__(256) void k_trmv_upper_scatter(int nx, int ny, int nz, int wy, int wz,
                                                            const int *__restrict__ ipos,
                                                            const T *__restrict__ WTm, int n,
                                                            const T *__restrict__ y,
                                                            T *__restrict__ out) {
  const int row = blockIdx.x;
  const T *a = WTm + (int64_t)row * n;
  double s = 0.0;
#pragma unroll 8
  for (int c = row + threadIdx.x; c < n; c += 256) s += (double)a[c] * (double)y[c];
  s = row_block_sum(s);
  const int d = ipos[row];
  if (threadIdx.x == 0 && d >= 0) {
    const int64_t p = d / 3;
    const int j = (int)(p % ny), k = (int)((p / ny) % nz), i = (int)(p / ((int64_t)ny * nz));
    out[3 * (((int64_t)(i + 1) * wz + (k + 1)) * wy + (j + 1)) + d % 3] = (T)s;
  }
}

template <typename T>
void launch_coarse_apply(const DevGrid &g, int l, const CoarseFactor &cf, const T *in, T *out) {
  const Level &L = g.lv[l];
  const int n = cf.n, np = cf.np;
  const T *Wm = cf.w<T>(), *WTm = cf.wt<T>();
  T *x = cf.cw<T>(), *y = x + np;  // padding entries of x stay zero
  const int bs = 256;
  hipLaunchKernelGGL((k_gather_interior<T>), dim3((n + bs - 1) / bs), dim3(bs), 0, stream(),
                     L.nx, L.ny, L.nz, L.wy, L.wz, cf.pos, in, x);
  hipLaunchKernelGGL((k_trmv_lower<T>), dim3(np), dim3(256), 0, stream(), Wm, np, x, y);
  hipLaunchKernelGGL((k_trmv_upper_scatter<T>), dim3(np), dim3(256), 0, stream(), L.nx,
                     L.ny, L.nz, L.wy, L.wz, cf.ipos, WTm, np, y, out);
}

template void launch_coarse_apply<double>(const DevGrid &, int, const CoarseFactor &,
                                          const double *, double *);
template void launch_coarse_apply<float>(const DevGrid &, int, const CoarseFactor &,
                                         const float *, float *);

}  // namespace t3d

// A (n x n FP64, device, symmetric positive definite with half bandwidth bw)
// <- A^{-1}.
extern "C" int top3d_dense_spd_inverse(double *A_dev, int n, int bw) {
  using namespace t3d;
  if (n <= 0 || bw < 0 || !A_dev) {
    set_error(TOP3D_ERR_ARG, "top3d_dense_spd_inverse: n=%d bw=%d", n, bw);
    return TOP3D_ERR_ARG;
  }
  CoarseFactor cf;
  if (!cf.allocate(n, bw, false, 0)) {
    cf.release();
    return top3d_last_error() ? top3d_last_error() : TOP3D_ERR_HIP;
  }
  cf.begin_assembly();
  int rc = T3D_HIP(hipMemcpy2DAsync(cf.A, sizeof(double) * cf.np, A_dev, sizeof(double) * n,
                                    sizeof(double) * n, n, hipMemcpyDeviceToDevice, stream()))
               ? cf.factor()
               : TOP3D_ERR_HIP;
  if (rc == TOP3D_OK) {
    const int nblk = (n + kNB - 1) / kNB;
    hipLaunchKernelGGL(k_gemm_wtw, dim3(nblk * nblk), dim3(256), 0, stream(), cf.WT, cf.W, n,
                       cf.np, A_dev);
    if (!T3D_HIP(hipStreamSynchronize(stream()))) rc = TOP3D_ERR_HIP;
  }
  cf.release();
  return rc;
}
