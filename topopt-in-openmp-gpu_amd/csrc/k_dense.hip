// Dense FP64 direct solver of the coarsest level (replaces CHOLMOD,
// stencil_assembly.c:366-503).
//
// The coarsest matrix (n = 3 * interior nodes, 4131 at 16x8x8 elements) is
// banded: with the x-slowest node order its half bandwidth is
// 3*(ny*nz + ny + 1) + 2, and Cholesky keeps that band.  Per design iteration:
//   1. banded right-looking Cholesky A = L L^T over 64x64 tiles, one launch
//      per block column b: each workgroup owns one tile (I, J) of the trailing
//      band, forms the panel tiles Q_I = A_Ib L_bb^{-T} and Q_J on FP64 MFMA
//      (v_mfma_f64_16x16x4f64) and applies C_IJ -= Q_I Q_J^T; the diagonal
//      workgroups write the panel of L, and the workgroup of tile
//      (b+1, b+1) factors it in LDS right after its update (Cholesky, then
//      the inverse of the 64x64 factor);
//   2. W = L^{-1} (dense lower triangle), one workgroup per 16-column strip
//      marching down the block rows with the last band of block rows in LDS,
//      every step a band GEMM plus a product with the inverse diagonal block,
//      both on MFMA; W^T is written alongside.
// A V-cycle applies A^{-1} = W^T W as two triangular GEMVs.
#include <algorithm>

#include "t3d_internal.hpp"

namespace t3d {

constexpr int kNB = 64;         // tile size
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
//   (a) wave 0, lanes 0..15 (one row each, cross-lane reads): the 16x16
//       Cholesky L_D and its inverse X_D;
//   (b) the panel below: P <- P X_D^T;
//   (c) the trailing lower triangle: C -= P P^T as 16x16 MFMA tiles;
// then the off-diagonal 16x16 blocks of the inverse by distance.
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

__device__ __forceinline__ void factor_tile(double *T, double *X, double *Ys, double *wbuf,
                                            int *info) {
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int li = lane & 15, lk = lane >> 4;
  for (int e = t; e < kNB * kNB; e += 256) X[(e >> 6) * kLd + (e & 63)] = 0.0;
  __syncthreads();
  for (int kb = 0; kb < 4; kb++) {
    const int d0 = 16 * kb;
    if (t < 16) {  // (a): one row per lane, broadcasts by v_readlane
      const int r = t;
      double v[16];
#pragma unroll
      for (int c = 0; c < 16; c++) v[c] = (c <= r) ? T[(d0 + r) * kLd + d0 + c] : 0.0;
#pragma unroll
      for (int j = 0; j < 16; j++) {
        double piv = readlane_d(v[j], j);
        if (!(piv > 0.0)) {
          if (r == 0) info[0] = 1;
          piv = 1.0;
        }
        const double ljj = sqrt(piv);
        const double inv = 1.0 / ljj;
        const double lrj = (r == j) ? ljj : v[j] * inv;
        v[j] = (r >= j) ? lrj : 0.0;
#pragma unroll
        for (int c = j + 1; c < 16; c++) {
          const double lcj = readlane_d(lrj, c);
          if (c <= r) v[c] -= lrj * lcj;
        }
      }
#pragma unroll
      for (int c = 0; c < 16; c++) T[(d0 + r) * kLd + d0 + c] = v[c];
      // column r of X_D = L_D^{-1}: forward substitution with the rows of
      // L_D broadcast from their lanes
      double dinv[16], x[16];
#pragma unroll
      for (int i = 0; i < 16; i++) dinv[i] = 1.0 / readlane_d(v[i], i);
#pragma unroll
      for (int i = 0; i < 16; i++) {
        double acc = (i == r) ? 1.0 : 0.0;
#pragma unroll
        for (int m = 0; m < i; m++) acc -= readlane_d(v[m], i) * x[m];
        x[i] = acc * dinv[i];
        X[(d0 + i) * kLd + d0 + r] = x[i];
      }
    }
    __syncthreads();
    const int npan = kNB - d0 - 16;  // panel rows below the diagonal block
    if (npan > 0) {
      // (b) P[p][c] <- sum_{m <= c} P[p][m] X_D[c][m]
      double pv[3];
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const int e = t + 256 * q;
        pv[q] = 0.0;
        if (e < npan * 16) {
          const int p = d0 + 16 + (e >> 4), c = e & 15;
          double acc = 0.0;
          for (int m = 0; m <= c; m++) acc += T[p * kLd + d0 + m] * X[(d0 + c) * kLd + d0 + m];
          pv[q] = acc;
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const int e = t + 256 * q;
        if (e < npan * 16) T[(d0 + 16 + (e >> 4)) * kLd + d0 + (e & 15)] = pv[q];
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
  // off-diagonal blocks of the inverse by distance d:
  //   X_IJ = -X_II sum_{K=J}^{I-1} L_IK X_KJ
  const int rr = t >> 4, cc = t & 15;
  for (int d = 1; d < 4; d++) {
    const int ntile = 4 - d;
    for (int a = 0; a < ntile; a++) {
      const int I = d + a, J = a;
      double y = 0.0;
      for (int K = J; K < I; K++)
#pragma unroll
        for (int m = 0; m < 16; m++)
          y += T[(16 * I + rr) * kLd + 16 * K + m] * X[(16 * K + m) * kLd + 16 * J + cc];
      Ys[a * 256 + t] = y;
    }
    __syncthreads();
    for (int a = 0; a < ntile; a++) {
      const int I = d + a, J = a;
      double xv = 0.0;
#pragma unroll
      for (int m = 0; m < 16; m++) xv -= X[(16 * I + rr) * kLd + 16 * I + m] * Ys[a * 256 + m * 16 + cc];
      X[(16 * I + rr) * kLd + 16 * J + cc] = xv;
    }
    __syncthreads();
  }
}

// Write the factored diagonal tile (blk, blk): L to Lf, L^{-1} to Dinv.
__device__ inline void store_diag(const double *T, const double *X, double *__restrict__ Lf,
                                  double *__restrict__ Dinv, int n, int blk) {
  const int k0 = blk * kNB;
  for (int e = threadIdx.x; e < kNB * kNB; e += 256) {
    const int r = e >> 6, c = e & 63;
    if (k0 + r < n && k0 + c < n) Lf[(int64_t)(k0 + r) * n + k0 + c] = T[r * kLd + c];
    Dinv[(int64_t)blk * kNB * kNB + e] = X[r * kLd + c];
  }
}

// One block column b of the banded Cholesky (b = -1: factor tile (0, 0)).
// Workgroup p handles trailing tile (I, J), b+1 <= J <= I <= t1, enumerated
// row by row.  A keeps the Schur-complement updates; Lf receives L.
__global__ __launch_bounds__(256) void k_chol_step(double *__restrict__ A, double *__restrict__ Lf,
                                                   double *__restrict__ Dinv, int n, int b,
                                                   int *__restrict__ info) {
  __shared__ double PI[kNB * kLd];
  __shared__ double PJ[kNB * kLd];
  __shared__ double DV[kNB * kLd];
  __shared__ double CT[kNB * kLd];
  __shared__ double Ys[3 * 256];
  __shared__ double wbuf[32];
  const int tid = threadIdx.x;
  int fblk = 0;  // diagonal block factored by this workgroup (if any)
  if (b < 0) {
    double rv[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int e = tid + 256 * q, r = e >> 6, c = e & 63;
      rv[q] = (r < n && c < n) ? A[(int64_t)r * n + c] : (r == c ? 1.0 : 0.0);
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int e = tid + 256 * q;
      DV[(e >> 6) * kLd + (e & 63)] = rv[q];
    }
  } else {
    int p = blockIdx.x, I = b + 1;
    while (p >= I - b) {
      p -= I - b;
      I++;
    }
    const int J = b + 1 + p;
    const int k0 = b * kNB;
    // stage the panel tiles P_I, P_J, L_bb^{-1} and the target tile C_IJ;
    // all loads of a thread are issued before its LDS stores
    {
      double ri[16], rj[16], rd[16], rc[16];
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int e = tid + 256 * q, r = e >> 6, c = e & 63;
        const int64_t gi = (int64_t)I * kNB + r, gj = (int64_t)J * kNB + r,
                      gjc = (int64_t)J * kNB + c;
        ri[q] = (gi < n) ? A[gi * n + k0 + c] : 0.0;
        rj[q] = (J != I && gj < n) ? A[gj * n + k0 + c] : 0.0;
        rd[q] = Dinv[(int64_t)b * kNB * kNB + e];
        rc[q] = (gi < n && gjc < n) ? A[gi * n + gjc] : (r == c ? 1.0 : 0.0);
      }
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int e = tid + 256 * q, o = (e >> 6) * kLd + (e & 63);
        PI[o] = ri[q];
        if (J != I) PJ[o] = rj[q];
        DV[o] = rd[q];
        CT[o] = rc[q];
      }
    }
    __syncthreads();
    const int w = tid >> 6, lane = tid & 63;
    const int qr = (w >> 1) * 32, qc = (w & 1) * 32;
    double4_t qi[2][2], qj[2][2];
    zero_acc(qi);
    zero_acc(qj);
    mm_nt_tile(PI, DV, qr, qc, lane, qi);  // Q_I = P_I L_bb^{-T}
    if (J != I) mm_nt_tile(PJ, DV, qr, qc, lane, qj);
    __syncthreads();
    store_acc(PI, qi, qr, qc, lane);
    if (J != I) store_acc(PJ, qj, qr, qc, lane);
    __syncthreads();
    if (I == J) {
#pragma unroll 4
      for (int e = tid; e < kNB * kNB; e += 256) {
        const int r = e >> 6, c = e & 63;
        const int64_t gi = (int64_t)I * kNB + r;
        if (gi < n) Lf[gi * n + k0 + c] = PI[r * kLd + c];
      }
    }
    double4_t cacc[2][2];
    zero_acc(cacc);
    mm_nt_tile(PI, (J != I) ? PJ : PI, qr, qc, lane, cacc);  // Q_I Q_J^T
    const bool next_diag = (I == b + 1);  // then J == I as well
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 2; y++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int r = qr + 16 * x + lk + 4 * v, c = qc + 16 * y + li;
          const int64_t gi = (int64_t)I * kNB + r, gj = (int64_t)J * kNB + c;
          const double a = CT[r * kLd + c] - cacc[x][y][v];
          if (next_diag) DV[r * kLd + c] = (gi < n && gj < n) ? a : (r == c ? 1.0 : 0.0);
          else if (gi < n && gj < n) A[gi * n + gj] = a;
        }
    if (!next_diag) return;
    fblk = b + 1;
  }
  __syncthreads();
  factor_tile(DV, PJ, Ys, wbuf, info);
  store_diag(DV, PJ, Lf, Dinv, n, fblk);
}

// ---------------------------------------------------------------------------
// W = L^{-1}, WT = W^T.  Workgroup s owns columns [16s, 16s+16), starting at
// block row J = 16s/64 (W_J = columns of L_JJ^{-1}); below it
//   W_I = -L_II^{-1} sum_{K=max(J, I-kband)}^{I-1} L_IK W_K,
// with the last kband+1 block rows of the strip kept in an LDS ring (for
// bands too wide for LDS, nring = 0, they are read back from W).  The
// L_IK tiles are staged through LDS with every thread's loads of a tile in
// flight together, the next tile loading while the current one is used;
// the rows of L_II^{-1} a wave needs are loaded at the start of the step.
__global__ __launch_bounds__(256) void k_linv_strip(const double *__restrict__ Lf,
                                                    const double *__restrict__ Dinv, int n,
                                                    int kband, int nring,
                                                    double *__restrict__ W,
                                                    double *__restrict__ WT) {
  extern __shared__ double lds[];
  double *ring = lds;                               // [nring][64][16]
  double *Bx = ring + (size_t)nring * kNB * kStrip;  // [64][17]
  double *Lt = Bx + kNB * 17;                        // [2][64][kLd]
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int nblk = (n + kNB - 1) / kNB;
  const int c0 = blockIdx.x * kStrip;
  const int J = c0 / kNB, cofs = c0 - J * kNB;
  const int gc = c0 + li;
  {
    double *slot = nring ? ring + (size_t)(J % nring) * kNB * kStrip : nullptr;
    for (int e = tid; e < kNB * kStrip; e += 256) {
      const int r = e >> 4, j = e & 15;
      const double v = Dinv[(int64_t)J * kNB * kNB + r * kNB + cofs + j];
      if (nring) slot[e] = v;
      const int64_t gr = (int64_t)J * kNB + r, gcol = c0 + j;
      if (gr < n && gcol < n) {
        W[gr * n + gcol] = v;
        WT[gcol * n + gr] = v;
      }
    }
  }
  auto load_tile = [&](int I, int K, double (&reg)[16]) {
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int e = tid + 256 * q, r = e >> 6, c = e & 63;
      const int64_t row = (int64_t)I * kNB + r;
      reg[q] = row < n ? Lf[row * n + (int64_t)K * kNB + c] : 0.0;
    }
  };
  auto store_tile = [&](double *dst, const double (&reg)[16]) {
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int e = tid + 256 * q;
      dst[(e >> 6) * kLd + (e & 63)] = reg[q];
    }
  };
  __syncthreads();
  for (int I = J + 1; I < nblk; I++) {
    const int Kmin = max(J, I - kband);
    // rows 16w.. of L_II^{-1}: columns k < 16(w+1) are non-zero
    const double *D = Dinv + (int64_t)I * kNB * kNB;
    double dreg[16];
#pragma unroll
    for (int q = 0; q < 16; q++)
      dreg[q] = (4 * q < 16 * (w + 1)) ? D[(16 * w + li) * kNB + 4 * q + lk] : 0.0;
    double treg[16];
    load_tile(I, Kmin, treg);
    store_tile(Lt, treg);
    __syncthreads();
    double4_t acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    for (int K = Kmin; K < I; K++) {
      const double *cur = Lt + (size_t)((K - Kmin) & 1) * kNB * kLd;
      double *nxt = Lt + (size_t)(((K - Kmin) & 1) ^ 1) * kNB * kLd;
      if (K + 1 < I) load_tile(I, K + 1, treg);
      if (nring) {
        const double *slot = ring + (size_t)(K % nring) * kNB * kStrip;
#pragma unroll
        for (int k = 0; k < kNB; k += 8) {
          acc0 = mfma4(cur[(16 * w + li) * kLd + k + lk], slot[(k + lk) * kStrip + li], acc0);
          acc1 = mfma4(cur[(16 * w + li) * kLd + k + 4 + lk], slot[(k + 4 + lk) * kStrip + li],
                       acc1);
        }
      } else {
        const double *wk = W + (int64_t)K * kNB * n + gc;
#pragma unroll
        for (int k = 0; k < kNB; k += 8) {
          const double b0 = gc < n ? wk[(int64_t)(k + lk) * n] : 0.0;
          const double b1 = gc < n ? wk[(int64_t)(k + 4 + lk) * n] : 0.0;
          acc0 = mfma4(cur[(16 * w + li) * kLd + k + lk], b0, acc0);
          acc1 = mfma4(cur[(16 * w + li) * kLd + k + 4 + lk], b1, acc1);
        }
      }
      if (K + 1 < I) store_tile(nxt, treg);
      __syncthreads();
    }
#pragma unroll
    for (int v = 0; v < 4; v++) Bx[(16 * w + lk + 4 * v) * 17 + li] = acc0[v] + acc1[v];
    __syncthreads();
    double4_t o = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 16; q++)
      if (4 * q < 16 * (w + 1)) o = mfma4(dreg[q], Bx[(4 * q + lk) * 17 + li], o);
    double *slot = nring ? ring + (size_t)(I % nring) * kNB * kStrip : nullptr;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = 16 * w + lk + 4 * v;
      const double val = -o[v];
      if (nring) slot[r * kStrip + li] = val;
      const int64_t gr = (int64_t)I * kNB + r;
      if (gr < n && gc < n) {
        W[gr * n + gc] = val;
        WT[(int64_t)gc * n + gr] = val;
      }
    }
    __syncthreads();
  }
}

// C = WT W restricted to m >= max(i, j) (the explicit inverse A^{-1}),
// 64x64 tiles, k staged through LDS in chunks of 16.
__global__ __launch_bounds__(256) void k_gemm_wtw(const double *__restrict__ WT,
                                                  const double *__restrict__ W, int n,
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
      As[r][k] = (i0 + r < n && m0 + k < n) ? WT[(int64_t)(i0 + r) * n + m0 + k] : 0.0;
      const int kk = e >> 6, c = e & 63;
      Bs[kk][c] = (m0 + kk < n && j0 + c < n) ? W[(int64_t)(m0 + kk) * n + j0 + c] : 0.0;
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
static size_t strip_lds_bytes(int nring) {
  return sizeof(double) * ((size_t)nring * kNB * kStrip + (size_t)kNB * 17 + 2 * (size_t)kNB * kLd);
}

static void *zalloc(size_t bytes) { return dalloc(bytes); }

bool CoarseFactor::allocate(int n_, int bw_, bool f32) {
  n = n_;
  bw = std::min(bw_, n_ - 1);
  if (bw < 0) bw = 0;
  nblk = (n + kNB - 1) / kNB;
  kband = (bw + kNB - 1) / kNB;
  nring = std::min(kband + 1, nblk);
  if (strip_lds_bytes(nring) > 158 * 1024) nring = 0;  // band rows read back from W
  const size_t nn = (size_t)n * n;
  A = (double *)zalloc(sizeof(double) * nn);
  L = (double *)zalloc(sizeof(double) * nn);
  W = (double *)zalloc(sizeof(double) * nn);
  WT = (double *)zalloc(sizeof(double) * nn);
  Dinv = (double *)zalloc(sizeof(double) * (size_t)nblk * kNB * kNB);
  info = (int *)zalloc(sizeof(int) * 4);
  cw64 = (double *)zalloc(sizeof(double) * 2 * (size_t)n);
  if (f32) {
    W32 = (float *)zalloc(sizeof(float) * nn);
    WT32 = (float *)zalloc(sizeof(float) * nn);
    cw32 = (float *)zalloc(sizeof(float) * 2 * (size_t)n);
  }
  return A && L && W && WT && Dinv && info && cw64 && (!f32 || (W32 && WT32 && cw32));
}

void CoarseFactor::release() {
  for (void *p : {(void *)A, (void *)L, (void *)W, (void *)WT, (void *)Dinv, (void *)info,
                  (void *)cw64, (void *)W32, (void *)WT32, (void *)cw32})
    if (p) dfree(p);
  A = L = W = WT = Dinv = cw64 = nullptr;
  W32 = WT32 = cw32 = nullptr;
  info = nullptr;
}

int CoarseFactor::factor() {
  hipStream_t s = stream();
  hipMemsetAsync(info, 0, sizeof(int), s);
  hipLaunchKernelGGL(k_chol_step, dim3(1), dim3(256), 0, s, A, L, Dinv, n, -1, info);
  for (int b = 0; b + 1 < nblk; b++) {
    const int k0 = b * kNB;
    const int t1 = std::min(nblk - 1, (k0 + kNB - 1 + bw) / kNB);
    const int ntile = t1 - b;
    hipLaunchKernelGGL(k_chol_step, dim3(ntile * (ntile + 1) / 2), dim3(256), 0, s, A, L, Dinv,
                       n, b, info);
  }
  const size_t lds = strip_lds_bytes(nring);
  static bool attr = false;
  if (!attr) {
    hipFuncSetAttribute((const void *)k_linv_strip, hipFuncAttributeMaxDynamicSharedMemorySize,
                        158 * 1024);
    attr = true;
  }
  hipLaunchKernelGGL(k_linv_strip, dim3((n + kStrip - 1) / kStrip), dim3(256), lds, s, L, Dinv,
                     n, kband, nring, W, WT);
  if (W32) {
    launch_convert<float, double>((int64_t)n * n, W, W32);
    launch_convert<float, double>((int64_t)n * n, WT, WT32);
  }
  int hinfo = 0;
  hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s);
  if (!T3D_HIP(hipStreamSynchronize(s))) return TOP3D_ERR_HIP;
  if (hinfo) {
    set_error(TOP3D_ERR_NOT_SPD, "coarse Cholesky: non-positive pivot");
    return TOP3D_ERR_NOT_SPD;
  }
  return TOP3D_OK;
}

// ---------------------------------------------------------------------------
// Coarsest solve z = W^T (W r): gather the interior DOFs, lower-triangular
// GEMV, upper-triangular GEMV scattered back to the halo layout (the halo of
// z stays zero).  One workgroup per row, FP64 accumulation.
template <typename T>
__global__ void k_gather_interior(int nx, int ny, int nz, int wy, int wz,
                                  const T *__restrict__ in, T *__restrict__ out) {
  const int64_t nint = (int64_t)nx * ny * nz;
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= 3 * nint) return;
  const int64_t p = d / 3;
  const int j = (int)(p % ny), k = (int)((p / ny) % nz), i = (int)(p / ((int64_t)ny * nz));
  out[d] = in[3 * (((int64_t)(i + 1) * wz + (k + 1)) * wy + (j + 1)) + d % 3];
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
  for (int c = threadIdx.x; c <= row; c += 256) s += (double)a[c] * (double)x[c];
  s = row_block_sum(s);
  if (threadIdx.x == 0) y[row] = (T)s;
}

template <typename T>
__global__ __launch_bounds__(256) void k_trmv_upper_scatter(int nx, int ny, int nz, int wy, int wz,
                                                            const T *__restrict__ WTm, int n,
                                                            const T *__restrict__ y,
                                                            T *__restrict__ out) {
  const int row = blockIdx.x;
  const T *a = WTm + (int64_t)row * n;
  double s = 0.0;
  for (int c = row + threadIdx.x; c < n; c += 256) s += (double)a[c] * (double)y[c];
  s = row_block_sum(s);
  if (threadIdx.x == 0) {
    const int64_t p = row / 3;
    const int j = (int)(p % ny), k = (int)((p / ny) % nz), i = (int)(p / ((int64_t)ny * nz));
    out[3 * (((int64_t)(i + 1) * wz + (k + 1)) * wy + (j + 1)) + row % 3] = (T)s;
  }
}

template <typename T>
void launch_coarse_apply(const DevGrid &g, int l, const CoarseFactor &cf, const T *in, T *out) {
  const Level &L = g.lv[l];
  const int n = cf.n;
  const T *Wm = cf.w<T>(), *WTm = cf.wt<T>();
  T *x = cf.cw<T>(), *y = x + n;
  const int bs = 256;
  hipLaunchKernelGGL((k_gather_interior<T>), dim3((n + bs - 1) / bs), dim3(bs), 0, stream(),
                     L.nx, L.ny, L.nz, L.wy, L.wz, in, x);
  hipLaunchKernelGGL((k_trmv_lower<T>), dim3(n), dim3(256), 0, stream(), Wm, n, x, y);
  hipLaunchKernelGGL((k_trmv_upper_scatter<T>), dim3(n), dim3(256), 0, stream(), L.nx,
                     L.ny, L.nz, L.wy, L.wz, WTm, n, y, out);
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
  if (!cf.allocate(n, bw, false)) {
    cf.release();
    return top3d_last_error() ? top3d_last_error() : TOP3D_ERR_HIP;
  }
  hipMemcpyAsync(cf.A, A_dev, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToDevice, stream());
  int rc = cf.factor();
  if (rc == TOP3D_OK) {
    const int nblk = (n + kNB - 1) / kNB;
    hipLaunchKernelGGL(k_gemm_wtw, dim3(nblk * nblk), dim3(256), 0, stream(), cf.WT, cf.W, n,
                       A_dev);
    if (!T3D_HIP(hipStreamSynchronize(stream()))) rc = TOP3D_ERR_HIP;
  }
  cf.release();
  return rc;
}
