// Dense FP64 direct solver of the coarsest level (replaces CHOLMOD,
// stencil_assembly.c:366-503).
//
// The coarsest matrix (n = 3 * nodes, <= ~4k) is banded: with the x-slowest
// node order its half bandwidth is 3*(ny*nz + ny + 1) + 2, and Cholesky
// keeps that band.  Per design iteration:
//   1. banded right-looking blocked Cholesky A = L L^T (64-wide blocks): the
//      diagonal block on one workgroup, the panel by row-parallel
//      substitution, the trailing band update on FP64 MFMA
//      (v_mfma_f64_16x16x4f64), the only dense contraction of the solver;
//   2. inverses of the 64x64 diagonal blocks of L;
//   3. the explicit inverse W = A^{-1} from W = I by a forward and a backward
//      banded solve with n right-hand sides, one workgroup per 16-column
//      strip, every triangular step turned into small GEMMs.
// Each V-cycle then applies A^{-1} as one parallel GEMV.
#include <algorithm>

#include "t3d_internal.hpp"

namespace t3d {

constexpr int kNB = 64;

typedef double double4_t __attribute__((ext_vector_type(4)));

// ---- 1a. Cholesky of the diagonal block (left-looking, in LDS) -------------
__global__ __launch_bounds__(64) void k_potrf_diag(double *__restrict__ A, int64_t n,
                                                   int k0, int kb, int *__restrict__ info) {
  __shared__ double L[kNB][kNB + 1];
  const int t = threadIdx.x;
  for (int r = 0; r < kb; r++)
    if (t < kb) L[r][t] = (t <= r) ? A[(k0 + r) * n + k0 + t] : 0.0;
  __syncthreads();
  for (int j = 0; j < kb; j++) {
    if (t == 0) {
      double d = L[j][j];
      for (int m = 0; m < j; m++) d -= L[j][m] * L[j][m];
      if (!(d > 0.0)) {
        atomicExch(info, 1);
        d = 1.0;
      }
      L[j][j] = sqrt(d);
    }
    __syncthreads();
    if (t > j && t < kb) {
      double v = L[t][j];
      for (int m = 0; m < j; m++) v -= L[t][m] * L[j][m];
      L[t][j] = v / L[j][j];
    }
    __syncthreads();
  }
  for (int r = 0; r < kb; r++)
    if (t <= r) A[(k0 + r) * n + k0 + t] = L[r][t];
}

// ---- 1b. panel: A[i, k0:k0+kb] <- A[i, k0:k0+kb] L_kk^{-T}, rows in band --
__global__ __launch_bounds__(64) void k_trsm_panel(double *__restrict__ A, int64_t n, int k0,
                                                   int kb, int r0, int r1) {
  __shared__ double L[kNB][kNB + 1];
  const int t = threadIdx.x;
  for (int r = 0; r < kb; r++)
    if (t < kb) L[r][t] = (t <= r) ? A[(k0 + r) * n + k0 + t] : 0.0;
  __syncthreads();
  const int i = r0 + blockIdx.x * 64 + t;
  if (i >= r1) return;
  double x[kNB];
  double *row = A + (int64_t)i * n + k0;
#pragma unroll
  for (int j = 0; j < kNB; j++) x[j] = j < kb ? row[j] : 0.0;
#pragma unroll
  for (int j = 0; j < kNB; j++) {
    if (j < kb) {
      double v = x[j];
#pragma unroll
      for (int m = 0; m < j; m++) v -= x[m] * L[j][m];
      x[j] = v / L[j][j];
    }
  }
#pragma unroll
  for (int j = 0; j < kNB; j++)
    if (j < kb) row[j] = x[j];
}

// ---- 1c. trailing update C_IJ -= P_I P_J^T on FP64 MFMA --------------------
// One workgroup (4 waves) per 64x64 tile of the trailing band, each wave a
// 32x32 quadrant as 2x2 blocks of v_mfma_f64_16x16x4f64.  MFMA operand
// layout (16x16x4, f64, measured on gfx950): lane l feeds A[l%16][l/16] and
// B[l/16][l%16]; accumulator register v of lane l holds D[l/16 + 4v][l%16].
__global__ __launch_bounds__(256) void k_band_update(double *__restrict__ A, int64_t n, int k0,
                                                     int kb, int t0, int ntiles, int bwt) {
  __shared__ double Pi[kNB][kNB + 1];
  __shared__ double Pj[kNB][kNB + 1];
  // tile enumeration: row tile I = t0 + a, col tile J = I - d, d in [0, bwt]
  const int a = blockIdx.x / (bwt + 1), d = blockIdx.x % (bwt + 1);
  const int I = t0 + a, J = I - d;
  if (a >= ntiles || J < t0) return;
  const int i0 = I * kNB, j0 = J * kNB;
  const int tid = threadIdx.x;
  for (int e = tid; e < kNB * kNB; e += 256) {
    const int r = e / kNB, c = e % kNB;
    Pi[r][c] = (i0 + r < n && c < kb) ? A[(int64_t)(i0 + r) * n + k0 + c] : 0.0;
    Pj[r][c] = (j0 + r < n && c < kb) ? A[(int64_t)(j0 + r) * n + k0 + c] : 0.0;
  }
  __syncthreads();
  const int w = tid >> 6, lane = tid & 63;
  const int qr = (w >> 1) * 32, qc = (w & 1) * 32;
  double4_t acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) acc[x][y] = double4_t{0.0, 0.0, 0.0, 0.0};
  const int li = lane & 15, lk = lane >> 4;
  for (int k = 0; k < kNB; k += 4) {
    double av[2], bv[2];
#pragma unroll
    for (int x = 0; x < 2; x++) av[x] = Pi[qr + 16 * x + li][k + lk];
#pragma unroll
    for (int y = 0; y < 2; y++) bv[y] = Pj[qc + 16 * y + li][k + lk];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 2; y++)
        acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x], bv[y], acc[x][y], 0, 0, 0);
  }
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = i0 + qr + 16 * x + lk + 4 * v;
        const int c = j0 + qc + 16 * y + li;
        if (r < n && c < n && c <= r) A[(int64_t)r * n + c] -= acc[x][y][v];
      }
}

// ---- 2. inverse of each 64x64 diagonal block of L (lower) ------------------
__global__ __launch_bounds__(64) void k_trtri_diag(const double *__restrict__ A, int64_t n,
                                                   double *__restrict__ Dinv) {
  __shared__ double L[kNB][kNB + 1];
  const int b = blockIdx.x, t = threadIdx.x;
  const int k0 = b * kNB, kb = (int)std::min<int64_t>(kNB, n - k0);
  for (int r = 0; r < kNB; r++)
    L[r][t] = (r < kb && t < kb && t <= r) ? A[(int64_t)(k0 + r) * n + k0 + t] : (r == t ? 1.0 : 0.0);
  __syncthreads();
  // column t of L^{-1}: forward substitution with e_t
  double x[kNB];
#pragma unroll
  for (int r = 0; r < kNB; r++) {
    double v = (r == t) ? 1.0 : 0.0;
#pragma unroll
    for (int m = 0; m < r; m++) v -= L[r][m] * x[m];
    x[r] = v / L[r][r];
  }
  double *D = Dinv + (int64_t)b * kNB * kNB;
#pragma unroll
  for (int r = 0; r < kNB; r++) D[r * kNB + t] = x[r];
}

// ---- 3. banded solves with n right-hand sides, in place on W ---------------
// forward: W <- L^{-1} W; backward: W <- L^{-T} W.  One workgroup per
// 16-column strip; per 64-row block: B -= L_band B_band (GEMM over the
// band), then B <- Dinv B.
constexpr int kStrip = 16;

template <bool FWD>
__global__ __launch_bounds__(256) void k_band_solve(const double *__restrict__ A, int64_t n,
                                                    int bw, const double *__restrict__ Dinv,
                                                    double *__restrict__ W) {
  __shared__ double B[kNB][kStrip + 1];
  __shared__ double Di[kNB][kNB + 1];
  const int c0 = blockIdx.x * kStrip;
  const int tid = threadIdx.x;
  const int col = tid % kStrip, rg = tid / kStrip;  // 16 row groups x 4 rows
  const int nblk = (int)((n + kNB - 1) / kNB);
  for (int s = 0; s < nblk; s++) {
    const int b = FWD ? s : nblk - 1 - s;
    const int r0 = b * kNB;
    const int kb = (int)std::min<int64_t>(kNB, n - r0);
    // B = W[block rows, strip] - (band coupling) * W[band rows, strip]
    for (int q = 0; q < 4; q++) {
      const int r = rg * 4 + q;
      const int gr = r0 + r;
      double v = 0.0;
      const int gc = c0 + col;
      if (r < kb && gc < n) {
        v = W[(int64_t)gr * n + gc];
        if (FWD) {  // rows m in [max(0, gr-bw), r0): L[gr][m] W[m]
          const int m0 = (int)std::max<int64_t>(0, (int64_t)gr - bw);
          for (int m = m0; m < r0; m++) v -= A[(int64_t)gr * n + m] * W[(int64_t)m * n + gc];
        } else {    // rows m in [r0+kb, min(n, gr+bw+1)): L[m][gr] W[m]
          const int64_t m1 = std::min<int64_t>(n, (int64_t)gr + bw + 1);
          for (int64_t m = r0 + kb; m < m1; m++) v -= A[m * n + gr] * W[m * n + gc];
        }
      }
      B[r][col] = v;
    }
    const double *D = Dinv + (int64_t)b * kNB * kNB;
    for (int e = tid; e < kNB * kNB; e += 256) {
      const int r = e / kNB, c = e % kNB;
      Di[r][c] = FWD ? D[r * kNB + c] : D[c * kNB + r];  // L_bb^{-1} or its transpose
    }
    __syncthreads();
    for (int q = 0; q < 4; q++) {
      const int r = rg * 4 + q;
      const int gc = c0 + col;
      if (r < kb && gc < n) {
        double v = 0.0;
        for (int m = 0; m < kb; m++) v += Di[r][m] * B[m][col];
        W[(int64_t)(r0 + r) * n + gc] = v;
      }
    }
    __syncthreads();
  }
}

__global__ void k_identity(double *__restrict__ W, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n * n) W[e] = (e / n == e % n) ? 1.0 : 0.0;
}

// A (n x n, row-major, symmetric) <- A^{-1}, via work (>= n*n + 65*64*64 +
// 1 doubles).  Returns TOP3D_OK or TOP3D_ERR_NOT_SPD.
int dense_cholesky_inverse_banded(double *A, int n, int bw, double *work) {
  const int64_t nn = (int64_t)n * n;
  double *W = work;
  const int nblk = (n + kNB - 1) / kNB;
  double *Dinv = work + nn;
  int *info = (int *)(Dinv + (int64_t)nblk * kNB * kNB);
  hipMemsetAsync(info, 0, sizeof(int), stream());
  const int bwt = (bw + kNB - 1) / kNB + 1;  // tile half-bandwidth
  for (int b = 0; b < nblk; b++) {
    const int k0 = b * kNB, kb = std::min(kNB, n - k0);
    hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(64), 0, stream(), A, (int64_t)n, k0, kb,
                       info);
    const int r0 = k0 + kb, r1 = std::min(n, k0 + kb + bw + 1);
    if (r1 > r0) {
      hipLaunchKernelGGL(k_trsm_panel, dim3((r1 - r0 + 63) / 64), dim3(64), 0, stream(), A,
                         (int64_t)n, k0, kb, r0, r1);
      const int t0 = b + 1, t1 = (r1 - 1) / kNB;  // trailing tiles touched
      const int ntiles = t1 - t0 + 1;
      if (ntiles > 0)
        hipLaunchKernelGGL(k_band_update, dim3(ntiles * (bwt + 1)), dim3(256), 0, stream(),
                           A, (int64_t)n, k0, kb, t0, ntiles, bwt);
    }
  }
  hipLaunchKernelGGL(k_trtri_diag, dim3(nblk), dim3(64), 0, stream(), A, (int64_t)n, Dinv);
  const int bs = 256;
  hipLaunchKernelGGL(k_identity, dim3((nn + bs - 1) / bs), dim3(bs), 0, stream(), W,
                     (int64_t)n);
  const int nstrip = (n + kStrip - 1) / kStrip;
  hipLaunchKernelGGL((k_band_solve<true>), dim3(nstrip), dim3(256), 0, stream(), A,
                     (int64_t)n, bw, Dinv, W);
  hipLaunchKernelGGL((k_band_solve<false>), dim3(nstrip), dim3(256), 0, stream(), A,
                     (int64_t)n, bw, Dinv, W);
  int hinfo = 0;
  hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, stream());
  hipMemcpyAsync(A, W, sizeof(double) * nn, hipMemcpyDeviceToDevice, stream());
  if (!T3D_HIP(hipStreamSynchronize(stream()))) return TOP3D_ERR_HIP;
  if (hinfo) {
    set_error(TOP3D_ERR_NOT_SPD, "coarse Cholesky: non-positive pivot");
    return TOP3D_ERR_NOT_SPD;
  }
  return TOP3D_OK;
}

int dense_cholesky_inverse(double *A, int n, double *work) {
  return dense_cholesky_inverse_banded(A, n, n, work);
}

}  // namespace t3d

extern "C" int top3d_dense_spd_inverse(double *A_dev, int n, int bw) {
  using namespace t3d;
  if (n <= 0 || bw < 0) {
    set_error(TOP3D_ERR_ARG, "top3d_dense_spd_inverse: n=%d bw=%d", n, bw);
    return TOP3D_ERR_ARG;
  }
  const int nblk = (n + kNB - 1) / kNB;
  double *work = (double *)dalloc(sizeof(double) * ((size_t)n * n + (size_t)nblk * kNB * kNB + 8));
  if (!work) return TOP3D_ERR_HIP;
  const int rc = dense_cholesky_inverse_banded(A_dev, n, std::min(bw, n), work);
  hipStreamSynchronize(stream());
  dfree(work);
  return rc;
}
