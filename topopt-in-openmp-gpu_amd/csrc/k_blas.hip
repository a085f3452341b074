// Vector kernels of the CG recurrence: dot products and norms as wavefront
// reductions with a fixed-partition second stage (deterministic), the fused
// CG updates with scalars kept on the device, conversions and fills.
#include "t3d_internal.hpp"

namespace t3d {

constexpr int kRedBlock = 256;
constexpr int kRedGrid = 1024;  // partial sums per reduction

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sm[kRedBlock / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sm[w] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kRedBlock / 64; i++) s += sm[i];
  return s;  // valid in thread 0
}

// stage 1: partials[block] = sum over a grid-stride range of a[i]*b[i]
template <typename T>
__global__ __launch_bounds__(kRedBlock) void k_dot_partial(const T *__restrict__ a,
                                                           const T *__restrict__ b,
                                                           int64_t n,
                                                           double *__restrict__ partials) {
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += (double)a[i] * (double)b[i];
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// stage 2: one block sums the partials in a fixed order (r.r, b.b), also into
// pinned host memory when host_rr is given (in place of a device-to-host copy)
__global__ __launch_bounds__(kRedBlock) void k_sum_rr(const double *__restrict__ partials, int n,
                                                      double *__restrict__ out,
                                                      double *host_rr) {
  double s = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += partials[i];
  s = block_sum(s);
  if (threadIdx.x == 0) {
    *out = s;
    if (host_rr) __hip_atomic_store(host_rr, s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

template <typename T>
void launch_dot(const T *a, const T *b, int64_t n, double *partials, double *result,
                double *host) {
  hipLaunchKernelGGL((k_dot_partial<T>), dim3(kRedGrid), dim3(kRedBlock), 0, stream(), a,
                     b, n, partials);
  hipLaunchKernelGGL(k_sum_rr, dim3(1), dim3(kRedBlock), 0, stream(), partials, kRedGrid,
                     result, host);
}

// Device-resident CG scalars (layout shared with the solver):
//   s[0] rho, s[1] rho_old, s[2] p.q, s[3] r.r, s[4] b.b
// ---- fused CG recurrence ---------------------------------------------------
// The CG scalars are reduced from per-workgroup partial sums by every block
// of the consuming kernel, in one fixed order, so no separate reduction
// launch sits between the producer and the consumer.

// block-wide sum returned to every thread
__device__ __forceinline__ double block_sum_all(double v) {
  __shared__ double sm[kRedBlock / 64 + 1];
  __syncthreads();
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sm[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int i = 0; i < (int)(blockDim.x + 63) / 64; i++) s += sm[i];
    sm[kRedBlock / 64] = s;
  }
  __syncthreads();
  return sm[kRedBlock / 64];
}

__device__ __forceinline__ double sum_partials_all(const double *__restrict__ partials, int n) {
  double s = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += partials[i];
  return block_sum_all(s);
}

// rho partials: r . z, z in the V-cycle's precision (and r too in the mixed
// form, whose FP32 z limits rho to FP32 accuracy anyway), summed in FP64
template <typename TR, typename TZ>
__global__ __launch_bounds__(kRedBlock) void k_dot_rz(const TR *__restrict__ r,
                                                      const TZ *__restrict__ z, int64_t n,
                                                      double *__restrict__ partials) {
  double s = 0;
#pragma unroll 4
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += (double)r[i] * (double)z[i];
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// rho = sum(partials); p = z (first) or p = (rho/rho_old) p + z
// (stencil_solvers.c:350-362); block 0 stores s[0] = rho.  After the first
// iteration it also applies the previous iteration's u += alpha p
// (stencil_solvers.c:366-371) with the p it replaces, alpha = s[1] / s[2] as
// the CG update left them, so that p is read once for both.  u is loaded and
// stored non-temporally: it is not needed again until the next iteration, and
// the memory-side cache (256 MiB) then keeps more of the new p for q = K p,
// which follows (95 against 105 us for that apply at 256x128x128).
template <typename TZ>
__global__ __launch_bounds__(kRedBlock) void k_pupdate_fused(
    int64_t n, const double *__restrict__ partials, int npart, double *__restrict__ s,
    double *__restrict__ p, const TZ *__restrict__ z, int first, double *__restrict__ u) {
  const double rho = sum_partials_all(partials, npart);
  if (blockIdx.x == 0 && threadIdx.x == 0) s[0] = rho;
  const double beta = first ? 0.0 : rho / s[1];
  const double alpha = first ? 0.0 : s[1] / s[2];
#pragma unroll 4
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double zi = (double)z[i];
    if (first) {
      p[i] = zi;
    } else {
      const double pi = p[i];
      __builtin_nontemporal_store(__builtin_nontemporal_load(&u[i]) + alpha * pi, &u[i]);
      p[i] = beta * pi + zi;
    }
  }
}

// u += alpha p with alpha = s[1] / s[2]: the last CG update's step, which
// k_pupdate_fused would have applied in the next iteration
__global__ void k_u_update(int64_t n, const double *__restrict__ s, double *__restrict__ u,
                           const double *__restrict__ p) {
  const double alpha = s[1] / s[2];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    u[i] += alpha * p[i];
}

// pq = sum(partials); alpha = rho / pq; r -= alpha q (stencil_solvers.c:
// 366-377); r32 = (float) r when given; r.r partials.  Block 0 stores
// rho_old = rho and pq; u += alpha p is left to k_pupdate_fused (or, after
// the last iteration, k_u_update).  r and q are accessed non-temporally (not
// needed again before the next CG update), r32 normally: the V-cycle reads it
// next.
__global__ __launch_bounds__(kRedBlock) void k_axpy_fused(
    int64_t n, const double *__restrict__ pq_part, int npart, double *__restrict__ s,
    double *__restrict__ r, const double *__restrict__ q, float *__restrict__ r32,
    double *__restrict__ rr_part) {
  const double pq = sum_partials_all(pq_part, npart);
  const double rho = s[0];
  const double alpha = rho / pq;
  double acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double ri =
        __builtin_nontemporal_load(&r[i]) - alpha * __builtin_nontemporal_load(&q[i]);
    __builtin_nontemporal_store(ri, &r[i]);
    if (r32) r32[i] = (float)ri;
    acc += ri * ri;
  }
  acc = block_sum_all(acc);
  if (threadIdx.x == 0) {
    rr_part[blockIdx.x] = acc;
    if (blockIdx.x == 0) {
      s[1] = rho;
      s[2] = pq;
    }
  }
}

template <typename TZ>
void launch_dot_rz(const TZ *r, const TZ *z, int64_t n, double *partials) {
  hipLaunchKernelGGL((k_dot_rz<TZ, TZ>), dim3(kRedGrid), dim3(kRedBlock), 0, stream(), r, z, n,
                     partials);
}
template <typename TZ>
void launch_pupdate_fused(int64_t n, const double *partials, int npart, double *scal, double *p,
                          const TZ *z, int first, double *u) {
  // 2048 blocks: 8 waves per SIMD on 256 CUs for this streaming update
  hipLaunchKernelGGL((k_pupdate_fused<TZ>), dim3(2 * kRedGrid), dim3(kRedBlock), 0, stream(), n,
                     partials, npart, scal, p, z, first, u);
}
void launch_u_update(int64_t n, const double *scal, double *u, const double *p) {
  hipLaunchKernelGGL(k_u_update, dim3(2 * kRedGrid), dim3(kRedBlock), 0, stream(), n, scal, u,
                     p);
}
void launch_axpy_fused(int64_t n, const double *pq_part, int npart, double *scal, double *r,
                       const double *q, float *r32, double *rr_part, double *rr,
                       double *host_rr) {
  hipLaunchKernelGGL(k_axpy_fused, dim3(kRedGrid), dim3(kRedBlock), 0, stream(), n, pq_part,
                     npart, scal, r, q, r32, rr_part);
  hipLaunchKernelGGL(k_sum_rr, dim3(1), dim3(kRedBlock), 0, stream(), rr_part, kRedGrid, rr,
                     host_rr);
}
template void launch_dot_rz<double>(const double *, const double *, int64_t, double *);
template void launch_dot_rz<float>(const float *, const float *, int64_t, double *);
template void launch_pupdate_fused<double>(int64_t, const double *, int, double *, double *,
                                           const double *, int, double *);
template void launch_pupdate_fused<float>(int64_t, const double *, int, double *, double *,
                                          const float *, int, double *);

template <typename TO, typename TI>
__global__ void k_convert(int64_t n, const TI *__restrict__ in, TO *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (TO)in[i];
}
template <typename TO, typename TI>
void launch_convert(int64_t n, const TI *in, TO *out) {
  const int bs = 256;
  hipLaunchKernelGGL((k_convert<TO, TI>), dim3((n + bs - 1) / bs), dim3(bs), 0, stream(), n,
                     in, out);
}

template <typename T>
__global__ void k_fill(int64_t n, T *__restrict__ out, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = v;
}
template <typename T>
void launch_fill(int64_t n, T *out, T v) {
  const int bs = 256;
  hipLaunchKernelGGL((k_fill<T>), dim3((n + bs - 1) / bs), dim3(bs), 0, stream(), n, out, v);
}

template <typename T>
__global__ void k_sub(int64_t n, const T *__restrict__ a, const T *__restrict__ b,
                      T *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] - b[i];
}
template <typename T>
void launch_sub(int64_t n, const T *a, const T *b, T *out) {
  const int bs = 256;
  hipLaunchKernelGGL((k_sub<T>), dim3((n + bs - 1) / bs), dim3(bs), 0, stream(), n, a, b,
                     out);
}

template <typename T>
__global__ void k_add(int64_t n, const T *__restrict__ a, T *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += a[i];
}
template <typename T>
void launch_add(int64_t n, const T *a, T *out) {
  const int bs = 256;
  hipLaunchKernelGGL((k_add<T>), dim3((n + bs - 1) / bs), dim3(bs), 0, stream(), n, a, out);
}

// out[f] = in[f] for the fixed DOFs of level l (BC rows of the operators)
__global__ void k_copy_fixed(int64_t n, const int64_t *__restrict__ idx,
                             const double *__restrict__ in, double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[idx[i]] = in[idx[i]];
}
void launch_copy_fixed(const DevGrid &g, int l, const double *in, double *out) {
  const int64_t n = g.nfixed[l];
  if (n <= 0) return;
  const int bs = 256;
  hipLaunchKernelGGL(k_copy_fixed, dim3((n + bs - 1) / bs), dim3(bs), 0, stream(), n,
                     g.fixed[l], in, out);
}

template void launch_dot<double>(const double *, const double *, int64_t, double *, double *,
                                 double *);
template void launch_dot<float>(const float *, const float *, int64_t, double *, double *,
                                double *);
template void launch_convert<float, double>(int64_t, const double *, float *);
template void launch_convert<double, float>(int64_t, const float *, double *);
template void launch_convert<double, double>(int64_t, const double *, double *);
template void launch_fill<double>(int64_t, double *, double);
template void launch_fill<float>(int64_t, float *, float);
template void launch_sub<double>(int64_t, const double *, const double *, double *);
template void launch_sub<float>(int64_t, const float *, const float *, float *);
template void launch_add<double>(int64_t, const double *, double *);
template void launch_add<float>(int64_t, const float *, float *);

}  // namespace t3d
