// Assembled coarse levels on the GPU (replacing the host CSR assembly of
// stencil_assembly.c:126-260 and the PCIe round trips of
// stencil_solvers.c:172-250).
//
// Level-l element matrices are built by Galerkin recursion,
//   Ke_l(E) = sum_{children c} Q_c^T Ke_{l-1}(c) Q_c,
// with Q_c the trilinear interpolation from the parent's 8 nodes to child
// c's nodes.  Because the reference's sub-cell matrices at level l use the
// fine elements' own quadrature points (local_matrix.c:268-293), this equals
// its sum_f E_f KE_l[f] exactly in exact arithmetic.  Element matrices are
// then gathered into a 27-point node stencil with symmetric elimination of
// the fixed DOFs (stencil_assembly.c:226-259), which serves the smoother
// levels and the dense coarsest matrix.
#include <mutex>

#include <rocblas/rocblas.h>

#include "t3d_internal.hpp"

namespace t3d {

__host__ __device__ constexpr int cn_x(int n) { return (n == 1 || n == 2 || n == 5 || n == 6) ? 1 : 0; }
__host__ __device__ constexpr int cn_y(int n) { return (n == 0 || n == 1 || n == 4 || n == 5) ? 1 : 0; }
__host__ __device__ constexpr int cn_z(int n) { return n >= 4 ? 1 : 0; }
__host__ __device__ constexpr int cn_of(int ox, int oy, int oz) {
  return oz * 4 + (oy ? (ox ? 1 : 0) : (ox ? 2 : 3));
}

// weight of parent node B at child c's node b (trilinear, per direction)
__device__ __forceinline__ double qweight(int cx, int cy, int cz, int b, int B) {
  const int p[3] = {cx + cn_x(b), cy + cn_y(b), cz + cn_z(b)};  // in half units 0..2
  const int q[3] = {cn_x(B), cn_y(B), cn_z(B)};
  double w = 1.0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double t = 0.5 * p[d];
    w *= q[d] ? t : 1.0 - t;
  }
  return w;
}

// One workgroup (576 threads, one per matrix entry) per parent element.
// SRC = 0: children are level-1 elements built from fine moduli and KE_1;
// SRC = 1: children are stored level-(l-1) element matrices.  The 8 children
// are staged in LDS together and the two products run for all of them
// between two barriers.
template <int SRC>
__global__ __launch_bounds__(576) void k_galerkin(
    int nelx, int nely, int nelz,        // parent level elements
    int cnely, int cnelz,                // child level elements (SRC = 1)
    int few, int fez, const double *__restrict__ KE1, const double *__restrict__ E,
    const double *__restrict__ Kchild, double *__restrict__ Kparent) {
  __shared__ double Kall[8 * 576];  // child matrices
  __shared__ double Tm[8 * 576];    // Kc Q_c per child
  __shared__ double Wq[8 * 64];     // Q_c: weight of parent node B at child node b, [c][b][B]
  const int t = threadIdx.x;
  const int r = t / 24, s = t % 24;
  const int64_t pe = blockIdx.x;  // compact parent index ((I*nelz + K)*nely + J), 0-based
  const int J = (int)(pe % nely);
  const int K = (int)((pe / nely) % nelz);
  const int I = (int)(pe / ((int64_t)nely * nelz));
  {
    double kv[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int ci = 2 * I + cn_x(c), cj = 2 * J + cn_y(c), ck = 2 * K + cn_z(c);
      if (SRC == 0) {
        double v = 0.0;
        for (int ii = 0; ii < 2; ii++)
          for (int kk = 0; kk < 2; kk++)
            for (int jj = 0; jj < 2; jj++) {
              const int64_t fe = ((int64_t)(2 * ci + ii + 1) * fez + (2 * ck + kk + 1)) * few +
                                 (2 * cj + jj + 1);
              v += E[fe] * KE1[576 * (4 * ii + 2 * kk + jj) + t];
            }
        kv[c] = v;
      } else {
        kv[c] = Kchild[576 * (((int64_t)ci * cnelz + ck) * cnely + cj) + t];
      }
    }
#pragma unroll
    for (int c = 0; c < 8; c++) Kall[576 * c + t] = kv[c];
  }
  if (t < 512) {
    const int c = t >> 6, q = t & 63;
    Wq[t] = qweight(cn_x(c), cn_y(c), cn_z(c), q >> 3, q & 7);
  }
  __syncthreads();
  // Tm_c[a][S] = sum_b Kc[a][3b + cs] Q_c[b][Snode]
  {
    const int a = r, Sn = s / 3, cs = s % 3;
#pragma unroll 1
    for (int c = 0; c < 8; c++) {
      double v = 0.0;
#pragma unroll
      for (int b = 0; b < 8; b++)
        v += Kall[576 * c + 24 * a + 3 * b + cs] * Wq[64 * c + 8 * b + Sn];
      Tm[576 * c + t] = v;
    }
  }
  __syncthreads();
  // P[R][S] = sum_c sum_a Q_c[a][Rnode] Tm_c[3a + cr][S]
  double acc = 0.0;
  {
    const int Rn = r / 3, cr = r % 3;
#pragma unroll 1
    for (int c = 0; c < 8; c++) {
      double v = 0.0;
#pragma unroll
      for (int a = 0; a < 8; a++) v += Wq[64 * c + 8 * a + Rn] * Tm[576 * c + 24 * (3 * a + cr) + s];
      acc += v;
    }
  }
  Kparent[576 * pe + t] = acc;
}

// Level-2 element matrices as one GEMM on FP64 MFMA:
//   Ke2[e][q] = sum_{f<64} E[fine(e, f)] KE_2[f][q],  q < 576,
// with f = 16 ii + 4 kk + jj the reference's sub-cell index
// (local_matrix.c:283).  A workgroup owns 128 level-2 elements: their 64
// moduli are gathered into LDS, and each wave sweeps 9 of the 36 16-column
// tiles of KE_2, holding the tile's 64x16 slice in registers while it runs
// over the 8 16-element row tiles (v_mfma_f64_16x16x4f64, K = 64).
constexpr int kL2Rows = 128;

__global__ __launch_bounds__(256) void k_elem_level2_mfma(
    int nelx, int nely, int nelz, int few, int fez, const double *__restrict__ KE2,
    const double *__restrict__ E, double *__restrict__ Ke2) {
  __shared__ double As[kL2Rows][65];
  const int64_t nel = (int64_t)nelx * nely * nelz;
  const int64_t e0 = (int64_t)blockIdx.x * kL2Rows;
  const int tid = threadIdx.x;
  for (int q = tid; q < kL2Rows * 64; q += 256) {
    const int r = q >> 6, f = q & 63;
    const int64_t e = e0 + r;
    double v = 0.0;
    if (e < nel) {
      const int J = (int)(e % nely), K = (int)((e / nely) % nelz),
                I = (int)(e / ((int64_t)nely * nelz));
      const int ii = f >> 4, kk = (f >> 2) & 3, jj = f & 3;
      v = E[((int64_t)(4 * I + ii + 1) * fez + (4 * K + kk + 1)) * few + (4 * J + jj + 1)];
    }
    As[r][f] = v;
  }
  __syncthreads();
  const int w = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
  typedef double double4_v __attribute__((ext_vector_type(4)));
  for (int nt = w; nt < 36; nt += 4) {
    double bfrag[16];
#pragma unroll
    for (int k = 0; k < 16; k++) bfrag[k] = KE2[(4 * k + lk) * 576 + 16 * nt + li];
#pragma unroll 2
    for (int mt = 0; mt < kL2Rows / 16; mt++) {
      double4_v acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 16; k++)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(As[16 * mt + li][4 * k + lk], bfrag[k], acc,
                                                   0, 0, 0);
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int64_t e = e0 + 16 * mt + lk + 4 * v;
        if (e < nel) Ke2[e * 576 + 16 * nt + li] = acc[v];
      }
    }
  }
}

void launch_elem_level2(const DevGrid &g, const double *E, double *Ke2) {
  const Level &L = g.lv[2];
  const int64_t nel = L.elements();
  if (g.KE2) {
    hipLaunchKernelGGL(k_elem_level2_mfma, dim3((nel + kL2Rows - 1) / kL2Rows), dim3(256), 0,
                       stream(), L.nelx, L.nely, L.nelz, g.ewy, g.ewz, g.KE2, E, Ke2);
    return;
  }
  hipLaunchKernelGGL((k_galerkin<0>), dim3(nel), dim3(576), 0, stream(), L.nelx,
                     L.nely, L.nelz, 0, 0, g.ewy, g.ewz, g.KE[1], E, nullptr, Ke2);
}

// Moduli of the 512 fine elements of each level-3 element, row e (compact
// level-3 element index (I*nelz + K)*nely + J), column f = 64 ii + 8 kk + jj
// (the sub-cell order of getKEsubspace, local_matrix.c:283).  One workgroup
// per kG3Run level-3 elements consecutive in y: it reads the 64 (ii, kk)
// rows of 8 kG3Run fine moduli (contiguous in E), and writes the elements'
// rows of Eg, which are consecutive, as one contiguous run from LDS (rows
// padded to 520 so that the row-wise LDS stores of the reads spread over
// the banks).
constexpr int kG3Run = 4;
constexpr int kG3Ld = 520;

__global__ __launch_bounds__(256) void k_gather_level3(int nelx, int nely, int nelz, int few,
                                                       int fez, const double *__restrict__ E,
                                                       double *__restrict__ Eg) {
  __shared__ double buf[kG3Run * kG3Ld];
  const int nrun = (nely + kG3Run - 1) / kG3Run;
  const int64_t ik = blockIdx.x / nrun;  // I * nelz + K
  const int J0 = kG3Run * (int)(blockIdx.x % nrun);
  const int K = (int)(ik % nelz), I = (int)(ik / nelz);
  const int nj = min(kG3Run, nely - J0);
  for (int q = threadIdx.x; q < 64 * 8 * kG3Run; q += blockDim.x) {
    const int row = q / (8 * kG3Run), c = q % (8 * kG3Run);  // row = 8 ii + kk
    if (c < 8 * nj) {
      const int fi = 8 * I + (row >> 3), fk = 8 * K + (row & 7), fj = 8 * J0 + c;
      buf[(c >> 3) * kG3Ld + 8 * row + (c & 7)] =
          E[((int64_t)(fi + 1) * fez + (fk + 1)) * few + (fj + 1)];
    }
  }
  __syncthreads();
  double *dst = Eg + (ik * nely + J0) * 512;
  for (int q = threadIdx.x; q < 512 * nj; q += blockDim.x)
    dst[q] = buf[(q >> 9) * kG3Ld + (q & 511)];
}

// one rocBLAS handle per device, created on first use
static rocblas_handle blas_handle() {
  static std::mutex mu;
  static rocblas_handle handles[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (!handles[dev] && rocblas_create_handle(&handles[dev]) != rocblas_status_success)
    handles[dev] = nullptr;
  return handles[dev];
}

bool launch_elem_level3(const DevGrid &g, const double *E, double *Eg, double *Ke3) {
  const Level &L = g.lv[3];
  const int64_t nel = L.elements();
  if (!g.KE3 || nel <= 0 || nel > (int64_t)INT32_MAX / 576) return false;
  rocblas_handle h = blas_handle();
  if (!h || rocblas_set_stream(h, stream()) != rocblas_status_success) return false;
  const int64_t nwg = (int64_t)L.nelx * L.nelz * ((L.nely + kG3Run - 1) / kG3Run);
  hipLaunchKernelGGL(k_gather_level3, dim3((unsigned)nwg), dim3(256), 0, stream(), L.nelx,
                     L.nely, L.nelz, g.ewy, g.ewz, E, Eg);
  // row-major Ke3 [nel x 576] = Eg [nel x 512] KE_3 [512 x 576], i.e. in
  // column-major terms Ke3^T (576 x nel) = KE_3^T (576 x 512) Eg^T (512 x nel)
  const double one = 1.0, zero = 0.0;
  return rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, 576, (rocblas_int)nel,
                       512, &one, g.KE3, 576, Eg, 512, &zero, Ke3, 576) ==
         rocblas_status_success;
}

void launch_elem_galerkin(const DevGrid &g, int l, const double *Kchild, double *Kparent) {
  const Level &L = g.lv[l], &C = g.lv[l - 1];
  hipLaunchKernelGGL((k_galerkin<1>), dim3(L.elements()), dim3(576), 0, stream(), L.nelx,
                     L.nely, L.nelz, C.nely, C.nelz, g.ewy, g.ewz, nullptr, nullptr, Kchild,
                     Kparent);
}

// 27-point stencil from element matrices, one thread per (node, neighbour):
// threadIdx.x = neighbour nb, kNsNodes consecutive nodes per workgroup, so
// that the threads of a node read the same three rows of each of its
// element matrices; the workgroup's 243 x kNsNodes values are staged in LDS
// and written as kNsNodes-long runs of each S plane.
// S[(nb*9 + r*3 + c) * nint + n]; symmetric elimination of fixed DOFs.
// bc_symmetric = 0 keeps the columns of fixed DOFs and only replaces fixed
// rows by identity rows (the matrix-free convention, stencil_methods.c:515-518).
// S32: an FP32 copy written alongside when given.
// HALF: the symmetric half layout of a matrix with symmetric elimination
// (kHalfPlanes planes): the 13 neighbours nb = 14..26 as full 3x3 blocks at
// planes (nb-14)*9 + 3r + c, the centre's upper triangle r <= c at planes
// 117 + half_centre_index(r, c); one thread per stored neighbour.
constexpr int kNsNodes = 32;

template <bool HALF>
__global__ __launch_bounds__(27 * kNsNodes) void k_node_stencil(
    int nelx, int nely, int nelz, int wy, int wz, const double *__restrict__ Ke,
    const uint8_t *__restrict__ fix, double *__restrict__ S, float *__restrict__ S32,
    int bc_symmetric) {
  constexpr int kPlanes = HALF ? kHalfPlanes : 243;
  __shared__ double tile[kPlanes * kNsNodes];
  const int64_t nx = nelx + 1, ny = nely + 1, nz = nelz + 1, nint = nx * ny * nz;
  const int nb = HALF ? 13 + threadIdx.x : threadIdx.x;
  const int64_t n0 = (int64_t)blockIdx.x * kNsNodes;
  const int64_t n = n0 + threadIdx.y;
  const int nloc = (int)min((int64_t)kNsNodes, nint - n0);
  if (n < nint) {
    const int j = (int)(n % ny), k = (int)((n / ny) % nz), i = (int)(n / (ny * nz));
    const int dx = nb / 9 - 1, dz = (nb / 3) % 3 - 1, dy = nb % 3 - 1;
    const int i2 = i + dx, j2 = j + dy, k2 = k + dz;
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const bool nb_in = i2 >= 0 && i2 < nx && j2 >= 0 && j2 < ny && k2 >= 0 && k2 < nz;
    if (nb_in) {
      for (int e = 0; e < 8; e++) {  // elements (0-based) containing node n
        const int ex = (e & 1) ? 0 : -1, ey = (e & 2) ? 0 : -1, ez = (e & 4) ? 0 : -1;
        const int ie = i + ex, je = j + ey, ke = k + ez;
        if (ie < 0 || ie >= nelx || je < 0 || je >= nely || ke < 0 || ke >= nelz) continue;
        const int ox = dx - ex, oy = dy - ey, oz = dz - ez;  // neighbour offset in element
        if (ox < 0 || ox > 1 || oy < 0 || oy > 1 || oz < 0 || oz > 1) continue;
        const int a = cn_of(-ex, -ey, -ez), b = cn_of(ox, oy, oz);
        const double *K = Ke + 576 * (((int64_t)ie * nelz + ke) * nely + je);
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int c = 0; c < 3; c++) v[3 * r + c] += K[24 * (3 * a + r) + 3 * b + c];
      }
      const uint8_t m1 = fix[((int64_t)(i + 1) * wz + (k + 1)) * wy + (j + 1)];
      const uint8_t m2 = fix[((int64_t)(i2 + 1) * wz + (k2 + 1)) * wy + (j2 + 1)];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
          const bool fr = (m1 >> r) & 1, fc = bc_symmetric && ((m2 >> c) & 1);
          if (fr || fc) v[3 * r + c] = (nb == 13 && r == c) ? 1.0 : 0.0;
        }
    }
    if (!HALF) {
#pragma unroll
      for (int q = 0; q < 9; q++) tile[(nb * 9 + q) * kNsNodes + threadIdx.y] = v[q];
    } else if (nb > 13) {
#pragma unroll
      for (int q = 0; q < 9; q++) tile[((nb - 14) * 9 + q) * kNsNodes + threadIdx.y] = v[q];
    } else {
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = r; c < 3; c++)
          tile[(117 + half_centre_index(r, c)) * kNsNodes + threadIdx.y] = v[3 * r + c];
    }
  }
  __syncthreads();
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  for (int e = tid; e < kPlanes * kNsNodes; e += blockDim.x * kNsNodes) {
    const int q = e / kNsNodes, nn = e % kNsNodes;
    if (nn < nloc) {
      const double val = tile[e];
      S[(int64_t)q * nint + n0 + nn] = val;
      if (S32) S32[(int64_t)q * nint + n0 + nn] = (float)val;
    }
  }
}

void launch_node_stencil(const DevGrid &g, int l, const double *Ke, double *S,
                         int bc_symmetric, float *S32, bool half) {
  const Level &L = g.lv[l];
  const int64_t nint = L.interior_nodes();
  const dim3 grid((unsigned)((nint + kNsNodes - 1) / kNsNodes));
  if (half)
    hipLaunchKernelGGL((k_node_stencil<true>), grid, dim3(14, kNsNodes), 0, stream(), L.nelx,
                       L.nely, L.nelz, L.wy, L.wz, Ke, g.fixmask[l], S, S32, 1);
  else
    hipLaunchKernelGGL((k_node_stencil<false>), grid, dim3(27, kNsNodes), 0, stream(), L.nelx,
                       L.nely, L.nelz, L.wy, L.wz, Ke, g.fixmask[l], S, S32, bc_symmetric);
}

// the assembled coarsest level into the coarse factor's np x np storage
// (row-major; logical DOF d at storage index pos[d])
__global__ void k_stencil_to_dense(int nx, int ny, int nz, const double *__restrict__ S,
                                   const int *__restrict__ pos, double *__restrict__ A,
                                   int64_t np) {
  const int64_t nint = (int64_t)nx * ny * nz;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= 27 * nint) return;
  const int nb = (int)(tid / nint);
  const int64_t p = tid % nint;
  const int j = (int)(p % ny), k = (int)((p / ny) % nz), i = (int)(p / ((int64_t)ny * nz));
  const int i2 = i + nb / 9 - 1, k2 = k + (nb / 3) % 3 - 1, j2 = j + nb % 3 - 1;
  if (i2 < 0 || i2 >= nx || j2 < 0 || j2 >= ny || k2 < 0 || k2 >= nz) return;
  const int64_t q = ((int64_t)i2 * nz + k2) * ny + j2;
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++)
      A[(int64_t)pos[3 * p + r] * np + pos[3 * q + c]] =
          S[(int64_t)(nb * 9 + 3 * r + c) * nint + p];
}

void launch_stencil_to_dense(const DevGrid &g, int l, const double *S, CoarseFactor &cf) {
  const Level &L = g.lv[l];
  cf.begin_assembly();
  const int64_t nthreads = 27 * L.interior_nodes();
  const int bs = 256;
  hipLaunchKernelGGL(k_stencil_to_dense, dim3((nthreads + bs - 1) / bs), dim3(bs), 0,
                     stream(), L.nx, L.ny, L.nz, S, cf.pos, cf.A, (int64_t)cf.np);
}

}  // namespace t3d
