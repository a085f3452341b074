/*
 * top3d_oracle.c -- CPU oracle for the MGCG state-solve path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference
 * (topopt/TopOpt-in-OpenMP-GPU) used as the checker for the HIP library.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library never does.  The reference itself is not
 * buildable in this image (it needs cblas.h and cholmod.h, which are absent),
 * so this restatement is the oracle, and it is pinned by independent checks
 * in tests/test_oracle_*.py:
 *
 *   or_ke            independent 8-point tensor rule with the same 1-D moments
 *                    as the 10-point midpoint rule; symmetry; rigid-body null
 *                    space; sub-cell sum equals the Galerkin product
 *   or_matvec0       scipy element-by-element assembly (BC rows = identity)
 *   or_matvec_l      Galerkin identity K_l = R K_{l-1} P (pre-BC), random rho
 *   or_prolong/restrict  P reproduces trilinear fields; R = P^T (adjointness)
 *   or_invdiag       diagonal of the scipy-assembled operator
 *   or_csr_*         scipy assembly with symmetric BC elimination
 *   or_coarse_solve  numpy dense solve
 *   or_vcycle        numpy composition of the independently pinned operators
 *   or_solve         residual test and scipy direct solve of the BC system
 *   or_sensitivity   c = u^T K u with the scipy operator; central FD of c(x)
 *                    with direct solves
 *   or_filter(_grad) explicit scipy filter matrix H; adjointness
 *   or_oc            the OC optimality conditions: one multiplier for every
 *                    unclamped element (fixes the exponent 1/2), clamped
 *                    elements beyond their bound, volume balance to the
 *                    bisection's resolution, a uniform closed form; copies
 *                    built with a wrong move limit, exponent, bisection
 *                    direction or a dropped volume term fail them
 *   or_volume        numpy sum over the interior index set (halo ignored)
 *   or_change        numpy max |x - xnew| in float, and the copy into x
 *   or_load          total force, node set, count
 *   or_top3d         composition of the above (trajectory fixtures are
 *                    regression data from this file, not independent)
 *   or_write_vtu     XML parse; point count and integer coordinates; every
 *                    cell is the unit cube at its element corner; the rigid
 *                    motions of each cell's node list (in file order) are in
 *                    the null space of KE_0; offsets, types; densities read
 *                    back in element order
 *
 * Citations are to files of the reference at the top level of its tree.
 */
#include "top3d_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OR_MAX(a, b) ((a) > (b) ? (a) : (b))
#define OR_MIN(a, b) ((a) > (b) ? (b) : (a))

int or_sizeof_ctype(void) { return (int)sizeof(or_ctype); }

/* ------------------------------------------------------------------------ */
/* Padding and halo layout: computePadding, gpu_definitions.c:67-92.  Node
 * (i,j,k) in halo coordinates lives at i*wrapy*wrapz + k*wrapy + j, DOF
 * 3*node+comp; real nodes are 1..nel+1 in each direction.                   */
void or_padding(const or_grid *g, int l, int *ncell, int *wrapx, int *wrapy,
                int *wrapz, int64_t *ndof) {
  const int nc = 1 << l;
  const int nelx = g->nelx / nc, nely = g->nely / nc, nelz = g->nelz / nc;
  const int a = g->yalign;
  const int pady = (a - ((nely + 1) % a)) % a;
  *ncell = nc;
  *wrapx = nelx + 3;
  *wrapy = nely + pady + 3;
  *wrapz = nelz + 3;
  *ndof = 3 * (int64_t)(*wrapx) * (*wrapy) * (*wrapz);
}

/* Fixed DOFs, setFixedDof_halo, stencil_grid_utility.c:7-65.
 * support 0, the "new cantilever" (:34-57): node plane i=1 (the x=0 face),
 * all z, and y in [1, nely_l/4+1] and [nely_l+2-(nely_l/4+1), nely_l+1].
 * support 1, the "classic cantilever" left in comments there (:19-32): every
 * node of plane i=1, k outer and j inner.                                   */
static void fixed_dofs(or_grid *g, int l) {
  int ncell, wx, wy, wz;
  int64_t ndof;
  or_padding(g, l, &ncell, &wx, &wy, &wz, &ndof);
  const int nelyc = g->nely / ncell, nelzc = g->nelz / ncell;
  const int nyc = nelyc + 1, nzc = nelzc + 1;
  if (g->support == 1) {
    g->nfixed[l] = 3 * nyc * nzc;
    g->fixed[l] = (int64_t *)malloc(sizeof(int64_t) * g->nfixed[l]);
    int off = 0;
    for (int64_t k = 1; k <= nzc; k++)
      for (int64_t j = 1; j <= nyc; j++)
        for (int c = 0; c < 3; c++)
          g->fixed[l][off++] = 3 * ((int64_t)wy * wz + wy * k + j) + c;
    return;
  }
  const int nodelimit = nelyc / 4 + 1;
  g->nfixed[l] = 3 * nzc * 2 * nodelimit;
  g->fixed[l] = (int64_t *)malloc(sizeof(int64_t) * g->nfixed[l]);
  int off = 0;
  const int64_t i = 1;
  for (int64_t k = 1; k < nzc + 1; k++) {
    for (int64_t j = 1; j < nodelimit + 1; j++)
      for (int c = 0; c < 3; c++)
        g->fixed[l][off++] = 3 * (i * wy * wz + wy * k + j) + c;
    for (int64_t j = (nyc + 1) - nodelimit; j < nyc + 1; j++)
      for (int c = 0; c < 3; c++)
        g->fixed[l][off++] = 3 * (i * wy * wz + wy * k + j) + c;
  }
}

int or_nfixed(const or_grid *g, int l) { return g->nfixed[l]; }
void or_fixed(const or_grid *g, int l, int64_t *out) {
  memcpy(out, g->fixed[l], sizeof(int64_t) * g->nfixed[l]);
}

/* ------------------------------------------------------------------------ */
/* Element stiffness: getC local_matrix.c:7-29, getB :33-247,
 * getKEsubspace :251-319.  Q1 hexahedron of half-size a=b=c=0.5*2^l,
 * integrated per sub-cell with an npts^3 midpoint rule (the reference uses
 * npts=10), eta evaluated at -y (local_matrix.c:298).                      */
static void constitutive(double C[6][6], double nu) {
  const double t1 = (1.0 - nu) / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double t2 = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double t3 = 1.0 / (2.0 * (1.0 + nu));
  memset(C, 0, sizeof(double) * 36);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) C[i][j] = (i == j) ? t1 : t2;
  C[3][3] = C[4][4] = C[5][5] = t3;
}

/* Corner signs of the iso-parametric nodes n1..n8 (local_matrix.c:45-73). */
static const int corner[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1},
                                 {-1, 1, -1},  {-1, -1, 1}, {1, -1, 1},
                                 {1, 1, 1},    {-1, 1, 1}};

/* Strain-displacement matrix at an iso point for an axis-aligned brick of
 * half sizes (a,b,c): rows (xx, yy, zz, xy, yz, xz), columns 3*node+comp.   */
static void strain_matrix(double B[6][24], double *jdet, const double iso[3],
                          double a, double b, double c) {
  double dN[8][3];
  for (int n = 0; n < 8; n++) {
    const double s0 = corner[n][0], s1 = corner[n][1], s2 = corner[n][2];
    dN[n][0] = 0.125 * s0 * (1 + s1 * iso[1]) * (1 + s2 * iso[2]);
    dN[n][1] = 0.125 * s1 * (1 + s0 * iso[0]) * (1 + s2 * iso[2]);
    dN[n][2] = 0.125 * s2 * (1 + s0 * iso[0]) * (1 + s1 * iso[1]);
  }
  /* The Jacobian of an axis-aligned brick is diag(a,b,c). */
  *jdet = a * b * c;
  const double ia = (1.0 / (*jdet)) * (b * c), ib = (1.0 / (*jdet)) * (a * c),
               ic = (1.0 / (*jdet)) * (a * b);
  memset(B, 0, sizeof(double) * 6 * 24);
  for (int n = 0; n < 8; n++) {
    const double dx = ia * dN[n][0], dy = ib * dN[n][1], dz = ic * dN[n][2];
    B[0][3 * n + 0] = dx;
    B[1][3 * n + 1] = dy;
    B[2][3 * n + 2] = dz;
    B[3][3 * n + 0] = dy;
    B[3][3 * n + 1] = dx;
    B[4][3 * n + 1] = dz;
    B[4][3 * n + 2] = dy;
    B[5][3 * n + 0] = dz;
    B[5][3 * n + 2] = dx;
  }
}

void or_ke(double nu, int l, int npts, double *KE) {
  const int ncell = 1 << l;
  const double a = 0.5 * ncell;
  double C[6][6];
  constitutive(C, nu);
  const double spacing = 2.0 / (double)ncell / (double)npts;
  const double vol = spacing * spacing * spacing;
  memset(KE, 0, sizeof(double) * 576 * ncell * ncell * ncell);
#pragma omp parallel for collapse(3) schedule(dynamic)
  for (int ii = 0; ii < ncell; ii++)
    for (int kk = 0; kk < ncell; kk++)
      for (int jj = 0; jj < ncell; jj++) {
        double *K = KE + 576 * (ncell * ncell * ii + ncell * kk + jj);
        double igp = -1.0 + spacing / 2.0 + 2.0 / (double)ncell * ii;
        for (int p = 0; p < npts; p++) {
          double jgp = -1.0 + spacing / 2.0 + 2.0 / (double)ncell * jj;
          for (int q = 0; q < npts; q++) {
            double kgp = -1.0 + spacing / 2.0 + 2.0 / (double)ncell * kk;
            for (int r = 0; r < npts; r++) {
              const double iso[3] = {igp, -1.0 * jgp, kgp};
              double B[6][24], jdet, BtC[24][6];
              strain_matrix(B, &jdet, iso, a, a, a);
              for (int i = 0; i < 24; i++)
                for (int j = 0; j < 6; j++) {
                  double s = 0.0;
                  for (int k = 0; k < 6; k++) s += B[k][i] * C[k][j];
                  BtC[i][j] = s;
                }
              const double w = jdet * vol;
              for (int i = 0; i < 24; i++)
                for (int j = 0; j < 24; j++) {
                  double s = 0.0;
                  for (int k = 0; k < 6; k++) s += BtC[i][k] * B[k][j];
                  K[24 * i + j] = w * s + K[24 * i + j];
                }
              kgp += spacing;
            }
            jgp += spacing;
          }
          igp += spacing;
        }
      }
}

void or_ke_copy(const or_grid *g, int l, double *out) {
  const int n = 1 << l;
  memcpy(out, g->KE[l], sizeof(double) * 576 * n * n * n);
}

/* setupGC, stencil_grid_utility.c:67-108. */
or_grid *or_create(int nelx, int nely, int nelz, int nl, int yalign) {
  if (nl < 3 || nl > OR_MAXL) return NULL;
  const int f = 1 << (nl - 1);
  if (nelx % f || nely % f || nelz % f) return NULL;
  or_grid *g = (or_grid *)calloc(1, sizeof(or_grid));
  g->nelx = nelx;
  g->nely = nely;
  g->nelz = nelz;
  g->nl = nl;
  g->yalign = yalign > 0 ? yalign : 1;
  g->E0 = 1.0;
  g->Emin = 1e-6;
  g->nu = 0.3;
  for (int l = 0; l < nl; l++) {
    const int n = 1 << l;
    g->KE[l] = (double *)malloc(sizeof(double) * 576 * n * n * n);
    or_ke(g->nu, l, 10, g->KE[l]);
    fixed_dofs(g, l);
  }
  return g;
}

/* Extension beyond the reference: with bcsym set, the matrix-free levels
 * also drop the columns of fixed DOFs (the input is read as zero there),
 * which makes the V-cycle symmetric like the assembled levels.  The
 * reference's identity-row convention (stencil_methods.c:424-427, 515-518)
 * makes PCG stagnate on strongly heterogeneous densities.                   */
void or_set_bcsym(or_grid *g, int on) { g->bcsym = on ? 1 : 0; }

void or_set_support(or_grid *g, int mode) {
  g->support = mode == 1 ? 1 : 0;
  for (int l = 0; l < g->nl; l++) {
    free(g->fixed[l]);
    fixed_dofs(g, l);
  }
}

static const or_ctype *masked_input(const or_grid *g, int l, const or_ctype *in,
                                    int64_t ndof, or_ctype **buf) {
  *buf = NULL;
  if (!g->bcsym) return in;
  *buf = (or_ctype *)malloc(sizeof(or_ctype) * ndof);
  memcpy(*buf, in, sizeof(or_ctype) * ndof);
  for (int f = 0; f < g->nfixed[l]; f++) (*buf)[g->fixed[l][f]] = 0;
  return *buf;
}

void or_destroy(or_grid *g) {
  if (!g) return;
  for (int l = 0; l < g->nl; l++) {
    free(g->KE[l]);
    free(g->fixed[l]);
  }
  free(g);
}

/* Line loads of the MATLAB specification (top3dmgcg_reference.m:57-59) on
 * the edge x = nelx, z = 0, i.e. the halo nodes (nelx+1, jj+1, 1) for
 * jj = 0..nely: F_z = -w_jj with w = 1 (profile 0, "constant load"), w = 1
 * but 1/2 at the two ends (profile 1), or w = sin(pi jj / nely) (profile 2,
 * "sine load").                                                             */
void or_line_load_profile(const or_grid *g, or_ctype *F, int profile) {
  int ncell, wx, wy, wz;
  int64_t ndof;
  or_padding(g, 0, &ncell, &wx, &wy, &wz, &ndof);
  memset(F, 0, sizeof(or_ctype) * ndof);
  const int64_t i = g->nelx + 1, k = 1, ny = g->nely;
  for (int64_t jj = 0; jj <= ny; jj++) {
    double w = 1.0;
    if (profile == 1 && (jj == 0 || jj == ny)) w = 0.5;
    if (profile == 2) w = sin(M_PI * (double)jj / (double)ny);
    F[3 * (i * wy * wz + k * wy + jj + 1) + 2] = (or_ctype)(-w);
  }
}

void or_line_load(const or_grid *g, or_ctype *F) { or_line_load_profile(g, F, 0); }

/* Cantilever load, stencil_optimization.c:39-95: F_z = -1/(4*count) on the 4
 * z=0-face nodes of each element (i,j) inside the disc around
 * (nelx, (nely-1)/2) of radius (nely+1)/5.                                  */
void or_load(const or_grid *g, or_ctype *F) {
  int ncell, wx, wy, wz;
  int64_t ndof;
  or_padding(g, 0, &ncell, &wx, &wy, &wz, &ndof);
  memset(F, 0, sizeof(or_ctype) * ndof);
  const int nelx = g->nelx, nely = g->nely;
  const double radius = ((double)(nely + 1)) / 5.0;
  const double r2 = radius * radius;
  const double cx = (double)nelx, cy = ((double)nely - 1.0) / 2.0;
  int count = 0;
  for (int i = 0; i < nelx; i++)
    for (int j = 0; j < nely; j++) {
      const double dx = i - cx, dy = j - cy;
      if (dx * dx + dy * dy < r2) count++;
    }
  const double f = -1.0 / (4.0 * (double)count);
  const int64_t kk = 1;
  for (int i = 0; i < nelx; i++)
    for (int j = 0; j < nely; j++) {
      const double dx = i - cx, dy = j - cy;
      if (dx * dx + dy * dy < r2) {
        const int64_t ii = i + 1, jj = j + 1;
        const int64_t n1 = (ii + 1) * wy * wz + wy * kk + (jj + 1);
        const int64_t n2 = (ii + 1) * wy * wz + wy * kk + jj;
        const int64_t n3 = ii * wy * wz + wy * kk + (jj + 1);
        const int64_t n4 = ii * wy * wz + wy * kk + jj;
        F[3 * n1 + 2] += f;
        F[3 * n2 + 2] += f;
        F[3 * n3 + 2] += f;
        F[3 * n4 + 2] += f;
      }
    }
}

/* ------------------------------------------------------------------------ */
/* Local node numbering of an element (stencil_utility.h:8-52): node
 * 0:(i,j+1,k) 1:(i+1,j+1,k) 2:(i+1,j,k) 3:(i,j,k), 4..7 the same at k+1.    */
static const int lnode[8][3] = {{0, 1, 0}, {1, 1, 0}, {1, 0, 0}, {0, 0, 0},
                                {0, 1, 1}, {1, 1, 1}, {1, 0, 1}, {0, 0, 1}};

static inline int local_index(int ox, int oy, int oz) {
  /* getLocalNodeIndex, stencil_methods.h:31-60 */
  for (int n = 0; n < 8; n++)
    if (lnode[n][0] == ox && lnode[n][1] == oy && lnode[n][2] == oz) return n;
  return -1;
}

static inline double young(const or_grid *g, or_dtype xe) {
  /* E = Emin + x^3 (E0-Emin), x^3 in DTYPE (stencil_methods.h:105-106) */
  const or_dtype x3 = xe * xe * xe;
  return g->Emin + x3 * (g->E0 - g->Emin);
}

/* Fine matrix-free operator: applyStateOperator_stencil,
 * stencil_methods.c:7-428 with the spoke kernel stencil_methods.h:62-182.
 * out(node) = sum over the 8 elements around the node and their 8 nodes of
 * E_e * KE[recv][send] * in(send); halo elements have scale 0; padded y
 * nodes are zeroed; fixed DOFs get out = in.                                */
void or_matvec0(const or_grid *g, const or_dtype *x, const or_ctype *in,
                or_ctype *out) {
  int ncell, wx, wy, wz;
  int64_t ndof;
  or_padding(g, 0, &ncell, &wx, &wy, &wz, &ndof);
  const int nx = g->nelx + 1, ny = g->nely + 1, nz = g->nelz + 1;
  const double *KE = g->KE[0];
  const int64_t ewy = wy - 1, ewz = wz - 1;
  or_ctype *mbuf;
  const or_ctype *inm = masked_input(g, 0, in, ndof, &mbuf);
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 1; i < nx + 1; i++)
    for (int k = 1; k < nz + 1; k++)
      for (int j = 1; j < ny + 1; j++) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int ex = -1; ex <= 0; ex++)
          for (int ey = -1; ey <= 0; ey++)
            for (int ez = -1; ez <= 0; ez++) {
              const int ie = i + ex, je = j + ey, ke = k + ez;
              double scale =
                  young(g, x[ie * ewy * ewz + ke * ewy + je]);
              if (ie == 0 || ie > g->nelx || je == 0 || je > g->nely ||
                  ke == 0 || ke > g->nelz)
                scale = 0.0;
              const int recv = local_index(-ex, -ey, -ez);
              for (int s = 0; s < 8; s++) {
                const int si = ie + lnode[s][0], sj = je + lnode[s][1],
                          sk = ke + lnode[s][2];
                const int64_t sn = 3 * ((int64_t)si * wy * wz + sk * wy + sj);
                const double vx = inm[sn], vy = inm[sn + 1], vz = inm[sn + 2];
                for (int c = 0; c < 3; c++) {
                  const double *row = KE + 24 * (3 * recv + c) + 3 * s;
                  double t = 0.0;
                  t += row[0] * vx;
                  t += row[1] * vy;
                  t += row[2] * vz;
                  acc[c] += scale * t;
                }
              }
            }
        const int64_t o = 3 * ((int64_t)i * wy * wz + k * wy + j);
        out[o] = (or_ctype)acc[0];
        out[o + 1] = (or_ctype)acc[1];
        out[o + 2] = (or_ctype)acc[2];
      }
  /* padded y nodes, stencil_methods.c:407-418 */
  for (int i = 0; i < wx; i++)
    for (int k = 0; k < wz; k++)
      for (int j = ny + 1; j < wy; j++) {
        const int64_t o = 3 * ((int64_t)i * wy * wz + k * wy + j);
        out[o] = out[o + 1] = out[o + 2] = 0;
      }
  /* BC rows, stencil_methods.c:424-427 */
  for (int f = 0; f < g->nfixed[0]; f++) out[g->fixed[0][f]] = in[g->fixed[0][f]];
  free(mbuf);
}

/* Level-l matrix-free operator: applyStateOperatorSubspace_halo,
 * stencil_methods.c:436-519.  Coarse element = sum over its 8^l fine
 * elements of E_fine * KE_l[subcell]; fixed DOFs get out = in.              */
void or_matvec_l(const or_grid *g, int l, const or_dtype *x,
                 const or_ctype *in, or_ctype *out) {
  int ncell, wx, wy, wz, fwx, fwy, fwz;
  int64_t ndof, fndof;
  or_padding(g, l, &ncell, &wx, &wy, &wz, &ndof);
  or_padding(g, 0, &ncell, &fwx, &fwy, &fwz, &fndof);
  ncell = 1 << l;
  const int nelxc = g->nelx / ncell, nelyc = g->nely / ncell,
            nelzc = g->nelz / ncell;
  const double *KE = g->KE[l];
  const int64_t ewy = fwy - 1, ewz = fwz - 1;
  or_ctype *mbuf;
  const or_ctype *inm = masked_input(g, l, in, ndof, &mbuf);
  for (int64_t i = 0; i < ndof; i++) out[i] = 0;
  /* 8-colour element loop (stencil_methods.c:457-459): elements of one
   * colour share no node, so each colour is a parallel loop */
  for (int bx = 0; bx < 2; bx++)
    for (int bz = 0; bz < 2; bz++)
      for (int by = 0; by < 2; by++)
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 1 + bx; i < nelxc + 1; i += 2)
    for (int k = 1 + bz; k < nelzc + 1; k += 2)
      for (int j = 1 + by; j < nelyc + 1; j += 2) {
        int64_t edof[8];
        for (int n = 0; n < 8; n++)
          edof[n] = 3 * ((int64_t)(i + lnode[n][0]) * wy * wz +
                         (k + lnode[n][2]) * wy + (j + lnode[n][1]));
        for (int row = 0; row < 24; row++) {
          double acc = 0.0;
          for (int n = 0; n < 8; n++)
            for (int ii = 0; ii < ncell; ii++)
              for (int kk = 0; kk < ncell; kk++)
                for (int jj = 0; jj < ncell; jj++) {
                  const int64_t ifn = (int64_t)(i - 1) * ncell + ii + 1;
                  const int64_t jfn = (int64_t)(j - 1) * ncell + jj + 1;
                  const int64_t kfn = (int64_t)(k - 1) * ncell + kk + 1;
                  const double *Kc =
                      KE + 576 * (ncell * ncell * ii + ncell * kk + jj);
                  const double sc = young(g, x[ifn * ewy * ewz + kfn * ewy + jfn]);
                  for (int c = 0; c < 3; c++)
                    acc += sc * Kc[24 * row + 3 * n + c] * (double)inm[edof[n] + c];
                }
          out[edof[row / 3] + row % 3] += (or_ctype)acc;
        }
      }
  for (int f = 0; f < g->nfixed[l]; f++) out[g->fixed[l][f]] = in[g->fixed[l][f]];
  free(mbuf);
}

/* Trilinear prolongation from level l+1 to level l:
 * projectToFinerGrid_halo, stencil_methods.c:521-590.                      */
void or_prolong(const or_grid *g, int l, const or_ctype *coarse,
                or_ctype *fine) {
  int nc, fwx, fwy, fwz, cwx, cwy, cwz;
  int64_t fnd, cnd;
  or_padding(g, l, &nc, &fwx, &fwy, &fwz, &fnd);
  or_padding(g, l + 1, &nc, &cwx, &cwy, &cwz, &cnd);
  const int ncf = 1 << l;
  const int nxf = g->nelx / ncf + 1, nyf = g->nely / ncf + 1,
            nzf = g->nelz / ncf + 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 1; i < nxf + 1; i++)
    for (int k = 1; k < nzf + 1; k++)
      for (int j = 1; j < nyf + 1; j++) {
        const int64_t fi = (int64_t)i * fwy * fwz + k * fwy + j;
        const int64_t ic[2] = {(i - 1) / 2 + 1, i / 2 + 1};
        const int64_t jc[2] = {(j - 1) / 2 + 1, j / 2 + 1};
        const int64_t kc[2] = {(k - 1) / 2 + 1, k / 2 + 1};
        for (int c = 0; c < 3; c++) {
          double s = 0.0;
          for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
              for (int d = 0; d < 2; d++)
                s += 0.125 * coarse[3 * (ic[a] * cwy * cwz + kc[d] * cwy + jc[b]) + c];
          fine[3 * fi + c] = (or_ctype)s;
        }
      }
}

/* Full-weighting restriction from level l to level l+1:
 * projectToCoarserGrid_halo, stencil_methods.c:595-688; weights 1, 1/2,
 * 1/4, 1/8 by squared distance over the 3x3x3 fine patch.                   */
void or_restrict(const or_grid *g, int l, const or_ctype *fine,
                 or_ctype *coarse) {
  int nc, fwx, fwy, fwz, cwx, cwy, cwz;
  int64_t fnd, cnd;
  or_padding(g, l, &nc, &fwx, &fwy, &fwz, &fnd);
  or_padding(g, l + 1, &nc, &cwx, &cwy, &cwz, &cnd);
  const int ncc = 1 << (l + 1);
  const int nxc = g->nelx / ncc + 1, nyc = g->nely / ncc + 1,
            nzc = g->nelz / ncc + 1;
  const double vals[4] = {1.0, 0.5, 0.25, 0.125};
#pragma omp parallel for collapse(2) schedule(static)
  for (int ic = 1; ic < nxc + 1; ic++)
    for (int kc = 1; kc < nzc + 1; kc++)
      for (int jc = 1; jc < nyc + 1; jc++) {
        const int64_t ci = (int64_t)ic * cwy * cwz + kc * cwy + jc;
        const int nx1 = (ic - 1) * 2 + 1, ny1 = (jc - 1) * 2 + 1,
                  nz1 = (kc - 1) * 2 + 1;
        or_ctype acc[3] = {0, 0, 0};
        for (int fi = nx1 - 1; fi < nx1 + 2; fi++)
          for (int fk = nz1 - 1; fk < nz1 + 2; fk++)
            for (int fj = ny1 - 1; fj < ny1 + 2; fj++) {
              const int64_t fn = (int64_t)fi * fwy * fwz + fk * fwy + fj;
              const int d2 = (nx1 - fi) * (nx1 - fi) + (ny1 - fj) * (ny1 - fj) +
                             (nz1 - fk) * (nz1 - fk);
              for (int c = 0; c < 3; c++) {
                const or_ctype prod = (or_ctype)(vals[d2] * fine[3 * fn + c]);
                acc[c] += prod;
              }
            }
        for (int c = 0; c < 3; c++) coarse[3 * ci + c] = acc[c];
      }
}

/* Inverse diagonal of K_l: assembleInvertedMatrixDiagonalSubspace_halo,
 * stencil_methods.c:693-773.  Fixed DOFs get diagonal 1; interior nodes are
 * inverted, the halo keeps 0.                                               */
void or_invdiag(const or_grid *g, const or_dtype *x, int l, double *diag) {
  int ncell, wx, wy, wz, fwx, fwy, fwz;
  int64_t ndof, fndof;
  or_padding(g, l, &ncell, &wx, &wy, &wz, &ndof);
  or_padding(g, 0, &ncell, &fwx, &fwy, &fwz, &fndof);
  ncell = 1 << l;
  const int nelxc = g->nelx / ncell, nelyc = g->nely / ncell,
            nelzc = g->nelz / ncell;
  const double *KE = g->KE[l];
  const int64_t ewy = fwy - 1, ewz = fwz - 1;
  for (int64_t i = 0; i < ndof; i++) diag[i] = 0.0;
  for (int bx = 0; bx < 2; bx++)
    for (int bz = 0; bz < 2; bz++)
      for (int by = 0; by < 2; by++)
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 1 + bx; i < nelxc + 1; i += 2)
    for (int k = 1 + bz; k < nelzc + 1; k += 2)
      for (int j = 1 + by; j < nelyc + 1; j += 2)
        for (int ii = 0; ii < ncell; ii++)
          for (int kk = 0; kk < ncell; kk++)
            for (int jj = 0; jj < ncell; jj++) {
              const int64_t ifn = (int64_t)(i - 1) * ncell + ii + 1;
              const int64_t jfn = (int64_t)(j - 1) * ncell + jj + 1;
              const int64_t kfn = (int64_t)(k - 1) * ncell + kk + 1;
              const double sc = young(g, x[ifn * ewy * ewz + kfn * ewy + jfn]);
              const double *Kc = KE + 576 * (ncell * ncell * ii + ncell * kk + jj);
              for (int n = 0; n < 8; n++) {
                const int64_t e = 3 * ((int64_t)(i + lnode[n][0]) * wy * wz +
                                       (k + lnode[n][2]) * wy + (j + lnode[n][1]));
                for (int c = 0; c < 3; c++)
                  diag[e + c] += sc * Kc[(3 * n + c) * 24 + 3 * n + c];
              }
            }
  for (int f = 0; f < g->nfixed[l]; f++) diag[g->fixed[l][f]] = 1.0;
  for (int i = 1; i < nelxc + 2; i++)
    for (int k = 1; k < nelzc + 2; k++)
      for (int j = 1; j < nelyc + 2; j++) {
        const int64_t n = 3 * ((int64_t)i * wy * wz + wy * k + j);
        for (int c = 0; c < 3; c++) diag[n + c] = 1.0 / diag[n + c];
      }
}

/* ------------------------------------------------------------------------ */
/* Assembled coarse levels: allocateSubspaceMatrix stencil_assembly.c:8-117,
 * assembleSubspaceMatrix :126-260, applyStateOperatorSubspaceMatrix
 * :262-364.  Unpadded node order i*ny*nz + k*ny + j, 3 DOFs per node, the
 * columns of a row are the 3x3x3 neighbour nodes in (i, k, j) order.        */
or_csr *or_csr_create(const or_grid *g, int l) {
  const int ncell = 1 << l;
  const int nxc = g->nelx / ncell + 1, nyc = g->nely / ncell + 1,
            nzc = g->nelz / ncell + 1;
  or_csr *M = (or_csr *)calloc(1, sizeof(or_csr));
  M->nrows = 3 * nxc * nyc * nzc;
  M->rowOffsets = (int *)malloc(sizeof(int) * (M->nrows + 1));
  M->rowOffsets[0] = 0;
  for (int i = 0; i < nxc; i++)
    for (int k = 0; k < nzc; k++)
      for (int j = 0; j < nyc; j++) {
        const int nidx = i * nyc * nzc + nyc * k + j;
        const int sz = 3 * (OR_MIN(i + 1, nxc - 1) - OR_MAX(i - 1, 0) + 1) *
                       (OR_MIN(j + 1, nyc - 1) - OR_MAX(j - 1, 0) + 1) *
                       (OR_MIN(k + 1, nzc - 1) - OR_MAX(k - 1, 0) + 1);
        for (int c = 0; c < 3; c++) M->rowOffsets[3 * nidx + c + 1] = sz;
      }
  for (int i = 1; i < M->nrows + 1; i++) M->rowOffsets[i] += M->rowOffsets[i - 1];
  M->nnz = M->rowOffsets[M->nrows];
  M->colIndex = (int *)malloc(sizeof(int) * M->nnz);
  M->vals = (double *)calloc(M->nnz, sizeof(double));
  for (int i = 0; i < nxc; i++)
    for (int k = 0; k < nzc; k++)
      for (int j = 0; j < nyc; j++) {
        const int rn = i * nyc * nzc + nyc * k + j;
        int off[3] = {M->rowOffsets[3 * rn], M->rowOffsets[3 * rn + 1],
                      M->rowOffsets[3 * rn + 2]};
        for (int ii = OR_MAX(i - 1, 0); ii < OR_MIN(i + 2, nxc); ii++)
          for (int kk = OR_MAX(k - 1, 0); kk < OR_MIN(k + 2, nzc); kk++)
            for (int jj = OR_MAX(j - 1, 0); jj < OR_MIN(j + 2, nyc); jj++) {
              const int cn = ii * nyc * nzc + nyc * kk + jj;
              for (int r = 0; r < 3; r++) {
                for (int c = 0; c < 3; c++) M->colIndex[off[r] + c] = 3 * cn + c;
                off[r] += 3;
              }
            }
      }
  return M;
}

void or_csr_destroy(or_csr *M) {
  if (!M) return;
  free(M->rowOffsets);
  free(M->colIndex);
  free(M->vals);
  free(M);
}

/* haloToTrue, stencil_utility.h:58-76 */
static int64_t halo_to_true(int64_t index, int wy, int wz, int ny, int nz) {
  const int64_t node = index / 3, comp = index % 3;
  const int64_t ih = node / ((int64_t)wy * wz), jh = node % wy,
                kh = (node % ((int64_t)wy * wz)) / wy;
  return 3 * ((ih - 1) * ny * nz + (kh - 1) * ny + (jh - 1)) + comp;
}

void or_csr_assemble(const or_grid *g, int l, const or_dtype *x, or_csr *M,
                     double *tmp) {
  int ncell, wx, wy, wz, fwx, fwy, fwz;
  int64_t ndof, fndof;
  or_padding(g, l, &ncell, &wx, &wy, &wz, &ndof);
  or_padding(g, 0, &ncell, &fwx, &fwy, &fwz, &fndof);
  ncell = 1 << l;
  const int nelxc = g->nelx / ncell, nelyc = g->nely / ncell,
            nelzc = g->nelz / ncell;
  const int nyc = nelyc + 1, nzc = nelzc + 1;
  const int64_t ewy = fwy - 1, ewz = fwz - 1;
  for (int64_t i = 0; i < M->nnz; i++) M->vals[i] = 0.0;
  /* 8-colour element loop, stencil_assembly.c:150-155 */
  for (int bx = 0; bx < 2; bx++)
    for (int bz = 0; bz < 2; bz++)
      for (int by = 0; by < 2; by++)
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = bx; i < nelxc; i += 2)
    for (int k = bz; k < nelzc; k += 2)
      for (int j = by; j < nelyc; j += 2) {
        double ke[24][24];
        int64_t edof[24];
        for (int n = 0; n < 8; n++)
          for (int c = 0; c < 3; c++)
            edof[3 * n + c] = 3 * ((int64_t)(i + lnode[n][0]) * nyc * nzc +
                                   (k + lnode[n][2]) * nyc + (j + lnode[n][1])) + c;
        memset(ke, 0, sizeof(ke));
        for (int ii = 0; ii < ncell; ii++)
          for (int kk = 0; kk < ncell; kk++)
            for (int jj = 0; jj < ncell; jj++) {
              const int64_t ifn = (int64_t)i * ncell + ii + 1;
              const int64_t jfn = (int64_t)j * ncell + jj + 1;
              const int64_t kfn = (int64_t)k * ncell + kk + 1;
              const double sc = young(g, x[ifn * ewy * ewz + kfn * ewy + jfn]);
              const double *Kc = g->KE[l] + 576 * (ncell * ncell * ii + ncell * kk + jj);
              for (int a = 0; a < 24; a++)
                for (int b = 0; b < 24; b++) ke[a][b] += sc * Kc[24 * a + b];
            }
        for (int a = 0; a < 24; a++) {
          const int rs = M->rowOffsets[edof[a]], re = M->rowOffsets[edof[a] + 1];
          for (int b = 0; b < 24; b++)
            for (int idx = rs; idx < re; idx++)
              if (M->colIndex[idx] == edof[b]) {
                M->vals[idx] += ke[a][b];
                break;
              }
        }
      }
  /* symmetric elimination of fixed DOFs, stencil_assembly.c:226-259 */
  for (int r = 0; r < M->nrows; r++) tmp[r] = 1.0;
  for (int f = 0; f < g->nfixed[l]; f++)
    tmp[halo_to_true(g->fixed[l][f], wy, wz, nyc, nzc)] = 0.0;
  for (int r = 0; r < M->nrows; r++)
    for (int c = M->rowOffsets[r]; c < M->rowOffsets[r + 1]; c++)
      M->vals[c] *= tmp[r] * tmp[M->colIndex[c]];
  for (int f = 0; f < g->nfixed[l]; f++) {
    const int64_t r = halo_to_true(g->fixed[l][f], wy, wz, nyc, nzc);
    for (int c = M->rowOffsets[r]; c < M->rowOffsets[r + 1]; c++)
      if (r == M->colIndex[c]) M->vals[c] = 1.0;
  }
}

void or_csr_apply(const or_grid *g, int l, const or_csr *M,
                  const or_ctype *in, or_ctype *out) {
  int ncell, wx, wy, wz;
  int64_t ndof;
  or_padding(g, l, &ncell, &wx, &wy, &wz, &ndof);
  const int nxc = g->nelx / ncell + 1, nyc = g->nely / ncell + 1,
            nzc = g->nelz / ncell + 1;
  for (int64_t i = 0; i < ndof; i++) out[i] = 0;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 1; i < nxc + 1; i++)
    for (int k = 1; k < nzc + 1; k++)
      for (int j = 1; j < nyc + 1; j++) {
        const int64_t hn = (int64_t)i * wy * wz + wy * k + j;
        const int rn = (i - 1) * nyc * nzc + nyc * (k - 1) + (j - 1);
        double acc[3] = {0, 0, 0};
        int off[3] = {M->rowOffsets[3 * rn], M->rowOffsets[3 * rn + 1],
                      M->rowOffsets[3 * rn + 2]};
        for (int ii = OR_MAX(i - 1, 1); ii < OR_MIN(i + 2, nxc + 1); ii++)
          for (int kk = OR_MAX(k - 1, 1); kk < OR_MIN(k + 2, nzc + 1); kk++)
            for (int jj = OR_MAX(j - 1, 1); jj < OR_MIN(j + 2, nyc + 1); jj++) {
              const int64_t cn = 3 * ((int64_t)ii * wy * wz + wy * kk + jj);
              for (int r = 0; r < 3; r++) {
                acc[r] += (double)in[cn] * M->vals[off[r]] +
                          (double)in[cn + 1] * M->vals[off[r] + 1] +
                          (double)in[cn + 2] * M->vals[off[r] + 2];
                off[r] += 3;
              }
            }
        for (int r = 0; r < 3; r++) out[3 * hn + r] = (or_ctype)acc[r];
      }
}

/* ------------------------------------------------------------------------ */
/* Damped Jacobi smoothers: smoothDampedJacobi_halo stencil_solvers.c:97-122,
 * smoothDampedJacobiSubspace_halo :11-49, ...SubspaceMatrix_halo :55-90.    */
static void jacobi_update(const or_grid *g, int l, or_ctype omega,
                          const double *invD, or_ctype *u, const or_ctype *b,
                          const or_ctype *tmp) {
  int ncell, wx, wy, wz;
  int64_t ndof;
  or_padding(g, l, &ncell, &wx, &wy, &wz, &ndof);
  const int nelxc = g->nelx / ncell, nelyc = g->nely / ncell,
            nelzc = g->nelz / ncell;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 1; i < nelxc + 2; i++)
    for (int k = 1; k < nelzc + 2; k++)
      for (int j = 1; j < nelyc + 2; j++) {
        const int64_t n = 3 * ((int64_t)i * wy * wz + wy * k + j);
        for (int c = 0; c < 3; c++)
          u[n + c] += omega * invD[n + c] * (b[n + c] - tmp[n + c]);
      }
}

void or_jacobi0(const or_grid *g, const or_dtype *x, int nswp, or_ctype omega,
                const double *invD, or_ctype *u, const or_ctype *b,
                or_ctype *tmp) {
  for (int s = 0; s < nswp; s++) {
    or_matvec0(g, x, u, tmp);
    jacobi_update(g, 0, omega, invD, u, b, tmp);
  }
}

void or_jacobi_l(const or_grid *g, const or_dtype *x, int l, int nswp,
                 or_ctype omega, const double *invD, or_ctype *u,
                 const or_ctype *b, or_ctype *tmp) {
  for (int s = 0; s < nswp; s++) {
    or_matvec_l(g, l, x, u, tmp);
    jacobi_update(g, l, omega, invD, u, b, tmp);
  }
}

static void jacobi_csr(const or_grid *g, int l, const or_csr *M, int nswp,
                       or_ctype omega, const double *invD, or_ctype *u,
                       const or_ctype *b, or_ctype *tmp) {
  for (int s = 0; s < nswp; s++) {
    or_csr_apply(g, l, M, u, tmp);
    jacobi_update(g, l, omega, invD, u, b, tmp);
  }
}

/* ------------------------------------------------------------------------ */
/* Coarsest-level direct solve.  The reference factorizes the CSR matrix
 * with CHOLMOD (stype=1: the entries with column <= row of the CSR rows are
 * used; stencil_assembly.c:366-503).  The oracle forms the dense symmetric
 * matrix from the same entries and uses a dense LL^T factor.               */
static int coarse_factor(or_solver *s, const or_csr *M) {
  const int n = M->nrows;
  s->ncoarse = n;
  free(s->chol);
  s->chol = (double *)calloc((size_t)n * n, sizeof(double));
  double *A = s->chol;
  int bw = 0;
  for (int r = 0; r < n; r++)
    for (int c = M->rowOffsets[r]; c < M->rowOffsets[r + 1]; c++) {
      const int col = M->colIndex[c];
      if (col <= r) {
        A[(size_t)r * n + col] = M->vals[c];
        if (r - col > bw) bw = r - col;
      }
    }
  s->bw = bw;
  /* the factor keeps the band of A (no fill outside |i-j| <= bw) */
  for (int j = 0; j < n; j++) {
    const int k0 = j - bw > 0 ? j - bw : 0;
    double d = A[(size_t)j * n + j];
    for (int k = k0; k < j; k++) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
    if (!(d > 0.0)) return -1;
    d = sqrt(d);
    A[(size_t)j * n + j] = d;
    const int i1 = j + bw < n - 1 ? j + bw : n - 1;
#pragma omp parallel for schedule(static) if (i1 - j > 64)
    for (int i = j + 1; i <= i1; i++) {
      double v = A[(size_t)i * n + j];
      const double *ri = A + (size_t)i * n, *rj = A + (size_t)j * n;
      for (int k = i - bw > 0 ? i - bw : 0; k < j; k++) v -= ri[k] * rj[k];
      A[(size_t)i * n + j] = v / d;
    }
  }
  return 0;
}

void or_coarse_solve(const or_grid *g, const or_solver *s, int l,
                     const or_ctype *in, or_ctype *out) {
  int ncell, wx, wy, wz;
  int64_t ndof;
  or_padding(g, l, &ncell, &wx, &wy, &wz, &ndof);
  const int nxc = g->nelx / ncell + 1, nyc = g->nely / ncell + 1,
            nzc = g->nelz / ncell + 1;
  const int n = s->ncoarse;
  double *y = (double *)malloc(sizeof(double) * n);
  for (int i = 1; i < nxc + 1; i++)
    for (int k = 1; k < nzc + 1; k++)
      for (int j = 1; j < nyc + 1; j++) {
        const int t = (i - 1) * nyc * nzc + (k - 1) * nyc + (j - 1);
        const int64_t h = (int64_t)i * wy * wz + wy * k + j;
        for (int c = 0; c < 3; c++) y[3 * t + c] = in[3 * h + c];
      }
  const double *L = s->chol;
  const int bw = s->bw;
  for (int i = 0; i < n; i++) {
    double v = y[i];
    for (int k = i - bw > 0 ? i - bw : 0; k < i; k++) v -= L[(size_t)i * n + k] * y[k];
    y[i] = v / L[(size_t)i * n + i];
  }
  for (int i = n - 1; i >= 0; i--) {
    double v = y[i];
    const int k1 = i + bw < n - 1 ? i + bw : n - 1;
    for (int k = i + 1; k <= k1; k++) v -= L[(size_t)k * n + i] * y[k];
    y[i] = v / L[(size_t)i * n + i];
  }
  for (int i = 1; i < nxc + 1; i++)
    for (int k = 1; k < nzc + 1; k++)
      for (int j = 1; j < nyc + 1; j++) {
        const int t = (i - 1) * nyc * nzc + (k - 1) * nyc + (j - 1);
        const int64_t h = (int64_t)i * wy * wz + wy * k + j;
        for (int c = 0; c < 3; c++) out[3 * h + c] = (or_ctype)y[3 * t + c];
      }
  free(y);
}

/* ------------------------------------------------------------------------ */
/* Solver buffers: allocateSolverData, stencil_solvers.c:399-441.  rmg[0]
 * and zmg[0] alias r and z.                                                 */
or_solver *or_solver_create(const or_grid *g) {
  or_solver *s = (or_solver *)calloc(1, sizeof(or_solver));
  for (int l = 0; l < g->nl; l++) {
    int nc, wx, wy, wz;
    int64_t nd;
    or_padding(g, l, &nc, &wx, &wy, &wz, &nd);
    s->invD[l] = (double *)calloc(nd, sizeof(double));
    s->dmg[l] = (or_ctype *)calloc(nd, sizeof(or_ctype));
    if (l == 0) {
      s->r = (or_ctype *)calloc(nd, sizeof(or_ctype));
      s->p = (or_ctype *)calloc(nd, sizeof(or_ctype));
      s->q = (or_ctype *)calloc(nd, sizeof(or_ctype));
      s->z = (or_ctype *)calloc(nd, sizeof(or_ctype));
      s->rmg[0] = s->r;
      s->zmg[0] = s->z;
    } else {
      s->rmg[l] = (or_ctype *)calloc(nd, sizeof(or_ctype));
      s->zmg[l] = (or_ctype *)calloc(nd, sizeof(or_ctype));
    }
    if (l >= 2) {
      or_csr *M = or_csr_create(g, l);
      s->csr[l] = *M;
      free(M);
    }
  }
  return s;
}

void or_solver_destroy(const or_grid *g, or_solver *s) {
  if (!s) return;
  for (int l = 0; l < g->nl; l++) {
    free(s->invD[l]);
    free(s->dmg[l]);
    if (l > 0) {
      free(s->rmg[l]);
      free(s->zmg[l]);
    }
    if (l >= 2) {
      free(s->csr[l].rowOffsets);
      free(s->csr[l].colIndex);
      free(s->csr[l].vals);
    }
  }
  free(s->p);
  free(s->q);
  free(s->r);
  free(s->z);
  free(s->chol);
  free(s);
}

/* Per-solve setup, stencil_solvers.c:283-302: CSR levels, inverse
 * diagonals of every level, coarse factorization.                           */
int or_solver_setup(const or_grid *g, or_solver *s, const or_dtype *x) {
  for (int l = 2; l < g->nl; l++) or_csr_assemble(g, l, x, &s->csr[l], s->invD[l]);
  for (int l = 0; l < g->nl; l++) or_invdiag(g, x, l, s->invD[l]);
  return coarse_factor(s, &s->csr[g->nl - 1]);
}

static int64_t level_ndof(const or_grid *g, int l) {
  int nc, wx, wy, wz;
  int64_t nd;
  or_padding(g, l, &nc, &wx, &wy, &wz, &nd);
  return nd;
}

/* V-cycle preconditioner, stencil_solvers.c:129-252.  Uses s->rmg[l] as
 * input and writes s->zmg[l].                                               */
void or_vcycle(const or_grid *g, or_solver *s, const or_dtype *x, int l,
               or_ctype omega, int nswp) {
  const int nl = g->nl;
  const int64_t nd = level_ndof(g, l), ndn = level_ndof(g, l + 1);
  or_ctype *z = s->zmg[l], *d = s->dmg[l], *r = s->rmg[l];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nd; i++) z[i] = 0;
  if (l == 0) {
    or_jacobi0(g, x, nswp, omega, s->invD[0], z, r, d);
    or_matvec0(g, x, z, d);
  } else if (l < 2) {
    or_jacobi_l(g, x, l, nswp, omega, s->invD[l], z, r, d);
    or_matvec_l(g, l, x, z, d);
  } else {
    jacobi_csr(g, l, &s->csr[l], nswp, omega, s->invD[l], z, r, d);
    or_csr_apply(g, l, &s->csr[l], z, d);
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nd; i++) d[i] = r[i] - d[i];
  or_restrict(g, l, d, s->rmg[l + 1]);
  if (nl == l + 2) {
    for (int64_t i = 0; i < ndn; i++) s->zmg[l + 1][i] = 0;
    or_coarse_solve(g, s, l + 1, s->rmg[l + 1], s->zmg[l + 1]);
  } else {
    or_vcycle(g, s, x, l + 1, omega, nswp);
  }
  or_prolong(g, l, s->zmg[l + 1], d);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nd; i++) z[i] += d[i];
  if (l == 0)
    or_jacobi0(g, x, nswp, omega, s->invD[0], z, r, d);
  else if (l < 2)
    or_jacobi_l(g, x, l, nswp, omega, s->invD[l], z, r, d);
  else
    jacobi_csr(g, l, &s->csr[l], nswp, omega, s->invD[l], z, r, d);
}

/* Reductions, stencil_solvers.h:36-58 (accumulated in CTYPE).  Chunks of a
 * fixed size are summed in parallel and the chunk sums in order, so the
 * result does not depend on the thread count.                               */
#define OR_CHUNK 8192
static or_ctype dot(const or_ctype *a, const or_ctype *b, int64_t n) {
  const int64_t nch = (n + OR_CHUNK - 1) / OR_CHUNK;
  or_ctype *part = (or_ctype *)malloc(sizeof(or_ctype) * (nch > 0 ? nch : 1));
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < nch; c++) {
    or_ctype s = 0;
    const int64_t e = (c + 1) * OR_CHUNK < n ? (c + 1) * OR_CHUNK : n;
    for (int64_t i = c * OR_CHUNK; i < e; i++) s += a[i] * b[i];
    part[c] = s;
  }
  or_ctype s = 0;
  for (int64_t c = 0; c < nch; c++) s += part[c];
  free(part);
  return s;
}
static or_ctype norm2(const or_ctype *v, int64_t n) {
  return (or_ctype)sqrt(dot(v, v, n));
}

#include <omp.h>

/* MGCG, solveStateMG_halo, stencil_solvers.c:259-397.  maxIter is 1000 in
 * the reference (:334); a smaller value bounds a CPU timing sample.  The
 * set-up and CG-loop wall times are left in s->setup_s and s->loop_s.
 * Returns 0, or -1 when the coarse factorization fails.                    */
int or_solve(const or_grid *g, or_solver *s, or_dtype *x, int nswp,
             or_ctype tol, int *finalIter, float *finalRes, or_ctype *b,
             or_stype *u, int maxIter) {
  const double t0 = omp_get_wtime();
  const int64_t ndof = level_ndof(g, 0);
  or_ctype *r = s->r, *p = s->p, *q = s->q, *z = s->z;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < ndof; i++) z[i] = (or_ctype)u[i];
  if (or_solver_setup(g, s, x)) return -1;
  or_ctype rhoold = 0, dpr, alpha, rho;
  or_matvec0(g, x, z, r);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < ndof; i++) r[i] = b[i] - r[i];
  const double omega = 0.6;
  const or_ctype bnorm = norm2(b, ndof);
  const double t1 = omp_get_wtime();
  s->setup_s = t1 - t0;
  for (int iter = 0; iter < maxIter; iter++) {
    or_vcycle(g, s, x, 0, (or_ctype)omega, nswp);
    rho = dot(r, z, ndof);
    if (iter == 0) {
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < ndof; i++) p[i] = z[i];
    } else {
      const or_ctype beta = rho / rhoold;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < ndof; i++) p[i] = beta * p[i] + z[i];
    }
    or_matvec0(g, x, p, q);
    dpr = dot(p, q, ndof);
    alpha = rho / dpr;
    rhoold = rho;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < ndof; i++) {
      u[i] += (or_stype)(alpha * p[i]);
      r[i] -= alpha * q[i];
    }
    const or_ctype rnorm = norm2(r, ndof);
    const or_ctype relres = rnorm / bnorm;
    *finalIter = iter;
    *finalRes = (float)relres;
    if (relres < tol) break;
  }
  s->loop_s = omp_get_wtime() - t1;
  return 0;
}

void or_solver_times(const or_solver *s, double *setup_s, double *loop_s) {
  *setup_s = s->setup_s;
  *loop_s = s->loop_s;
}

/* ------------------------------------------------------------------------ */
/* Compliance and sensitivity: getComplianceAndSensetivity_halo,
 * stencil_methods.c:808-862 (eight passes, one per local row node).        */
void or_sensitivity(const or_grid *g, const or_dtype *x, const or_stype *u,
                    or_dtype *c, or_dtype *dcdx) {
  int ncell, wx, wy, wz;
  int64_t ndof;
  or_padding(g, 0, &ncell, &wx, &wy, &wz, &ndof);
  const int64_t ewy = wy - 1, ewz = wz - 1;
  const double *KE = g->KE[0];
  const double E_diff = g->E0 - g->Emin, Emin = g->Emin;
  const int64_t nelx = g->nelx, nely = g->nely, nelz = g->nelz;
  const int64_t ne = nelx * nely * nelz;
  const int64_t nch = (ne + OR_CHUNK - 1) / OR_CHUNK;
  or_dtype *part = (or_dtype *)malloc(sizeof(or_dtype) * (nch > 0 ? nch : 1));
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < ne; t++) {
    const int64_t i = t / (nely * nelz) + 1, k = (t / nely) % nelz + 1, j = t % nely + 1;
    dcdx[i * ewy * ewz + k * ewy + j] = 0;
  }
  or_dtype cc = 0;
  for (int ii = 0; ii < 8; ii++) {
#pragma omp parallel for schedule(static)
    for (int64_t ch = 0; ch < nch; ch++) {
      or_dtype ps = 0;
      const int64_t t1 = (ch + 1) * OR_CHUNK < ne ? (ch + 1) * OR_CHUNK : ne;
      for (int64_t t = ch * OR_CHUNK; t < t1; t++) {
        const int64_t i = t / (nely * nelz) + 1, k = (t / nely) % nelz + 1, j = t % nely + 1;
        int64_t edof[8];
        for (int n = 0; n < 8; n++)
          edof[n] = 3 * ((i + lnode[n][0]) * wy * wz + (k + lnode[n][2]) * wy + (j + lnode[n][1]));
        const int64_t e = i * ewy * ewz + k * ewy + j;
        or_stype cl = 0.0;
        for (int a = 0; a < 3; a++) {
          double tt = 0.0;
          for (int n = 0; n < 8; n++) {
            tt += KE[(ii * 3 + a) * 24 + 3 * n] * u[edof[n]];
            tt += KE[(ii * 3 + a) * 24 + 3 * n + 1] * u[edof[n] + 1];
            tt += KE[(ii * 3 + a) * 24 + 3 * n + 2] * u[edof[n] + 2];
          }
          cl += u[edof[ii] + a] * tt;
        }
        const or_dtype xe = x[e];
        ps += cl * (Emin + xe * xe * xe * E_diff);
        dcdx[e] += cl * (-3.0 * E_diff * xe * xe);
      }
      part[ch] = ps;
    }
    for (int64_t ch = 0; ch < nch; ch++) cc += part[ch];
  }
  free(part);
  *c = cc;
}

/* Density filter xPhys = (H x)/Hs: applyDensityFilter,
 * stencil_optimization.c:228-285.  Neighbourhood i2 in
 * [i1-(ceil(r)-1), i1+ceil(r)] clipped to the grid, weight max(0, r-dist). */
void or_filter(const or_grid *g, or_dtype rmin, const or_dtype *rho,
               or_dtype *out) {
  int ncell, wx, wy, wz;
  int64_t ndof;
  or_padding(g, 0, &ncell, &wx, &wy, &wz, &ndof);
  const uint32_t ewy = wy - 1, ewz = wz - 1;
  const uint32_t nelx = g->nelx, nely = g->nely, nelz = g->nelz;
  const double cr = ceil(rmin);
#pragma omp parallel for collapse(2) schedule(static)
  for (uint32_t i1 = 1; i1 < nelx + 1; i1++)
    for (uint32_t k1 = 1; k1 < nelz + 1; k1++)
      for (uint32_t j1 = 1; j1 < nely + 1; j1++) {
        const uint32_t e1 = i1 * ewy * ewz + k1 * ewy + j1;
        double o = 0.0, w = 0.0;
        const uint32_t i2max = (uint32_t)OR_MIN(i1 + (cr + 1), nelx + 1);
        const uint32_t i2min = (uint32_t)OR_MAX(i1 - (cr - 1), 1);
        const uint32_t k2max = (uint32_t)OR_MIN(k1 + (cr + 1), nelz + 1);
        const uint32_t k2min = (uint32_t)OR_MAX(k1 - (cr - 1), 1);
        const uint32_t j2max = (uint32_t)OR_MIN(j1 + (cr + 1), nely + 1);
        const uint32_t j2min = (uint32_t)OR_MAX(j1 - (cr - 1), 1);
        for (uint32_t i2 = i2min; i2 < i2max; i2++)
          for (uint32_t k2 = k2min; k2 < k2max; k2++)
            for (uint32_t j2 = j2min; j2 < j2max; j2++) {
              const uint32_t e2 = i2 * ewy * ewz + k2 * ewy + j2;
              const uint32_t d2 = (i1 - i2) * (i1 - i2) + (j1 - j2) * (j1 - j2) +
                                  (k1 - k2) * (k1 - k2);
              const double fw = OR_MAX(0.0, rmin - sqrt((double)d2));
              o += fw * rho[e2];
              w += fw;
            }
        out[e1] = (or_dtype)(o / w);
      }
}

/* Transposed filter v <- H (v / Hs): applyDensityFilterGradient,
 * stencil_optimization.c:290-395.                                           */
void or_filter_grad(const or_grid *g, or_dtype rmin, or_dtype *v) {
  int ncell, wx, wy, wz;
  int64_t ndof;
  or_padding(g, 0, &ncell, &wx, &wy, &wz, &ndof);
  const uint32_t ewy = wy - 1, ewz = wz - 1;
  const uint32_t nelx = g->nelx, nely = g->nely, nelz = g->nelz;
  const int64_t N = (int64_t)(wx - 1) * ewy * ewz;
  or_dtype *tmp = (or_dtype *)calloc(N, sizeof(or_dtype));
  const double cr = ceil(rmin);
  for (int pass = 0; pass < 2; pass++) {
#pragma omp parallel for collapse(2) schedule(static)
    for (uint32_t i1 = 1; i1 < nelx + 1; i1++)
      for (uint32_t k1 = 1; k1 < nelz + 1; k1++)
        for (uint32_t j1 = 1; j1 < nely + 1; j1++) {
          const uint32_t e1 = i1 * ewy * ewz + k1 * ewy + j1;
          double acc = 0.0;
          const uint32_t i2max = (uint32_t)OR_MIN(i1 + (cr + 1), nelx + 1);
          const uint32_t i2min = (uint32_t)OR_MAX(i1 - (cr - 1), 1);
          const uint32_t k2max = (uint32_t)OR_MIN(k1 + (cr + 1), nelz + 1);
          const uint32_t k2min = (uint32_t)OR_MAX(k1 - (cr - 1), 1);
          const uint32_t j2max = (uint32_t)OR_MIN(j1 + (cr + 1), nely + 1);
          const uint32_t j2min = (uint32_t)OR_MAX(j1 - (cr - 1), 1);
          for (uint32_t i2 = i2min; i2 < i2max; i2++)
            for (uint32_t k2 = k2min; k2 < k2max; k2++)
              for (uint32_t j2 = j2min; j2 < j2max; j2++) {
                const uint32_t d2 = (i1 - i2) * (i1 - i2) +
                                    (j1 - j2) * (j1 - j2) + (k1 - k2) * (k1 - k2);
                const double fw = OR_MAX(0.0, ((double)rmin) - sqrt((double)d2));
                if (pass == 0)
                  acc += fw;
                else
                  acc += fw * tmp[i2 * ewy * ewz + k2 * ewy + j2];
              }
          if (pass == 0)
            tmp[e1] = (or_dtype)(((double)v[e1]) / acc);
          else
            v[e1] = (or_dtype)acc;
        }
  }
  free(tmp);
}

/* compute_volume, gpu_definitions.c:94-110 */
or_dtype or_volume(const or_grid *g, const or_dtype *xPhys) {
  int ncell, wx, wy, wz;
  int64_t ndof;
  or_padding(g, 0, &ncell, &wx, &wy, &wz, &ndof);
  or_dtype vol = 0;
  for (int i = 1; i < g->nelx + 1; i++)
    for (int k = 1; k < g->nelz + 1; k++)
      for (int j = 1; j < g->nely + 1; j++)
        vol += xPhys[(int64_t)i * (wy - 1) * (wz - 1) + k * (wy - 1) + j];
  return vol;
}

/* compute_change, gpu_definitions.c:112-125 (float result in both
 * typings, as in the reference), also copies xnew into x.                   */
float or_change(or_dtype *x, const or_dtype *xnew, int64_t nelem) {
  float change = 0.0f;
  for (int64_t i = 0; i < nelem; i++) {
    change = OR_MAX(change, fabs(x[i] - xnew[i]));
    x[i] = xnew[i];
  }
  return change;
}

/* Optimality-criteria update by bisection on lambda:
 * update_solution, gpu_definitions.c:127-157.  The acceptance test mixes the
 * mean volume error g with the sum of dv*(xnew-x) over the padded array, as
 * the reference does.
 *
 * OR_OC_MOVE, OR_OC_ETA, OR_OC_FLIP and OR_OC_DROP_G exist only so that the
 * tests can build deliberately wrong copies of this function (a wrong move
 * limit, exponent, bisection direction, or a dropped volume term) and show
 * that the OC checks reject each of them; the shipped builds never set them. */
#ifndef OR_OC_MOVE
#define OR_OC_MOVE 0.2
#endif
#ifdef OR_OC_ETA
#define OR_OC_ROOT(v) pow((v), OR_OC_ETA)
#else
#define OR_OC_ROOT(v) sqrt(v)
#endif
void or_oc(const or_dtype *x, or_dtype *xnew, const or_dtype *dv,
           const or_dtype *dc, int64_t nelem, or_dtype g) {
  or_dtype l1 = 0.0, l2 = 1e9, move = OR_OC_MOVE;
  const int64_t nch = (nelem + OR_CHUNK - 1) / OR_CHUNK;
  or_dtype *part = (or_dtype *)malloc(sizeof(or_dtype) * (nch > 0 ? nch : 1));
  while ((l2 - l1) / (l1 + l2) > 1e-6) {
    const or_dtype lmid = 0.5 * (l2 + l1);
    or_dtype gt = 0.0;
#pragma omp parallel for schedule(static)
    for (int64_t ch = 0; ch < nch; ch++) {
      or_dtype ps = 0;
      const int64_t i1 = (ch + 1) * OR_CHUNK < nelem ? (ch + 1) * OR_CHUNK : nelem;
      for (int64_t i = ch * OR_CHUNK; i < i1; i++) {
        xnew[i] = OR_MAX(0.0, OR_MAX(x[i] - move,
                                     OR_MIN(1.0, OR_MIN(x[i] + move,
                                                        x[i] * OR_OC_ROOT(-dc[i] / (dv[i] * lmid))))));
        ps += dv[i] * (xnew[i] - x[i]);
      }
      part[ch] = ps;
    }
    for (int64_t ch = 0; ch < nch; ch++) gt += part[ch];
#ifndef OR_OC_DROP_G
    gt += g;
#endif
#ifdef OR_OC_FLIP
    if (gt <= 0)
#else
    if (gt > 0)
#endif
      l1 = lmid;
    else
      l2 = lmid;
  }
  free(part);
}

/* ------------------------------------------------------------------------ */
/* Design loop: top3dmgcg, stencil_optimization.c:13-223 (without the VTU
 * output).  records[] receives one entry per design iteration; returns the
 * number of iterations run, or -1 on failure.                               */
int or_top3d(int nelx, int nely, int nelz, float volfrac, float rmin, int nl,
             float cgtol, int iters, int yalign, or_record *records,
             or_dtype *xPhys_out, or_stype *U_out) {
  or_grid *g = or_create(nelx, nely, nelz, nl, yalign);
  if (!g) return -1;
  int ncell, wx, wy, wz;
  int64_t ndof;
  or_padding(g, 0, &ncell, &wx, &wy, &wz, &ndof);
  const int64_t nelem = (int64_t)(wx - 1) * (wy - 1) * (wz - 1);
  or_ctype *F = (or_ctype *)malloc(sizeof(or_ctype) * ndof);
  or_stype *U = (or_stype *)calloc(ndof, sizeof(or_stype));
  or_load(g, F);
  or_dtype *dc = (or_dtype *)calloc(nelem, sizeof(or_dtype));
  or_dtype *dv = (or_dtype *)malloc(sizeof(or_dtype) * nelem);
  or_dtype *x = (or_dtype *)calloc(nelem, sizeof(or_dtype));
  or_dtype *xPhys = (or_dtype *)calloc(nelem, sizeof(or_dtype));
  or_dtype *xnew = (or_dtype *)calloc(nelem, sizeof(or_dtype));
  for (int64_t i = 0; i < nelem; i++) dv[i] = 1.0;
  for (int i = 1; i < nelx + 1; i++)
    for (int k = 1; k < nelz + 1; k++)
      for (int j = 1; j < nely + 1; j++) {
        const int64_t idx = (int64_t)i * (wy - 1) * (wz - 1) + k * (wy - 1) + j;
        x[idx] = volfrac;
        xPhys[idx] = volfrac;
      }
  or_solver *s = or_solver_create(g);
  or_filter_grad(g, rmin, dv);
  int loop = 0;
  float change = 1;
  int rc = 0;
  while ((change > 1e-2) && (loop < iters)) {
    loop++;
    int cgiter = 0;
    float cgres = 0;
    if (or_solve(g, s, xPhys, 4, (or_ctype)cgtol, &cgiter, &cgres, F, U, 1000)) {
      rc = -1;
      break;
    }
    or_dtype vol = or_volume(g, xPhys);
    or_dtype c = 0;
    or_sensitivity(g, xPhys, U, &c, dc);
    or_filter_grad(g, rmin, dc);
    vol /= (or_dtype)(nelx * nely * nelz);
    const or_dtype gg = vol - volfrac;
    or_oc(x, xnew, dv, dc, nelem, gg);
    change = or_change(x, xnew, nelem);
    or_filter(g, rmin, x, xPhys);
    if (records) {
      records[loop - 1].loop = loop;
      records[loop - 1].c = c;
      records[loop - 1].vol = vol;
      records[loop - 1].change = change;
      records[loop - 1].relres = cgres;
      records[loop - 1].cgiter = cgiter;
    }
  }
  if (xPhys_out) memcpy(xPhys_out, xPhys, sizeof(or_dtype) * nelem);
  if (U_out) memcpy(U_out, U, sizeof(or_stype) * ndof);
  or_solver_destroy(g, s);
  or_destroy(g);
  free(F);
  free(U);
  free(dc);
  free(dv);
  free(x);
  free(xPhys);
  free(xnew);
  return rc ? rc : loop;
}

/* ------------------------------------------------------------------------ */
/* Density snapshot in the ASCII VTK unstructured-grid format:
 * writeDensityVtkFile (halo = 0) and writeDensityVtkFileWithHalo (halo = 1),
 * stencil_optimization.c:399-598.  Points sit at integer node coordinates
 * (x=i, y=j, z=k), x slowest and y fastest.  Each element is a
 * VTK_HEXAHEDRON (type 12) listing its 8 nodes in the element's local node
 * order (getEdof_halo, stencil_utility.h:8-26).  One Float32 "density" value
 * per cell comes from the padded element layout.  halo = 0 writes the
 * (nel+1)^3 node box from 0 and the real elements 1..nel; halo = 1 the whole
 * padded grid.  Returns 0, or -1 if the file cannot be opened.             */
int or_write_vtu(const or_grid *g, const or_dtype *x, const char *path,
                 int halo) {
  /* local node q of element (i, j, k) sits at (i, j, k) + corner[q] */
  static const int corner[8][3] = {{0, 1, 0}, {1, 1, 0}, {1, 0, 0},
                                   {0, 0, 0}, {0, 1, 1}, {1, 1, 1},
                                   {1, 0, 1}, {0, 0, 1}}; /* (dx, dy, dz) */
  int ncell, wx, wy, wz;
  int64_t ndof;
  or_padding(g, 0, &ncell, &wx, &wy, &wz, &ndof);
  const int64_t ewy = wy - 1, ewz = wz - 1;
  const int px = halo ? wx : g->nelx + 1;
  const int py = halo ? wy : g->nely + 1;
  const int pz = halo ? wz : g->nelz + 1;
  const int cx = px - 1, cy = py - 1, cz = pz - 1;
  const int e0 = halo ? 0 : 1; /* element index of the first cell */
  FILE *f = fopen(path, "w");
  if (!f) return -1;
  fprintf(f, "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" "
             "byte_order=\"LittleEndian\">\n<UnstructuredGrid>\n");
  fprintf(f, "<Piece NumberOfPoints=\"%i\" NumberOfCells=\"%i\">\n",
          px * py * pz, cx * cy * cz);
  fprintf(f, "<Points>\n<DataArray type=\"Float32\" NumberOfComponents=\"3\" "
             "format=\"ascii\">\n");
  for (int i = 0; i < px; i++)
    for (int k = 0; k < pz; k++)
      for (int j = 0; j < py; j++)
        fprintf(f, "%e %e %e\n", (double)(float)i, (double)(float)j,
                (double)(float)k);
  fprintf(f, "</DataArray>\n</Points>\n<Cells>\n");
  fprintf(f, "<DataArray type=\"Int32\" Name=\"connectivity\" "
             "format= \"ascii\">\n");
  for (int i = 0; i < cx; i++)
    for (int k = 0; k < cz; k++)
      for (int j = 0; j < cy; j++)
        for (int q = 0; q < 8; q++) {
          const int n = (i + corner[q][0]) * py * pz + (k + corner[q][2]) * py +
                        (j + corner[q][1]);
          fprintf(f, q < 7 ? "%d " : "%d\n", n);
        }
  fprintf(f, "</DataArray>\n");
  fprintf(f, "<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n");
  for (int c = 1; c <= cx * cy * cz; c++) fprintf(f, "%d\n", 8 * c);
  fprintf(f, "</DataArray>\n");
  fprintf(f, "<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
  for (int c = 0; c < cx * cy * cz; c++) fprintf(f, "12\n");
  fprintf(f, "</DataArray>\n</Cells>\n<CellData>\n");
  fprintf(f, "<DataArray type=\"Float32\" NumberOfComponents=\"1\" "
             "Name=\"density\" format=\"ascii\">\n");
  for (int i = 0; i < cx; i++)
    for (int k = 0; k < cz; k++)
      for (int j = 0; j < cy; j++)
        fprintf(f, "%e\n",
                (double)x[(i + e0) * ewy * ewz + (k + e0) * ewy + (j + e0)]);
  fprintf(f, "</DataArray>\n</CellData>\n</Piece>\n</UnstructuredGrid>\n"
             "</VTKFile>\n");
  fclose(f);
  return 0;
}
