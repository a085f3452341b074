/*
 * top3d_oracle.h -- CPU oracle for the MGCG state-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * program's algorithm (topopt/TopOpt-in-OpenMP-GPU), written from its
 * behaviour, and used solely as the checker for the HIP library: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load
 * it.  The product library never links or calls it.
 *
 * Two typings are built from the same source:
 *   liboracle_fp64.so    CTYPE = DTYPE = double  (the "FP64 oracle": the
 *                        reference re-typed, definitions.h:25-29)
 *   liboracle_asship.so  CTYPE = DTYPE = float   (the reference as shipped)
 * MTYPE (element matrices, inverse diagonal) and STYPE (displacement) are
 * double in both, as in the reference.
 *
 * Each function cites the reference file:line it restates.  Parity status of
 * every function (how it is pinned independently of this file) is listed in
 * the header comment of top3d_oracle.c.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef ORACLE_ASSHIPPED
typedef float or_ctype;
typedef float or_dtype;
#else
typedef double or_ctype;
typedef double or_dtype;
#endif
typedef double or_mtype;
typedef double or_stype;

#define OR_MAXL 8

typedef struct {
  int nrows;
  int64_t nnz;
  int *rowOffsets;
  int *colIndex;
  double *vals;
} or_csr;

typedef struct {
  int nelx, nely, nelz; /* fine-grid element counts */
  int nl;               /* number of multigrid levels */
  int yalign;           /* y padding alignment (the reference uses 128) */
  double E0, Emin, nu;
  double *KE[OR_MAXL];  /* 576 * 8^l doubles per level */
  int nfixed[OR_MAXL];
  int64_t *fixed[OR_MAXL];
  int bcsym; /* 0: reference BC rows (identity rows, columns kept);
                1: symmetric elimination on the matrix-free levels too */
  int support; /* 0: two clamped strips of the x=0 face (as shipped);
                  1: the whole x=0 face clamped (the commented variant) */
} or_grid;

typedef struct {
  or_ctype *r, *p, *q, *z;
  double *invD[OR_MAXL];
  or_ctype *dmg[OR_MAXL], *rmg[OR_MAXL], *zmg[OR_MAXL];
  or_csr csr[OR_MAXL];
  int ncoarse;     /* unpadded DOFs of the coarsest level */
  double *chol;    /* dense lower Cholesky factor of the coarsest matrix */
  int bw;          /* half bandwidth of the coarsest matrix */
  double setup_s, loop_s; /* wall time of the last or_solve: set-up, CG loop */
} or_solver;

typedef struct {
  int loop;
  double c, vol, change, relres;
  int cgiter;
} or_record;

#ifdef __cplusplus
extern "C" {
#endif

int or_sizeof_ctype(void);
or_grid *or_create(int nelx, int nely, int nelz, int nl, int yalign);
void or_destroy(or_grid *g);
void or_set_bcsym(or_grid *g, int on);
void or_set_support(or_grid *g, int mode);
void or_padding(const or_grid *g, int l, int *ncell, int *wrapx, int *wrapy,
                int *wrapz, int64_t *ndof);
int or_nfixed(const or_grid *g, int l);
void or_fixed(const or_grid *g, int l, int64_t *out);
void or_ke(double nu, int l, int npts, double *KE);
void or_ke_copy(const or_grid *g, int l, double *out);
void or_load(const or_grid *g, or_ctype *F);
void or_line_load(const or_grid *g, or_ctype *F);
void or_line_load_profile(const or_grid *g, or_ctype *F, int profile);

void or_matvec0(const or_grid *g, const or_dtype *x, const or_ctype *in,
                or_ctype *out);
void or_matvec_l(const or_grid *g, int l, const or_dtype *x,
                 const or_ctype *in, or_ctype *out);
void or_prolong(const or_grid *g, int l, const or_ctype *coarse,
                or_ctype *fine);
void or_restrict(const or_grid *g, int l, const or_ctype *fine,
                 or_ctype *coarse);
void or_invdiag(const or_grid *g, const or_dtype *x, int l, double *diag);

or_csr *or_csr_create(const or_grid *g, int l);
void or_csr_destroy(or_csr *M);
void or_csr_assemble(const or_grid *g, int l, const or_dtype *x, or_csr *M,
                     double *tmp);
void or_csr_apply(const or_grid *g, int l, const or_csr *M,
                  const or_ctype *in, or_ctype *out);

void or_jacobi0(const or_grid *g, const or_dtype *x, int nswp, or_ctype omega,
                const double *invD, or_ctype *u, const or_ctype *b,
                or_ctype *tmp);
void or_jacobi_l(const or_grid *g, const or_dtype *x, int l, int nswp,
                 or_ctype omega, const double *invD, or_ctype *u,
                 const or_ctype *b, or_ctype *tmp);

or_solver *or_solver_create(const or_grid *g);
void or_solver_destroy(const or_grid *g, or_solver *s);
int or_solver_setup(const or_grid *g, or_solver *s, const or_dtype *x);
void or_coarse_solve(const or_grid *g, const or_solver *s, int l,
                     const or_ctype *in, or_ctype *out);
void or_vcycle(const or_grid *g, or_solver *s, const or_dtype *x, int l,
               or_ctype omega, int nswp);
int or_solve(const or_grid *g, or_solver *s, or_dtype *x, int nswp,
             or_ctype tol, int *finalIter, float *finalRes, or_ctype *b,
             or_stype *u, int maxIter);
void or_solver_times(const or_solver *s, double *setup_s, double *loop_s);

void or_sensitivity(const or_grid *g, const or_dtype *x, const or_stype *u,
                    or_dtype *c, or_dtype *dcdx);
void or_filter(const or_grid *g, or_dtype rmin, const or_dtype *rho,
               or_dtype *out);
void or_filter_grad(const or_grid *g, or_dtype rmin, or_dtype *v);
or_dtype or_volume(const or_grid *g, const or_dtype *xPhys);
float or_change(or_dtype *x, const or_dtype *xnew, int64_t nelem);
void or_oc(const or_dtype *x, or_dtype *xnew, const or_dtype *dv,
           const or_dtype *dc, int64_t nelem, or_dtype g);

int or_top3d(int nelx, int nely, int nelz, float volfrac, float rmin, int nl,
             float cgtol, int iters, int yalign, or_record *records,
             or_dtype *xPhys_out, or_stype *U_out);
int or_write_vtu(const or_grid *g, const or_dtype *x, const char *path,
                 int halo);

#ifdef __cplusplus
}
#endif
