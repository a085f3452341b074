/* Prints sizeof/offsetof of every struct of include/top3d_hip.h, compiled as
 * C (the language of the reference's callers).  tests/test_abi.py compiles
 * and runs it and compares with the reference's layout on x86-64 glibc
 * (definitions.h:31-53, gpu_definitions.h:7-37, stencil_assembly.h:8-27). */
#include <stddef.h>
#include <stdio.h>

#include "top3d_hip.h"

#define F(S, m) printf("%s.%s %zu\n", #S, #m, offsetof(struct S, m))
#define Z(S) printf("%s %zu\n", #S, sizeof(struct S))

int main(void) {
  Z(FixedDofs); F(FixedDofs, n); F(FixedDofs, idx);
  Z(gridContext);
  F(gridContext, E0); F(gridContext, Emin); F(gridContext, nu);
  F(gridContext, elementSizeX); F(gridContext, elementSizeY); F(gridContext, elementSizeZ);
  F(gridContext, nelx); F(gridContext, nely); F(gridContext, nelz);
  F(gridContext, wrapx); F(gridContext, wrapy); F(gridContext, wrapz);
  F(gridContext, penal); F(gridContext, precomputedKE); F(gridContext, fixedDofs);
  Z(CSRMatrix); F(CSRMatrix, nnz); F(CSRMatrix, nrows); F(CSRMatrix, rowOffsets);
  F(CSRMatrix, colIndex); F(CSRMatrix, vals);
  Z(CoarseSolverData); F(CoarseSolverData, cholmodCommon); F(CoarseSolverData, factoredMatrix);
  F(CoarseSolverData, E_workspace);
  Z(SolverDataBuffer); F(SolverDataBuffer, r); F(SolverDataBuffer, p); F(SolverDataBuffer, q);
  F(SolverDataBuffer, z); F(SolverDataBuffer, invD); F(SolverDataBuffer, dmg);
  F(SolverDataBuffer, rmg); F(SolverDataBuffer, zmg); F(SolverDataBuffer, coarseMatrices);
  F(SolverDataBuffer, bottomSolver);
  Z(gpuNode); F(gpuNode, id); F(gpuNode, prev); F(gpuNode, next); F(gpuNode, gc);
  F(gpuNode, offset_x); F(gpuNode, x_global);
  Z(gpuGrid); F(gpuGrid, num_targets); F(gpuGrid, target);
  printf("uint_fast32_t %zu\n", sizeof(uint_fast32_t));
  return 0;
}
