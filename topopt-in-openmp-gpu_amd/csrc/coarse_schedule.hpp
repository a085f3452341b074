// Elimination schedule of the coarse-level tiled Cholesky (host code, no
// device calls): the storage order of the DOFs, the workgroup tasks of every
// launch, and the non-zero tile rows of W = L^{-1}.
#pragma once

#include <string>
#include <vector>

namespace t3d {

constexpr int kTile = 64;   // tile size of the coarse factorization
constexpr int kMaxSrc = 4;  // tile columns updating one tile in one launch
constexpr int kTaskFactor = 1;

// One workgroup task: tile (I, J), I >= J, minus the updates of the tile
// columns K[0..nsrc) eliminated in this launch; kTaskFactor: the (diagonal)
// tile is complete afterwards and is factored in place.  nsrc = 0 with
// kTaskFactor factors the tile as it is.
struct CholTask {
  int I, J, nsrc, flags;
  int K[kMaxSrc];
};

struct CoarseSchedule {
  int n = 0, np = 0, nblk = 0;         // logical size, storage size (64 nblk)
  std::vector<int> pos;                // logical DOF -> storage index
  std::vector<CholTask> tasks;         // all launches, concatenated
  std::vector<int> launch_first, launch_count;
  // W = L^{-1}: tile column J has colinfo[2J+1] rows starting at row
  // colinfo[2J]; row r is (I, klist offset, klist count, 0) at rows[4r..];
  // klist holds the K < I of the column's earlier rows with L_IK != 0
  std::vector<int> colinfo, rows, klist;
  std::string error;
};

// planes > 1 with n divisible by planes: the DOFs form `planes` consecutive
// blocks coupled only to adjacent blocks (x-planes of a grid), eliminated in
// nested-dissection order, each plane padded to whole tiles.  Otherwise the
// matrix is taken as banded (half bandwidth bw) in its own order.
bool build_coarse_schedule(int n, int bw, int planes, CoarseSchedule &out);

}  // namespace t3d
