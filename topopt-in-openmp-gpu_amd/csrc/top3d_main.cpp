// top3d command line (top3d.c:14-88): -x -y -z coarse elements, -l levels,
// -f volume fraction, -r filter radius, -i design iterations, -v verbose,
// -w write VTU; plus -p 0|1 (FP64 | mixed precision), -b 0|1 (reference
// | symmetric BC rows on the matrix-free levels) and -s 0|1 (two clamped
// strips of the x=0 face | the whole face clamped).
#include <getopt.h>

#include <cstdio>
#include <cstdlib>

#include "../../include/top3d_hip.h"

int main(int argc, char *argv[]) {
  int nelx_coarse = 12, nely_coarse = 6, nelz_coarse = 6;
  float volfrac = 0.2f, rmin = 1.5f;
  int iters = 20, nl = 4, verbose = 0, write_result = 0, precision = 0, bc = 0, support = 0;
  int opt;
  while ((opt = getopt(argc, argv, "x:y:z:r:f:i:l:v:w:p:b:s:")) != -1) {
    switch (opt) {
      case 'x': nelx_coarse = std::atoi(optarg); break;
      case 'y': nely_coarse = std::atoi(optarg); break;
      case 'z': nelz_coarse = std::atoi(optarg); break;
      case 'r': rmin = (float)std::atof(optarg); break;
      case 'f': volfrac = (float)std::atof(optarg); break;
      case 'i': iters = std::atoi(optarg); break;
      case 'l': nl = std::atoi(optarg); break;
      case 'v': verbose = std::atoi(optarg); break;
      case 'w': write_result = std::atoi(optarg); break;
      case 'p': precision = std::atoi(optarg); break;
      case 'b': bc = std::atoi(optarg); break;
      case 's': support = std::atoi(optarg); break;
      default:
        std::fprintf(stderr, "Usage: %s [-xyzrfilvwpbs]\n", argv[0]);
        return EXIT_FAILURE;
    }
  }
  int sizeIncr = 2;  // fine = coarse * 2^(nl-1), top3d.c:62-68
  for (int i = 2; i < nl; i++) sizeIncr *= 2;
  const int nelx = nelx_coarse * sizeIncr, nely = nely_coarse * sizeIncr,
            nelz = nelz_coarse * sizeIncr;
  if (verbose)
    std::printf("Running topopt with:\n number of coarse els x (-x): %i (fine els = %i)\n"
                " number of coarse els y (-y): %i (fine els = %i)\n number of coarse els z "
                "(-z): %i (fine els = %i)\n total number of elements: %i\n volume fraction "
                "(-f): %f\n filter radius in elements (-r): %f\n number of design iterations "
                "(-i): %i\n verbose (-v): %i\n write file (-w): %i\n precision (-p): %s\n\n",
                nelx_coarse, nelx, nely_coarse, nely, nelz_coarse, nelz, nelx * nely * nelz,
                volfrac, rmin, iters, verbose, write_result, precision ? "mixed" : "fp64");
  top3d_set_default_precision(precision ? TOP3D_MIXED : TOP3D_FP64);
  top3d_set_default_bc_mode(bc ? TOP3D_BC_SYMMETRIC : TOP3D_BC_REFERENCE);
  top3d_set_default_support(support ? TOP3D_SUPPORT_FACE : TOP3D_SUPPORT_STRIPS);
  const float cgtol = 1e-5f;
  const int cgmax = 200;
  top3dmgcg(nelx, nely, nelz, volfrac, rmin, nl, cgtol, cgmax, verbose, write_result, iters);
  if (top3d_last_error() != TOP3D_OK) {
    std::fprintf(stderr, "top3d: %s\n", top3d_last_error_string());
    return 1;
  }
  return 0;
}
