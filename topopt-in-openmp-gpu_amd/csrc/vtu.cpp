// ASCII VTK unstructured-grid output of the element densities, the format of
// writeDensityVtkFile / writeDensityVtkFileWithHalo
// (stencil_optimization.c:396-598): points at integer node coordinates
// (x=i, y=j, z=k), VTK_HEXAHEDRON (type 12) cells in the local node order of
// the solver, one Float32 "density" value per cell.  `densityArray` is a
// host array in this library's padded element layout.
#include <cstdio>

#include "t3d_internal.hpp"

namespace {

void header(FILE *f, long npts, long ncells) {
  std::fprintf(f, "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" "
                  "byte_order=\"LittleEndian\">\n<UnstructuredGrid>\n");
  std::fprintf(f, "<Piece NumberOfPoints=\"%li\" NumberOfCells=\"%li\">\n", npts, ncells);
}

void points(FILE *f, int nx, int ny, int nz) {
  std::fprintf(f, "<Points>\n<DataArray type=\"Float32\" NumberOfComponents=\"3\" "
                  "format=\"ascii\">\n");
  for (int i = 0; i < nx; i++)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++) std::fprintf(f, "%e %e %e\n", (float)i, (float)j, (float)k);
  std::fprintf(f, "</DataArray>\n</Points>\n");
}

// hexahedra over an ex x ey x ez block of cells of a (ny, nz)-strided node grid
void cells(FILE *f, int ex, int ey, int ez, long sy, long sz) {
  std::fprintf(f, "<Cells>\n<DataArray type=\"Int32\" Name=\"connectivity\" format= \"ascii\">\n");
  for (int i = 0; i < ex; i++)
    for (int k = 0; k < ez; k++)
      for (int j = 0; j < ey; j++) {
        const long a = (long)i * sy * sz, b = (long)(i + 1) * sy * sz;
        const long k0 = (long)k * sy, k1 = (long)(k + 1) * sy;
        std::fprintf(f, "%ld %ld %ld %ld %ld %ld %ld %ld\n", a + k0 + j + 1, b + k0 + j + 1,
                     b + k0 + j, a + k0 + j, a + k1 + j + 1, b + k1 + j + 1, b + k1 + j,
                     a + k1 + j);
      }
  const long n = (long)ex * ey * ez;
  std::fprintf(f, "</DataArray>\n<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n");
  for (long c = 1; c <= n; c++) std::fprintf(f, "%ld\n", c * 8);
  std::fprintf(f, "</DataArray>\n<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
  for (long c = 0; c < n; c++) std::fprintf(f, "%d\n", 12);
  std::fprintf(f, "</DataArray>\n</Cells>\n");
}

void footer(FILE *f) {
  std::fprintf(f, "</DataArray>\n</CellData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
}

void padded_dims(int nelx, int nely, int nelz, t3d::Level *L) {
  struct gridContext gc;
  gc.nelx = nelx;
  gc.nely = nely;
  gc.nelz = nelz;
  t3d::compute_level(&gc, 0, L);
}

}  // namespace

extern "C" {

void writeDensityVtkFile(const int nelx, const int nely, const int nelz,
                         const DTYPE *densityArray, const char *filename) {
  t3d::Level L;
  padded_dims(nelx, nely, nelz, &L);
  const long ewy = L.wy - 1, ewz = L.wz - 1;
  FILE *f = std::fopen(filename, "w");
  if (!f) {
    t3d::set_error(TOP3D_ERR_ARG, "writeDensityVtkFile: cannot open %s", filename);
    return;
  }
  header(f, (long)(nelx + 1) * (nely + 1) * (nelz + 1), (long)nelx * nely * nelz);
  points(f, nelx + 1, nely + 1, nelz + 1);
  cells(f, nelx, nely, nelz, nely + 1, nelz + 1);
  std::fprintf(f, "<CellData>\n<DataArray type=\"Float32\" NumberOfComponents=\"1\" "
                  "Name=\"density\" format=\"ascii\">\n");
  for (long i = 1; i < nelx + 1; i++)
    for (long k = 1; k < nelz + 1; k++)
      for (long j = 1; j < nely + 1; j++)
        std::fprintf(f, "%e\n", (double)densityArray[i * ewy * ewz + k * ewy + j]);
  footer(f);
  std::fclose(f);
}

void writeDensityVtkFileWithHalo(const int nelx, const int nely, const int nelz,
                                 const DTYPE *densityArray, const char *filename) {
  t3d::Level L;
  padded_dims(nelx, nely, nelz, &L);
  const int ewx = L.wx - 1, ewy = L.wy - 1, ewz = L.wz - 1;
  FILE *f = std::fopen(filename, "w");
  if (!f) {
    t3d::set_error(TOP3D_ERR_ARG, "writeDensityVtkFileWithHalo: cannot open %s", filename);
    return;
  }
  header(f, (long)L.wx * L.wy * L.wz, (long)ewx * ewy * ewz);
  points(f, L.wx, L.wy, L.wz);
  cells(f, ewx, ewy, ewz, L.wy, L.wz);
  std::fprintf(f, "<CellData>\n<DataArray type=\"Float32\" NumberOfComponents=\"1\" "
                  "Name=\"density\" format=\"ascii\">\n");
  for (long i = 0; i < ewx; i++)
    for (long k = 0; k < ewz; k++)
      for (long j = 0; j < ewy; j++)
        std::fprintf(f, "%e\n", (double)densityArray[i * ewy * ewz + k * ewy + j]);
  footer(f);
  std::fclose(f);
}

}  // extern "C"
