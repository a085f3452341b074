// CPU check of the coarse-level elimination schedule (coarse_schedule.cpp):
// runs the workgroup tasks of every launch in plain double arithmetic, the
// way the GPU kernels do, on a random SPD matrix with the plane (block
// tridiagonal) or band structure, and checks
//   - the tasks of one launch never write a tile another task of that launch
//     reads or writes, and every diagonal tile is factored once, before use;
//   - L L^T = P A P^T and W L = I with W built from the row lists.
// Usage: coarse_schedule_check n bw planes seed   (prints one summary line)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <set>
#include <vector>

#include "../../topopt-in-openmp-gpu_amd/csrc/coarse_schedule.hpp"

using namespace t3d;
constexpr int B = kTile;

struct Mat {
  int np;
  std::vector<double> v;
  double &at(int r, int c) { return v[(size_t)r * np + c]; }
};

static void tile_get(Mat &M, int I, int J, double *t) {
  for (int r = 0; r < B; r++)
    for (int c = 0; c < B; c++) t[r * B + c] = M.at(I * B + r, J * B + c);
}
static void tile_put(Mat &M, int I, int J, const double *t) {
  for (int r = 0; r < B; r++)
    for (int c = 0; c < B; c++) M.at(I * B + r, J * B + c) = t[r * B + c];
}
// out = X Y^T
static void mm_nt(const double *X, const double *Y, double *out) {
  for (int r = 0; r < B; r++)
    for (int c = 0; c < B; c++) {
      double s = 0;
      for (int k = 0; k < B; k++) s += X[r * B + k] * Y[c * B + k];
      out[r * B + c] = s;
    }
}

int main(int argc, char **argv) {
  if (argc < 5) return 2;
  const int n = std::atoi(argv[1]), bw = std::atoi(argv[2]), planes = std::atoi(argv[3]);
  std::mt19937_64 rng(std::atoi(argv[4]));
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  CoarseSchedule cs;
  if (!build_coarse_schedule(n, bw, planes, cs)) {
    std::printf("FAIL schedule: %s\n", cs.error.c_str());
    return 1;
  }
  const int np = cs.np, nb = cs.nblk;
  // random SPD matrix: planes coupled to neighbours only, or a band
  const int pm = planes > 1 ? n / planes : n;
  std::vector<double> A0((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < i; j++) {
      const bool nz = planes > 1 ? std::abs(i / pm - j / pm) <= 1 : i - j <= bw;
      if (nz) A0[(size_t)i * n + j] = A0[(size_t)j * n + i] = U(rng);
    }
  for (int i = 0; i < n; i++) {
    double s = 0;
    for (int j = 0; j < n; j++) s += std::fabs(A0[(size_t)i * n + j]);
    A0[(size_t)i * n + i] = s + 1.0;
  }
  Mat A{np, std::vector<double>((size_t)np * np, 0.0)};
  std::vector<int> ipos(np, -1);
  for (int d = 0; d < n; d++) ipos[cs.pos[d]] = d;
  for (int s = 0; s < np; s++)
    if (ipos[s] < 0) A.at(s, s) = 1.0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) A.at(cs.pos[i], cs.pos[j]) = A0[(size_t)i * n + j];
  Mat P = A;  // the permuted matrix, for the final check
  Mat Lf{np, std::vector<double>((size_t)np * np, 0.0)};
  std::vector<double> Dinv((size_t)nb * B * B, 0.0);
  std::vector<char> factored(nb, 0);
  std::vector<double> t0(B * B), t1(B * B), t2(B * B), t3(B * B), t4(B * B);
  // Cholesky of a tile in place (lower), inverse into X
  auto factor = [&](double *T, double *X) {
    for (int j = 0; j < B; j++) {
      double d = T[j * B + j];
      for (int k = 0; k < j; k++) d -= T[j * B + k] * T[j * B + k];
      if (!(d > 0)) return false;
      d = std::sqrt(d);
      T[j * B + j] = d;
      for (int i = j + 1; i < B; i++) {
        double s = T[i * B + j];
        for (int k = 0; k < j; k++) s -= T[i * B + k] * T[j * B + k];
        T[i * B + j] = s / d;
      }
      for (int c = j + 1; c < B; c++) T[j * B + c] = 0.0;
    }
    for (int c = 0; c < B; c++)
      for (int i = 0; i < B; i++) {
        double s = (i == c) ? 1.0 : 0.0;
        for (int m = 0; m < i; m++) s -= T[i * B + m] * X[m * B + c];
        X[i * B + c] = s / T[i * B + i];
      }
    return true;
  };
  int launches = 0, ntasks = 0;
  for (size_t q = 0; q < cs.launch_first.size(); q++) {
    launches++;
    // race check within the launch (tiles of A, Lf and Dinv as (matrix, I, J))
    std::map<std::tuple<int, int, int>, int> writer;
    std::vector<std::pair<std::tuple<int, int, int>, int>> reads;
    for (int t = cs.launch_first[q]; t < cs.launch_first[q] + cs.launch_count[q]; t++) {
      const CholTask &k = cs.tasks[t];
      if (k.I < 0 || k.I >= nb || k.J < 0 || k.J > k.I || k.nsrc < 0 || k.nsrc > kMaxSrc) {
        std::printf("FAIL task %d out of range\n", t);
        return 1;
      }
      auto W = [&](int m, int I, int J) {
        if (!writer.emplace(std::make_tuple(m, I, J), t).second) {
          std::printf("FAIL launch %zu: tile (%d,%d,%d) written twice\n", q, m, I, J);
          std::exit(1);
        }
      };
      auto R = [&](int m, int I, int J) { reads.push_back({std::make_tuple(m, I, J), t}); };
      for (int s = 0; s < k.nsrc; s++) {
        if (k.K[s] >= k.J || !factored[k.K[s]]) {
          std::printf("FAIL task %d: source %d not factored or not left of J\n", t, k.K[s]);
          return 1;
        }
        R(0, k.I, k.K[s]);
        R(0, k.J, k.K[s]);
        R(2, k.K[s], k.K[s]);
        if (k.I == k.J) W(1, k.I, k.K[s]);
      }
      R(0, k.I, k.J);
      if (k.flags & kTaskFactor) {
        if (k.I != k.J || factored[k.I]) {
          std::printf("FAIL task %d: factor of tile %d twice or off-diagonal\n", t, k.I);
          return 1;
        }
        W(1, k.I, k.I);
        W(2, k.I, k.I);
      } else {
        W(0, k.I, k.J);
      }
    }
    for (auto &r : reads) {
      auto it = writer.find(r.first);
      if (it != writer.end() && it->second != r.second) {
        std::printf("FAIL launch %zu: task %d reads a tile task %d writes\n", q, r.second,
                    it->second);
        return 1;
      }
    }
    // execute the launch (tasks are independent, run in order)
    std::vector<int> newly;
    for (int t = cs.launch_first[q]; t < cs.launch_first[q] + cs.launch_count[q]; t++) {
      const CholTask &k = cs.tasks[t];
      ntasks++;
      tile_get(A, k.I, k.J, t0.data());  // C
      for (int s = 0; s < k.nsrc; s++) {
        const double *D = &Dinv[(size_t)k.K[s] * B * B];
        tile_get(A, k.I, k.K[s], t1.data());
        mm_nt(t1.data(), D, t2.data());  // Q_I
        if (k.I == k.J) {
          tile_put(Lf, k.I, k.K[s], t2.data());
          mm_nt(t2.data(), t2.data(), t4.data());
        } else {
          tile_get(A, k.J, k.K[s], t1.data());
          mm_nt(t1.data(), D, t3.data());  // Q_J
          mm_nt(t2.data(), t3.data(), t4.data());
        }
        for (int e = 0; e < B * B; e++) t0[e] -= t4[e];
      }
      if (k.flags & kTaskFactor) {
        std::fill(t1.begin(), t1.end(), 0.0);
        if (!factor(t0.data(), t1.data())) {
          std::printf("FAIL non-positive pivot in tile %d\n", k.I);
          return 1;
        }
        tile_put(Lf, k.I, k.I, t0.data());
        std::copy(t1.begin(), t1.end(), Dinv.begin() + (size_t)k.I * B * B);
        newly.push_back(k.I);
      } else {
        tile_put(A, k.I, k.J, t0.data());
      }
    }
    for (int I : newly) factored[I] = 1;
  }
  for (int I = 0; I < nb; I++)
    if (!factored[I]) {
      std::printf("FAIL tile %d never factored\n", I);
      return 1;
    }
  // L L^T = P A P^T
  double err = 0, ref = 0;
  for (int i = 0; i < np; i++)
    for (int j = 0; j <= i; j++) {
      double s = 0;
      for (int k = 0; k <= j; k++) s += Lf.at(i, k) * Lf.at(j, k);
      err = std::fmax(err, std::fabs(s - P.at(i, j)));
      ref = std::fmax(ref, std::fabs(P.at(i, j)));
    }
  // W from the row lists; W L = I
  Mat Wm{np, std::vector<double>((size_t)np * np, 0.0)};
  size_t wtiles = 0;
  for (int J = 0; J < nb; J++) {
    for (int r = 0; r < B; r++)
      for (int c = 0; c < B; c++) Wm.at(J * B + r, J * B + c) = Dinv[(size_t)J * B * B + r * B + c];
    wtiles++;
    for (int q = 0; q < cs.colinfo[2 * J + 1]; q++) {
      const int *rw = &cs.rows[4 * (cs.colinfo[2 * J] + q)];
      const int I = rw[0];
      std::fill(t0.begin(), t0.end(), 0.0);
      for (int t = 0; t < rw[2]; t++) {
        const int K = cs.klist[rw[1] + t];
        for (int r = 0; r < B; r++)
          for (int c = 0; c < B; c++) {
            double s = 0;
            for (int m = 0; m < B; m++) s += Lf.at(I * B + r, K * B + m) * Wm.at(K * B + m, J * B + c);
            t0[r * B + c] += s;
          }
      }
      const double *D = &Dinv[(size_t)I * B * B];
      for (int r = 0; r < B; r++)
        for (int c = 0; c < B; c++) {
          double s = 0;
          for (int m = 0; m < B; m++) s += D[r * B + m] * t0[m * B + c];
          Wm.at(I * B + r, J * B + c) = -s;
        }
      wtiles++;
    }
  }
  double werr = 0;
  for (int i = 0; i < np; i++)
    for (int j = 0; j <= i; j++) {
      double s = 0;
      for (int k = j; k <= i; k++) s += Wm.at(i, k) * Lf.at(k, j);
      werr = std::fmax(werr, std::fabs(s - (i == j ? 1.0 : 0.0)));
    }
  const bool ok = err <= 1e-12 * ref && werr <= 1e-10;
  std::printf("%s n=%d np=%d launches=%d tasks=%d llt_err=%.3e wl_err=%.3e w_tiles=%zu\n",
              ok ? "OK" : "FAIL", n, np, launches, ntasks, err / ref, werr, wtiles);
  return ok ? 0 : 1;
}
