// Finite-size correction from the transition structure factor
// (FiniteSizeCorrection.cxx:136-311).  S(G) lives on the momentum grid; it
// is placed on the bounding box of the grid in direct coordinates
// (coefficients with respect to the reciprocal lattice vectors), interpolated
// there with clamped Catmull-Rom tricubic splines (the reference uses the
// vendored Geometric Tools IntpTricubic3 with catmullRom true,
// gte/TricubicInterpolation.hpp:231-353: the cell index is the truncated
// coordinate clamped into the box, the four neighbours per direction are
// clamped, the fraction is not), and
//   uncorrected = sum_{G != 0} c S(G) / |G|^2
//   corrected   = sum_{G, shifts s != -G} c S~(G + s) / |G + s|^2 / (8 N^3)
// over the (2N)^3 shifts s = (a B0 + b B1 + c B2) / (2N), a, b, c in [-N, N),
// with c = unit(Gamma)^2 / unit(grid)^3 / unit(V) * unit(grid) /
// unit(Gamma)^2 * Omega / (2 pi^2).  Host code, threaded over the grid.
#include "finite_size.hpp"

#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

namespace cc4s_amd {

namespace {

struct Vec3 {
  double x, y, z;
};
inline Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator*(double s, Vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(Vec3 a, Vec3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// weights of the points i-1, i, i+1, i+2 at fraction t
inline void catmull_rom(double t, double w[4]) {
  const double t2 = t * t, t3 = t2 * t;
  w[0] = 0.5 * (-t + 2.0 * t2 - t3);
  w[1] = 0.5 * (2.0 - 5.0 * t2 + 3.0 * t3);
  w[2] = 0.5 * (t + 4.0 * t2 - 3.0 * t3);
  w[3] = 0.5 * (-t2 + t3);
}

struct Box {
  int dims[3];
  double origin[3];
  std::vector<double> F;  // F[x + dims0 (y + dims1 z)]

  double operator()(const double p[3]) const {
    int cell[3];
    double w[3][4];
    for (int d = 0; d < 3; ++d) {
      const double idx = p[d] - origin[d];
      int c = static_cast<int>(idx);  // truncation, as the reference
      c = std::min(std::max(c, 0), dims[d] - 1);
      cell[d] = c;
      catmull_rom(idx - c, w[d]);
    }
    double r = 0.0;
    for (int k = 0; k < 4; ++k) {
      const int z = std::min(std::max(cell[2] - 1 + k, 0), dims[2] - 1);
      for (int j = 0; j < 4; ++j) {
        const int y = std::min(std::max(cell[1] - 1 + j, 0), dims[1] - 1);
        for (int i = 0; i < 4; ++i) {
          const int x = std::min(std::max(cell[0] - 1 + i, 0), dims[0] - 1);
          r += w[0][i] * w[1][j] * w[2][k] * F[x + dims[0] * (y + dims[1] * z)];
        }
      }
    }
    return r;
  }
};

}  // namespace

FiniteSizeEnergy finite_size_correction(const double *grid, int64_t NG, const double *Bm,
                                        const double *S, int64_t N, double vertex_unit,
                                        double grid_unit, double potential_unit) {
  CC4S_ASSERT(NG > 0 && N > 0, "finite-size correction needs a grid and N > 0");
  const Vec3 B[3] = {{Bm[0], Bm[1], Bm[2]}, {Bm[3], Bm[4], Bm[5]}, {Bm[6], Bm[7], Bm[8]}};
  const double omega = std::fabs(dot(cross(B[0], B[1]), B[2]));
  CC4S_ASSERT(omega > 0.0, "reciprocal lattice vectors are linearly dependent");
  const Vec3 A[3] = {(1.0 / omega) * cross(B[1], B[2]), (1.0 / omega) * cross(B[2], B[0]),
                     (1.0 / omega) * cross(B[0], B[1])};
  const double to_units =
      vertex_unit * vertex_unit / (grid_unit * grid_unit * grid_unit) / potential_unit;
  const double factor = grid_unit / (vertex_unit * vertex_unit) * omega / 2.0 / M_PI / M_PI;
  std::vector<Vec3> g(NG);
  for (int64_t i = 0; i < NG; ++i) g[i] = {grid[3 * i], grid[3 * i + 1], grid[3 * i + 2]};

  // bounding box in direct coordinates; it always holds the origin
  double dmin[3] = {0, 0, 0}, dmax[3] = {0, 0, 0};
  for (int64_t i = 0; i < NG; ++i)
    for (int d = 0; d < 3; ++d) {
      const double c = dot(A[d], g[i]);
      dmin[d] = std::min(dmin[d], c);
      dmax[d] = std::max(dmax[d], c);
    }
  Box box;
  size_t size = 1;
  for (int d = 0; d < 3; ++d) {
    box.dims[d] = static_cast<int>(std::floor(dmax[d] - dmin[d] + 1.5));
    box.origin[d] = std::floor(dmin[d] + 0.5);
    size *= static_cast<size_t>(box.dims[d]);
  }
  CC4S_ASSERT(box.dims[0] >= 4 && box.dims[1] >= 4 && box.dims[2] >= 4,
              "the tricubic interpolation needs at least 4 grid points per direction");
  box.F.assign(size, 0.0);
  for (int64_t i = 0; i < NG; ++i) {
    int64_t index = 0;
    for (int d = 2; d >= 0; --d) {
      index *= box.dims[d];
      index += static_cast<int64_t>(std::floor(dot(A[d], g[i]) + 0.5) - box.origin[d]);
    }
    box.F[index] = S[i];
  }

  FiniteSizeEnergy out;
  for (int64_t i = 0; i < NG; ++i) {
    if (std::sqrt(dot(g[i], g[i])) < 1e-8) continue;
    out.uncorrected += to_units * factor / dot(g[i], g[i]) * S[i];
  }

  // fine grid: every grid point shifted by the (2N)^3 sub-cell offsets;
  // threads take grid points, partial sums are added in a fixed order
  const int nthreads = static_cast<int>(
      std::max<unsigned>(1, std::min<unsigned>(16, std::thread::hardware_concurrency())));
  std::vector<double> part(static_cast<size_t>(NG), 0.0);
  // the sub-cell steps B_d / (2N), as the reference forms them
  const double inv2n = static_cast<double>(2 * N);
  const Vec3 step[3] = {{B[0].x / inv2n, B[0].y / inv2n, B[0].z / inv2n},
                        {B[1].x / inv2n, B[1].y / inv2n, B[1].z / inv2n},
                        {B[2].x / inv2n, B[2].y / inv2n, B[2].z / inv2n}};
  auto work = [&](int tid) {
    for (int64_t i = tid; i < NG; i += nthreads) {
      double sum = 0.0;
      for (int64_t a = -N; a < N; ++a)
        for (int64_t b = -N; b < N; ++b)
          for (int64_t c = -N; c < N; ++c) {
            const Vec3 s = static_cast<double>(a) * step[0] + static_cast<double>(b) * step[1] +
                           static_cast<double>(c) * step[2];
            const Vec3 q = s + g[i];
            const double q2 = dot(q, q);
            if (std::sqrt(q2) < 1e-8) continue;
            const double p[3] = {dot(A[0], q), dot(A[1], q), dot(A[2], q)};
            sum += to_units * box(p) * factor / q2;
          }
      part[i] = sum;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nthreads; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto &t : pool) t.join();
  double total = 0.0;
  for (double v : part) total += v;
  out.corrected = total / static_cast<double>(N) / N / N / 8.0;
  return out;
}

}  // namespace cc4s_amd
