// scenes.cpp -- seeded synthetic scene generators for the benchmark workloads
// (host-only; not part of the per-step pipeline). All randomness comes from
// splitmix64, so a seed reproduces a scene bit for bit on any machine.
//
//  * lattice2d: equal circles on a jittered square lattice in a box.
//  * rsa2d:     random sequential addition of circles with uniform radii, largest
//               first, into a square box sized for a target area fraction.
//  * rsa3d:     random sequential addition of spheres with log-normal radii into a
//               box of given base, height sized for a target volume fraction.
// Body order is shuffled after placement so that index order is unrelated to size.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {
    // Box-Muller
    double u1 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

template <int D>
struct Grid {
  double h, lo[D];
  int dims[D];
  std::vector<std::vector<uint32_t>> cells;
  Grid(double h_, const double* lo_, const double* hi_) : h(h_) {
    size_t total = 1;
    for (int a = 0; a < D; ++a) {
      lo[a] = lo_[a];
      dims[a] = std::max(1, (int)std::ceil((hi_[a] - lo_[a]) / h));
      total *= (size_t)dims[a];
    }
    cells.resize(total);
  }
  int coord(double x, int a) const { return std::min(dims[a] - 1, std::max(0, (int)((x - lo[a]) / h))); }
  size_t index(const int* c) const {
    size_t k = 0;
    for (int a = D - 1; a >= 0; --a) k = k * (size_t)dims[a] + (size_t)c[a];
    return k;
  }
};

template <int D>
bool free_at(const Grid<D>& g, const std::vector<double>& q, const std::vector<double>& r, const double* x, double rx) {
  int c[D];
  for (int a = 0; a < D; ++a) c[a] = g.coord(x[a], a);
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int a = 0; a < D; ++a) {
    lo[a] = std::max(0, c[a] - 1);
    hi[a] = std::min(g.dims[a] - 1, c[a] + 1);
  }
  int k[3];
  for (k[2] = (D == 3 ? lo[2] : 0); k[2] <= (D == 3 ? hi[2] : 0); ++k[2])
    for (k[1] = lo[1]; k[1] <= hi[1]; ++k[1])
      for (k[0] = lo[0]; k[0] <= hi[0]; ++k[0]) {
        for (uint32_t j : g.cells[g.index(k)]) {
          double d2 = 0.0;
          for (int a = 0; a < D; ++a) {
            const double dd = x[a] - q[(size_t)D * j + a];
            d2 += dd * dd;
          }
          const double rs = rx + r[j];
          if (d2 < rs * rs) return false;
        }
      }
  return true;
}

template <int D>
int rsa(uint32_t n, std::vector<double>& radii, const double* lo, const double* hi, uint64_t seed, double* q_out,
        double* r_out, uint64_t max_attempts) {
  std::sort(radii.begin(), radii.end(), [](double a, double b) { return a > b; });
  const double rmax = radii.empty() ? 1.0 : radii[0];
  Grid<D> g(2.0 * rmax, lo, hi);
  Rng rng(seed ^ 0xA5A5A5A5DEADBEEFull);
  std::vector<double> q((size_t)D * n), r(n);
  for (uint32_t i = 0; i < n; ++i) {
    const double ri = radii[i];
    double x[D];
    bool placed = false;
    for (uint64_t att = 0; att < max_attempts; ++att) {
      for (int a = 0; a < D; ++a) x[a] = (lo[a] + ri) + rng.uniform() * ((hi[a] - lo[a]) - 2.0 * ri);
      if (free_at<D>(g, q, r, x, ri)) {
        placed = true;
        break;
      }
    }
    if (!placed) return -(int)(i + 1);
    for (int a = 0; a < D; ++a) q[(size_t)D * i + a] = x[a];
    r[i] = ri;
    int c[D];
    for (int a = 0; a < D; ++a) c[a] = g.coord(x[a], a);
    g.cells[g.index(c)].push_back(i);
  }
  // shuffle body order (Fisher-Yates)
  std::vector<uint32_t> perm(n);
  for (uint32_t i = 0; i < n; ++i) perm[i] = i;
  for (uint32_t i = n; i > 1; --i) std::swap(perm[i - 1], perm[rng.next() % i]);
  for (uint32_t i = 0; i < n; ++i) {
    for (int a = 0; a < D; ++a) q_out[(size_t)D * i + a] = q[(size_t)D * perm[i] + a];
    r_out[i] = r[perm[i]];
  }
  return 0;
}

}  // namespace

extern "C" {

// Equal circles of radius r on an m x m lattice of spacing h = side / m centred in
// [-side/2, side/2]^2, each jittered uniformly by +-(h/2 - r - 1e-6) per axis; initial
// velocities N(0, vstd^2) per component.
int scn_lattice2d(uint32_t m, double side, double r, double vstd, uint64_t seed, double* q, double* v) {
  Rng rng(seed);
  const double h = side / m;
  const double jit = h / 2.0 - r - 1.0e-6;
  if (jit < 0.0) return -1;
  for (uint32_t iy = 0; iy < m; ++iy)
    for (uint32_t ix = 0; ix < m; ++ix) {
      const size_t i = (size_t)iy * m + ix;
      q[2 * i] = -side / 2.0 + h * (ix + 0.5) + jit * (2.0 * rng.uniform() - 1.0);
      q[2 * i + 1] = -side / 2.0 + h * (iy + 0.5) + jit * (2.0 * rng.uniform() - 1.0);
    }
  for (uint32_t i = 0; i < m * m; ++i) {
    v[2 * i] = vstd * rng.normal();
    v[2 * i + 1] = vstd * rng.normal();
  }
  return 0;
}

// Circles with radii U[rmin, rmax], placed largest first at random non-overlapping
// positions in a square box of side L = sqrt(sum(pi r^2) / phi) centred at the origin.
// Writes q (2n), r (n); returns 0, or -(i+1) if circle i found no free spot.
int scn_rsa2d(uint32_t n, double rmin, double rmax, double phi, uint64_t seed, uint64_t max_attempts, double* q,
              double* r, double* side_out) {
  Rng rng(seed);
  std::vector<double> radii(n);
  double area = 0.0;
  for (uint32_t i = 0; i < n; ++i) {
    radii[i] = rmin + (rmax - rmin) * rng.uniform();
    area += 3.141592653589793 * radii[i] * radii[i];
  }
  const double L = std::sqrt(area / phi);
  if (side_out) *side_out = L;
  const double lo[2] = {-L / 2, -L / 2}, hi[2] = {L / 2, L / 2};
  return rsa<2>(n, radii, lo, hi, seed, q, r, max_attempts);
}

// Spheres with radii rmed * exp(sigma Z), Z ~ N(0,1) (clamped to [rclamp_lo, rclamp_hi]),
// placed largest first in the box [-base/2, base/2] x [0, H] x [-base/2, base/2] with
// H = sum(4/3 pi r^3) / (phi base^2). Writes q (3n), r (n), the height.
int scn_rsa3d(uint32_t n, double rmed, double sigma, double rclamp_lo, double rclamp_hi, double base, double phi,
              uint64_t seed, uint64_t max_attempts, double* q, double* r, double* height_out) {
  Rng rng(seed);
  std::vector<double> radii(n);
  double vol = 0.0;
  for (uint32_t i = 0; i < n; ++i) {
    double x = rmed * std::exp(sigma * rng.normal());
    x = std::min(rclamp_hi, std::max(rclamp_lo, x));
    radii[i] = x;
    vol += 4.0 / 3.0 * 3.141592653589793 * x * x * x;
  }
  const double H = vol / (phi * base * base);
  if (height_out) *height_out = H;
  const double lo[3] = {-base / 2, 0.0, -base / 2}, hi[3] = {base / 2, H, base / 2};
  return rsa<3>(n, radii, lo, hi, seed, q, r, max_attempts);
}

// A poured pile: spheres with radii rmed * exp(sigma Z) (clamped) are dropped one at a
// time, in index order, at uniform (x, z) over the base [-base/2, base/2]^2 onto the
// floor y = 0 and the spheres already placed, and roll downhill until they rest.
// Rolling is a projected descent: a trial move straight down is pushed back out of every
// sphere it enters (and kept inside the walls and above the floor); the move is kept if
// it lowers the sphere without overlap, otherwise the step is halved, until the step is
// below 1e-7 r. A sphere so comes to rest in a pocket of (typically) three supports,
// which gives a pile with about six contacts per sphere. Each sphere is finally lowered
// by `sink` so that it overlaps its supports slightly and its contacts are active at the
// first step. Radii: uniform_radii = 0, rmed exp(sigma Z) clamped to [rclamp_lo,
// rclamp_hi]; otherwise uniform in [rclamp_lo, rclamp_hi]. Writes q (3n), r (n) and the
// pile height.
int scn_deposit3d(uint32_t n, double rmed, double sigma, double rclamp_lo, double rclamp_hi, int uniform_radii,
                  double base, uint64_t seed, double sink, double* q, double* r, double* height_out) {
  Rng rng(seed);
  std::vector<double> radii(n);
  double vol = 0.0, rmax = 0.0, dsum = 0.0;
  for (uint32_t i = 0; i < n; ++i) {
    double x = uniform_radii ? rclamp_lo + (rclamp_hi - rclamp_lo) * rng.uniform() : rmed * std::exp(sigma * rng.normal());
    x = std::min(rclamp_hi, std::max(rclamp_lo, x));
    radii[i] = x;
    rmax = std::max(rmax, x);
    dsum += 2 * x;
    vol += 4.0 / 3.0 * 3.141592653589793 * x * x * x;
  }
  if (n == 0) return 0;
  if (!(base > 4 * rmax)) return -1;
  // generous height bound for the grid: a loose pile (volume fraction 0.3) plus margin
  const double hmax = vol / (0.3 * base * base) + 4 * rmax;
  const double lo[3] = {-base / 2, 0.0, -base / 2}, hi[3] = {base / 2, hmax, base / 2};
  Grid<3> grid(dsum / n, lo, hi);  // cells of the mean diameter, one sphere each on average
  const int ncx = grid.dims[0];
  std::vector<std::vector<uint32_t>> cols((size_t)grid.dims[0] * grid.dims[2]);
  std::vector<double> colmax((size_t)grid.dims[0] * grid.dims[2], -1.0);
  double top = 0.0;
  std::vector<uint32_t> cand;
  // spheres whose inflated surface may be reached from within `margin` of g
  auto gather = [&](const double* g, double ri, double margin) {
    cand.clear();
    const double reach = ri + rmax + margin;
    int c0[3], c1[3];
    for (int a = 0; a < 3; ++a) {
      c0[a] = grid.coord(g[a] - reach, a);
      c1[a] = grid.coord(g[a] + reach, a);
    }
    int c[3];
    for (c[2] = c0[2]; c[2] <= c1[2]; ++c[2])
      for (c[1] = c0[1]; c[1] <= c1[1]; ++c[1])
        for (c[0] = c0[0]; c[0] <= c1[0]; ++c[0])
          for (uint32_t j : grid.cells[grid.index(c)]) {
            const double dx = g[0] - q[3 * j], dy = g[1] - q[3 * j + 1], dz = g[2] - q[3 * j + 2];
            const double R = ri + r[j] + margin;
            if (dx * dx + dy * dy + dz * dz < R * R) cand.push_back(j);
          }
  };
  for (uint32_t i = 0; i < n; ++i) {
    const double ri = radii[i];
    double p[3];
    for (int attempt = 0;; ++attempt) {
      p[0] = -base / 2 + ri + (base - 2 * ri) * rng.uniform();
      p[2] = -base / 2 + ri + (base - 2 * ri) * rng.uniform();
      // vertical drop: the highest first contact along the line, or the floor
      double y = ri;
      {
        const int cx0 = grid.coord(p[0] - ri - rmax, 0), cx1 = grid.coord(p[0] + ri + rmax, 0);
        const int cz0 = grid.coord(p[2] - ri - rmax, 2), cz1 = grid.coord(p[2] + ri + rmax, 2);
        for (int cz = cz0; cz <= cz1; ++cz)
          for (int cx = cx0; cx <= cx1; ++cx) {
            const size_t col = (size_t)cz * ncx + cx;
            if (colmax[col] + ri + rmax <= y) continue;
            for (uint32_t j : cols[col]) {
              const double dx = p[0] - q[3 * j], dz = p[2] - q[3 * j + 2], R = ri + r[j];
              const double h2 = dx * dx + dz * dz;
              if (h2 < R * R) y = std::max(y, q[3 * j + 1] + std::sqrt(R * R - h2));
            }
          }
      }
      p[1] = y;
      // roll downhill: trial moves of `step` straight down, projected out of the
      // candidates; candidates are gathered with margin ri around g and refreshed when
      // the sphere has moved ri/2 from g (a trial point then stays within ri of g)
      double g[3] = {p[0], p[1], p[2]};
      gather(g, ri, ri);
      double step = 0.5 * ri;
      // at most 400 trial moves; a sphere still creeping along a shallow groove after them
      // (supported, but not in a pocket) is dropped again elsewhere
      int trial = 0;
      for (trial = 0; trial < 400 && step > 1e-6 * ri; ++trial) {
        const double mx = p[0] - g[0], my = p[1] - g[1], mz = p[2] - g[2];
        if (mx * mx + my * my + mz * mz > 0.25 * ri * ri) {
          g[0] = p[0];
          g[1] = p[1];
          g[2] = p[2];
          gather(g, ri, ri);
        }
        double t[3] = {p[0], p[1] - step, p[2]};
        bool ok = false;
        for (int pass = 0; pass < 16; ++pass) {
          double worst = 0.0;
          t[1] = std::max(t[1], ri);
          for (int a = 0; a < 3; a += 2) t[a] = std::min(base / 2 - ri, std::max(-base / 2 + ri, t[a]));
          for (uint32_t j : cand) {
            const double dx = t[0] - q[3 * j], dy = t[1] - q[3 * j + 1], dz = t[2] - q[3 * j + 2];
            const double R = ri + r[j];
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < R * R) {
              const double d = std::sqrt(d2);
              worst = std::max(worst, R - d);
              if (d > 0) {
                const double sc = R / d;
                t[0] = q[3 * j] + dx * sc;
                t[1] = q[3 * j + 1] + dy * sc;
                t[2] = q[3 * j + 2] + dz * sc;
              } else {
                t[1] = q[3 * j + 1] + R;
              }
            }
          }
          if (worst <= 1e-12 * ri && t[1] >= ri && std::fabs(t[0]) <= base / 2 - ri && std::fabs(t[2]) <= base / 2 - ri) {
            ok = true;
            break;
          }
        }
        const double dmx = t[0] - g[0], dmy = t[1] - g[1], dmz = t[2] - g[2];
        // keep moves that descend by a fair part of the step; shallow ones halve it (a
        // sphere sliding down a nearly level groove would otherwise creep at full step)
        if (ok && t[1] < p[1] - 0.05 * step && dmx * dmx + dmy * dmy + dmz * dmz < ri * ri) {
          p[0] = t[0];
          p[1] = t[1];
          p[2] = t[2];
          step = std::min(0.5 * ri, 1.5 * step);
        } else {
          step *= 0.5;
        }
      }
      if (trial < 400 || attempt >= 63) break;  // in its pocket (no lower position within reach)
    }
    p[1] -= sink;
    q[3 * i] = p[0];
    q[3 * i + 1] = p[1];
    q[3 * i + 2] = p[2];
    r[i] = ri;
    if (p[1] + ri > hmax - 2 * rmax) return -(int)i - 2;  // grid too short
    int c[3];
    for (int a = 0; a < 3; ++a) c[a] = grid.coord(p[a], a);
    grid.cells[grid.index(c)].push_back(i);
    const size_t col = (size_t)c[2] * ncx + c[0];
    cols[col].push_back(i);
    colmax[col] = std::max(colmax[col], p[1]);
    top = std::max(top, p[1] + ri);
  }
  if (height_out) *height_out = top;
  return 0;
}

}  // extern "C"
