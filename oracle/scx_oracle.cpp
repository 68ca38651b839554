// scx_oracle.cpp -- FP64 CPU restatement of SCISim's per-step contact pipeline
// for circle (2D) and sphere (3D) scenes.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP library is
// compared against. It is imported by tests/, by __graft_entry__.smoke() and by
// the cpu_baseline leg of bench.py, and by nothing else. It is never the thing
// measured or shipped.
//
// Every function cites the reference file:line it restates (paths relative to
// the breannansmith/scisim source tree).
//
// Parity pinning (what checks each part):
//   * grid broad phase / brute force: the reference's own AABB fixtures
//     (ball2dtests/collision_detection_tests.cpp:12-16) with the reference's own
//     criterion (grid == all-pairs, :34-40), plus 3D generated scenes vs brute force.
//   * CCD: the 11 known-answer tests of scisimtests/narrowphase_tests.cpp.
//   * bases / H / Delassus diagonal: orthonormality + closed forms
//     (W_ii = (1/ma+1/mb) I in 2D; diag(w_n, w_t, w_t) in 3D).
//   * local Coulomb solve + FB residual: closed-form contact modes (separate /
//     stick / slide give a zero residual), brute-force enumeration of contact modes
//     on tiny problems, the analytic slope tests of
//     assets/ball2d/shell_scripts/{sliding,static}_particle_test.*.
//   * integrators: closed-form parabolas under constant gravity.
//   * the So-bogus iterate sequence itself (skip heuristics, exact local-solver
//     branches, initial-evaluation rules) is NOT available offline: parity unpinned
//     at that level; results are compared at the residual tolerance instead.

#include <algorithm>
#include <array>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <set>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();

// ---------------------------------------------------------------------------
// Broad phase: ball2d/SpatialGridDetector.cpp:15-147 (2D) and
// rigidbody3d/SpatialGridDetector.cpp:8-151 (3D).
// ---------------------------------------------------------------------------

// AABB::overlaps, ball2d/SpatialGridDetector.cpp:15-26: boxes overlap unless a
// separating axis exists with a strict max < min.
template <int D>
inline bool aabb_overlaps(const double* amin, const double* amax, const double* bmin, const double* bmax) {
  for (int a = 0; a < D; ++a)
    if (amax[a] < bmin[a]) return false;
  for (int a = 0; a < D; ++a)
    if (bmax[a] < amin[a]) return false;
  return true;
}

// initializeSpatialGrid, ball2d/SpatialGridDetector.cpp:76-104.
template <int D>
void init_grid(uint32_t n, const double* mn, const double* mx, double* min_coord, uint32_t* dims, double* h) {
  double max_coord[D];
  for (int a = 0; a < D; ++a) { min_coord[a] = kInf; max_coord[a] = -kInf; }
  for (uint32_t i = 0; i < n; ++i)
    for (int a = 0; a < D; ++a) {
      min_coord[a] = std::min(min_coord[a], mn[D * i + a]);
      max_coord[a] = std::max(max_coord[a], mx[D * i + a]);
    }
  for (int a = 0; a < D; ++a) { min_coord[a] -= 2.0e-6; max_coord[a] += 2.0e-6; }
  double delta[D];
  for (int a = 0; a < D; ++a) delta[a] = 0.0;
  for (uint32_t i = 0; i < n; ++i)
    for (int a = 0; a < D; ++a) delta[a] += mx[D * i + a] - mn[D * i + a];
  double dmax = delta[0];
  for (int a = 1; a < D; ++a) dmax = std::max(dmax, delta[a]);
  *h = dmax / double(n);
  for (int a = 0; a < D; ++a) dims[a] = (uint32_t)std::ceil((max_coord[a] - min_coord[a]) / *h);
}

// getPotentialOverlaps, ball2d/SpatialGridDetector.cpp:106-133 with
// rasterizeAABBs :42-74 and keyForIndex :36-40 (u32 row-major key, wrapping as
// the reference's unsigned arithmetic does). Output: lexicographic (i<j) set.
template <int D>
void grid_pairs(uint32_t n, const double* mn, const double* mx, std::vector<std::pair<uint32_t, uint32_t>>& out) {
  out.clear();
  if (n == 0) return;
  double min_coord[D];
  uint32_t dims[D];
  double h;
  init_grid<D>(n, mn, mx, min_coord, dims, &h);
  std::map<uint32_t, std::vector<uint32_t>> voxels;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t lo[D], hi[D];
    for (int a = 0; a < D; ++a) {
      lo[a] = (uint32_t)(((mn[D * i + a] - 1.0e-6) - min_coord[a]) / h);
      hi[a] = (uint32_t)(((mx[D * i + a] + 1.0e-6) - min_coord[a]) / h);
    }
    if constexpr (D == 2) {
      for (uint32_t x = lo[0]; x <= hi[0]; ++x)
        for (uint32_t y = lo[1]; y <= hi[1]; ++y) voxels[x + dims[0] * y].push_back(i);
    } else {
      for (uint32_t x = lo[0]; x <= hi[0]; ++x)
        for (uint32_t y = lo[1]; y <= hi[1]; ++y)
          for (uint32_t z = lo[2]; z <= hi[2]; ++z)
            voxels[x + dims[0] * y + dims[0] * dims[1] * z].push_back(i);
    }
  }
  std::set<std::pair<uint32_t, uint32_t>> s;
  for (const auto& kv : voxels) {
    const std::vector<uint32_t>& ids = kv.second;
    for (size_t a = 0; a + 1 < ids.size(); ++a)
      for (size_t b = a + 1; b < ids.size(); ++b) {
        const uint32_t i = ids[a], j = ids[b];
        if (aabb_overlaps<D>(mn + D * i, mx + D * i, mn + D * j, mx + D * j)) s.insert(std::make_pair(i, j));
      }
  }
  out.assign(s.begin(), s.end());
}

// getPotentialOverlapsAllPairs, ball2d/SpatialGridDetector.cpp:135-147.
template <int D>
void all_pairs(uint32_t n, const double* mn, const double* mx, std::vector<std::pair<uint32_t, uint32_t>>& out) {
  out.clear();
  std::vector<std::vector<uint32_t>> per(n);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    const uint32_t i = (uint32_t)ii;
    for (uint32_t j = i + 1; j < n; ++j)
      if (aabb_overlaps<D>(mn + D * i, mx + D * i, mn + D * j, mx + D * j)) per[i].push_back(j);
  }
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t j : per[i]) out.emplace_back(i, j);
}

// ---------------------------------------------------------------------------
// CCD: scisim/CollisionDetection/CollisionDetectionUtilities.cpp:3-126.
// ---------------------------------------------------------------------------
inline void ccd_coeffs(const double* q0a, const double* q1a, double ra, const double* q0b, const double* q1b, double rb,
                       double c[3]) {
  const double d0x = q0a[0] - q0b[0], d0y = q0a[1] - q0b[1];
  const double dqx = (q1a[0] - q1b[0]) - d0x, dqy = (q1a[1] - q1b[1]) - d0y;
  const double rs = ra + rb;
  c[0] = (d0x * d0x + d0y * d0y) - rs * rs;
  c[1] = 2.0 * (d0x * dqx + d0y * dqy);
  c[2] = dqx * dqx + dqy * dqy;
}

// ballBallCCDCollisionHappens(c), CollisionDetectionUtilities.cpp:60-121 with the
// stable root helpers :24-58.
inline bool ccd_happens(const double c[3], double* t) {
  if (c[2] != 0.0) {
    const double c1c1 = c[1] * c[1];
    const double fc2c0 = 4 * c[2] * c[0];
    if (c1c1 < fc2c0) return false;
    const double ds = std::sqrt(c1c1 - fc2c0);
    // secondRootOfQuadratic :42-58
    const double root1 = (c[1] > 0.0) ? (2.0 * c[0]) / (-c[1] - ds) : (-c[1] + ds) / (2.0 * c[2]);
    if (root1 < 0.0) return false;
    // firstRootOfQuadratic :24-40
    const double root0 = (c[1] >= 0.0) ? (-c[1] - ds) / (2.0 * c[2]) : (2.0 * c[0]) / (-c[1] + ds);
    if (root0 > 1.0) return false;
    *t = std::max(0.0, root0);
    return true;
  }
  if (c[0] <= 0.0) {
    *t = 0.0;
    return true;
  }
  return false;
}

// ---------------------------------------------------------------------------
// Small FP64 vector helpers. Sums are left to right and no FMA contraction is
// allowed (the Makefile builds with -ffp-contract=off), matching the
// reference's GCC -std=c++14 build (CMakeLists.txt:53-56).
// ---------------------------------------------------------------------------
inline double dot2(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1]; }
inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
template <int D>
inline void normalize(double* v) {
  double z = 0.0;
  for (int a = 0; a < D; ++a) z = (a == 0) ? v[0] * v[0] : z + v[a] * v[a];
  if (z > 0.0) {
    const double s = std::sqrt(z);
    for (int a = 0; a < D; ++a) v[a] = v[a] / s;
  }
}

// FrictionUtilities::orthogonalVector, rigidbody3d/Constraints/FrictionUtilities.cpp:19-29.
inline void orthogonal_vector(const double* n, double* o) {
  double e[3] = {0, 0, 0};
  if (std::fabs(n[0]) <= std::fabs(n[1]) && std::fabs(n[0]) <= std::fabs(n[2]))
    e[0] = 1.0;
  else if (std::fabs(n[1]) <= std::fabs(n[2]))
    e[1] = 1.0;
  else
    e[2] = 1.0;
  const double d = dot3(n, e);
  for (int a = 0; a < 3; ++a) o[a] = e[a] - d * n[a];
  normalize<3>(o);
}

// ---------------------------------------------------------------------------
// Contacts and the So-bogus-shaped problem.
// ---------------------------------------------------------------------------
struct Contact {
  int32_t a, b;      // body indices; b = -1 for a static plane
  int32_t plane;     // plane index or -1
  double basis[9];   // columns [n s t] (2D: [n t]), column-major D x D
  double pa[3];      // arm of body a (3D): contact point - x_a, at q0
  double pb[3];      // arm of body b (3D)
  double ndotv;      // evalNdotV at (q0, v0)
  double nrel;       // extra normal forcing (nrel_extra of FrictionSolver::solve)
  double drel[2];    // extra tangential forcing (drel_extra), D-1 used
};

struct Params {
  double dt;
  double g[3];
  double mu;
  double CoR;
  double tol;
  uint32_t max_iters;
  uint32_t eval_every;
  int integrator;  // 0 symplectic Euler, 1 Verlet (3D: SplitHam translational part)
  int order;       // 0 natural active-set order, 1 colour order
  int warm_start;  // ImpulsesToCache: 0 NONE, 1 NORMAL_AND_FRICTION (2D only), 2 NORMAL,
                   // 3 world impulse transported with the normal (SE-map cache)
};

// Fischer-Burmeister residual of one contact with the De Saxce change of variable
// (So-bogus SOCLaw<D,double,true>::eval, external; semantics restated from the
// public So-bogus source, see scisim/ConstrainedMaps/bogus/Ball2DSobogusInterface.cpp:150-168
// for the call site). Returns ||phi||^2; the aggregated error is its max over contacts.
// fb_phi writes the vector phi itself (D entries); fb_err returns ||phi||^2.
template <int D>
void fb_phi(double mu, const double* r, const double* u, double* phi) {
  double uT2 = 0.0;
  for (int a = 1; a < D; ++a) uT2 += u[a] * u[a];
  const double un_hat = u[0] + mu * std::sqrt(uT2);
  double xh[D], yh[D];
  xh[0] = mu * r[0];
  yh[0] = un_hat;
  for (int a = 1; a < D; ++a) { xh[a] = r[a]; yh[a] = mu * u[a]; }
  double z0 = 0.0;
  for (int a = 0; a < D; ++a) z0 += xh[a] * xh[a] + yh[a] * yh[a];
  double z1[D - 1], nz1 = 0.0;
  for (int a = 1; a < D; ++a) {
    z1[a - 1] = 2.0 * (xh[0] * xh[a] + yh[0] * yh[a]);
    nz1 += z1[a - 1] * z1[a - 1];
  }
  nz1 = std::sqrt(nz1);
  const double s1 = std::sqrt(z0 + nz1);
  const double s2 = std::sqrt(std::max(0.0, z0 - nz1));
  phi[0] = xh[0] + yh[0] - 0.5 * (s1 + s2);
  for (int a = 1; a < D; ++a) {
    const double dir = nz1 > 0.0 ? z1[a - 1] / nz1 : 0.0;
    phi[a] = xh[a] + yh[a] - 0.5 * (s1 - s2) * dir;
  }
}

template <int D>
double fb_err(double mu, const double* r, const double* u) {
  double phi[D];
  fb_phi<D>(mu, r, u, phi);
  double e = 0.0;
  for (int a = 0; a < D; ++a) e += phi[a] * phi[a];
  // NaN never reads as converged (std::max would drop it)
  return std::isnan(e) ? kInf : e;
}

// Exact Coulomb local solve for a diagonal Delassus block diag(w_n, w_t, w_t):
// separate (r = 0) if the normal pre-velocity is non-negative, stick if the
// unconstrained tangential impulse lies in the cone, else slide on the cone
// boundary opposing the tangential pre-velocity. This is the solution of
// So-bogus's local SOC problem for this block structure (see the closed-form
// derivation in oracle/README-level comments of tests/test_oracle_solver.py).
template <int D>
void local_solve(double mu, double wn, double wt, const double* bt, double* r) {
  const double rn = -bt[0] / wn;
  if (!(rn > 0.0)) {
    for (int a = 0; a < D; ++a) r[a] = 0.0;
    return;
  }
  r[0] = rn;
  double nrm2 = 0.0;
  for (int a = 1; a < D; ++a) {
    r[a] = -bt[a] / wt;
    nrm2 += r[a] * r[a];
  }
  const double lim = mu * rn;
  const double nrm = std::sqrt(nrm2);
  if (nrm > lim) {
    const double s = (nrm > 0.0) ? lim / nrm : 0.0;
    for (int a = 1; a < D; ++a) r[a] *= s;
  }
}

// Deterministic graph colouring shared (as a specification) with the GPU:
// Jones-Plassmann rounds with hashed priorities, each selected contact taking the
// smallest colour unused by its already-coloured neighbours. Neighbours are
// contacts sharing a body (static planes are not bodies). The reference itself
// does not colour (max_threads = 0, Sobogus.cpp:305,522).
inline uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
inline uint64_t prio_key(uint32_t c) { return ((uint64_t)fmix32(c * 0x9E3779B9u + 0x632BE5ABu) << 32) | c; }

int colour_contacts(int nbodies, const std::vector<Contact>& cs, std::vector<int>& colour, int* nrounds) {
  const int C = (int)cs.size();
  std::vector<std::vector<int>> adj(nbodies);
  for (int c = 0; c < C; ++c) {
    adj[cs[c].a].push_back(c);
    if (cs[c].b >= 0) adj[cs[c].b].push_back(c);
  }
  colour.assign(C, -1);
  std::vector<int> round_of(C, -1);
  int remaining = C, round = 0, ncol = 0;
  std::vector<int> sel;
  while (remaining > 0) {
    sel.clear();
    for (int c = 0; c < C; ++c) {
      if (colour[c] >= 0) continue;
      const uint64_t kc = prio_key(c);
      bool is_max = true;
      const int bodies[2] = {cs[c].a, cs[c].b};
      for (int k = 0; k < 2 && is_max; ++k) {
        if (bodies[k] < 0) continue;
        for (int d : adj[bodies[k]])
          if (d != c && colour[d] < 0 && prio_key(d) > kc) { is_max = false; break; }
      }
      if (is_max) sel.push_back(c);
    }
    for (int c : sel) {
      uint64_t used = 0;
      const int bodies[2] = {cs[c].a, cs[c].b};
      for (int k = 0; k < 2; ++k) {
        if (bodies[k] < 0) continue;
        for (int d : adj[bodies[k]])
          if (colour[d] >= 0) used |= (uint64_t)1 << colour[d];
      }
      int col = 0;
      while (col < 64 && (used >> col) & 1) ++col;
      if (col >= 64) return -1;
      colour[c] = col;  // assigned after every selection in the round was decided
      round_of[c] = round;
      ncol = std::max(ncol, col + 1);
    }
    remaining -= (int)sel.size();
    ++round;
  }
  if (nrounds) *nrounds = round;
  return ncol;
}

// Body description for the solver: per-body inverse mass and (3D) inverse isotropic
// inertia. Sphere inertia is isotropic, I = 2/5 m r^2 (rigidbody3d/Geometry/RigidBodySphere.cpp:10-18).
struct Bodies {
  int dim;
  int n;
  const double* m;
  const double* I;  // 3D only
};

// The per-contact H blocks in the local frame, rows [n; s; t] (2D: [n; t]) over the
// body dofs [lin | ang], restating evalH of ball2d/Constraints/BallBallConstraint.cpp:170-188,
// BallStaticPlaneConstraint.cpp:126-143, rigidbody3d/Constraints/SphereSphereConstraint.cpp:229-271,
// StaticPlaneSphereConstraint.cpp:197-223. Body b enters with a minus sign
// (scisim/ConstrainedMaps/bogus/Ball2DSobogusInterface.cpp:73-74).
template <int D>
struct HBlock {
  static constexpr int DOF = (D == 2) ? 2 : 6;
  double Ha[D][DOF];
  double Hb[D][DOF];
};

template <int D>
void eval_H(const Contact& c, HBlock<D>& H) {
  const double* B = c.basis;
  for (int row = 0; row < D; ++row)
    for (int k = 0; k < HBlock<D>::DOF; ++k) { H.Ha[row][k] = 0.0; H.Hb[row][k] = 0.0; }
  for (int row = 0; row < D; ++row)
    for (int k = 0; k < D; ++k) { H.Ha[row][k] = B[row * D + k]; H.Hb[row][k] = B[row * D + k]; }
  if (D == 3) {
    for (int row = 1; row < 3; ++row) {
      double cr[3];
      cross3(c.pa, &B[row * 3], cr);
      for (int k = 0; k < 3; ++k) H.Ha[row][3 + k] = cr[k];
      cross3(c.pb, &B[row * 3], cr);
      for (int k = 0; k < 3; ++k) H.Hb[row][3 + k] = cr[k];
    }
  }
}

template <int D>
struct Solver {
  static constexpr int DOF = HBlock<D>::DOF;
  const Bodies& B;
  const std::vector<Contact>& cs;
  std::vector<HBlock<D>> H;
  std::vector<double> wn, wt, forcing, mus;
  std::vector<double> minv;  // DOF per body diag of M^-1

  Solver(const Bodies& b, const std::vector<Contact>& c, const std::vector<double>& mu, double CoR)
      : B(b), cs(c) {
    const int C = (int)cs.size();
    H.resize(C);
    wn.resize(C);
    wt.resize(C);
    forcing.resize(C);
    mus = mu;
    minv.assign((size_t)B.n * DOF, 0.0);
    for (int i = 0; i < B.n; ++i) {
      for (int k = 0; k < D; ++k) minv[(size_t)i * DOF + k] = 1.0 / B.m[i];
      if (D == 3)
        for (int k = 3; k < 6; ++k) minv[(size_t)i * DOF + k] = 1.0 / B.I[i];
    }
#pragma omp parallel for schedule(static) if (C > 4096)
    for (int ci = 0; ci < C; ++ci) {
      eval_H<D>(cs[ci], H[ci]);
      // Delassus diagonal block W_ii = H_a Ma^-1 H_a^T + H_b Mb^-1 H_b^T
      double W[D][D];
      for (int r1 = 0; r1 < D; ++r1)
        for (int r2 = 0; r2 < D; ++r2) {
          double s = 0.0;
          for (int k = 0; k < DOF; ++k) s += H[ci].Ha[r1][k] * minv[(size_t)cs[ci].a * DOF + k] * H[ci].Ha[r2][k];
          if (cs[ci].b >= 0)
            for (int k = 0; k < DOF; ++k) s += H[ci].Hb[r1][k] * minv[(size_t)cs[ci].b * DOF + k] * H[ci].Hb[r2][k];
          W[r1][r2] = s;
        }
      wn[ci] = W[0][0];
      wt[ci] = (D == 2) ? W[1][1] : 0.5 * (W[1][1] + W[2][2]);
      // Constraint::computeForcingTerm (scisim/Constraints/Constraint.cpp:18-34) with
      // static geometry (kinematic nrel = drel = 0) plus the caller's extra terms
      // (Sobogus.cpp:893-905), in the local frame: (CoR * n.v + nrel, drel...).
      forcing[ci] = CoR * cs[ci].ndotv + cs[ci].nrel;
    }
    for (int ci = 0; ci < C; ++ci)
      for (int k = 0; k < D - 1; ++k) has_drel = has_drel || cs[ci].drel[k] != 0.0;
  }
  bool has_drel = false;

  // u = H_a v_a - H_b v_b + w, in the local frame.
  void rel_vel(int ci, const double* v, double* u) const {
    const Contact& c = cs[ci];
    for (int row = 0; row < D; ++row) {
      double s = 0.0;
      for (int k = 0; k < DOF; ++k) s += H[ci].Ha[row][k] * v[(size_t)c.a * DOF + k];
      if (c.b >= 0) {
        double sb = 0.0;
        for (int k = 0; k < DOF; ++k) sb += H[ci].Hb[row][k] * v[(size_t)c.b * DOF + k];
        s -= sb;
      }
      u[row] = s;
    }
    u[0] += forcing[ci];
    if (has_drel)
      for (int k = 1; k < D; ++k) u[k] += c.drel[k - 1];
  }

  void apply(int ci, const double* dr, double* v) const {
    const Contact& c = cs[ci];
    for (int k = 0; k < DOF; ++k) {
      double s = 0.0;
      for (int row = 0; row < D; ++row) s += H[ci].Ha[row][k] * dr[row];
      v[(size_t)c.a * DOF + k] += minv[(size_t)c.a * DOF + k] * s;
    }
    if (c.b >= 0)
      for (int k = 0; k < DOF; ++k) {
        double s = 0.0;
        for (int row = 0; row < D; ++row) s += H[ci].Hb[row][k] * dr[row];
        v[(size_t)c.b * DOF + k] -= minv[(size_t)c.b * DOF + k] * s;
      }
  }

  // The max over contacts (an exact, order-independent reduction, so the thread count
  // never changes the result); the reference's So-bogus interface parallelises its
  // per-contact work with OpenMP likewise (Ball2DSobogusInterface.cpp:66-67).
  double eval(const double* r, const double* v) const {
    double e = 0.0;
    const int64_t C = (int64_t)cs.size();
#pragma omp parallel for reduction(max : e) schedule(static) if (C > 4096)
    for (int64_t ci = 0; ci < C; ++ci) {
      double u[D];
      rel_vel((int)ci, v, u);
      e = std::max(e, fb_err<D>(mus[ci], r + D * ci, u));
    }
    return e;
  }

  // Block Gauss-Seidel in primal-velocity form (equivalent in exact arithmetic to
  // So-bogus GaussSeidel on W = H M^-1 H^T, b = E^T w + H v0): Sobogus.cpp:273-373
  // (2D), :486-594 (3D); Ball2DSobogusInterface.cpp:102-148. Keeps the best
  // iterate, stops when err < tol, returns err_best and the sweep count
  // (So-bogus patch scripts_include/sobogus_patch.patch:67-70).
  double solve(const std::vector<int>& order, const double* v0, double* r, double* v, uint32_t max_iters,
               uint32_t eval_every, double tol, uint32_t* sweeps, std::vector<double>* hist) const {
    const size_t C = cs.size();
    for (size_t i = 0; i < (size_t)B.n * DOF; ++i) v[i] = v0[i];
    for (size_t ci = 0; ci < C; ++ci) apply((int)ci, r + D * ci, v);
    double err = eval(r, v);
    if (hist) hist->push_back(err);
    double err_best = err;
    std::vector<double> r_best(r, r + D * C);
    if (err < tol) {
      *sweeps = 0;
      return err;
    }
    uint32_t it;
    for (it = 1; it <= max_iters; ++it) {
      for (int ci : order) {
        double u[D], bt[D], rn[D], dr[D];
        rel_vel(ci, v, u);
        double* rc = r + D * ci;
        bt[0] = u[0] - wn[ci] * rc[0];
        for (int a = 1; a < D; ++a) bt[a] = u[a] - wt[ci] * rc[a];
        local_solve<D>(mus[ci], wn[ci], wt[ci], bt, rn);
        for (int a = 0; a < D; ++a) { dr[a] = rn[a] - rc[a]; rc[a] = rn[a]; }
        apply(ci, dr, v);
      }
      if (eval_every > 0 && it % eval_every == 0) {
        err = eval(r, v);
        if (hist) hist->push_back(err);
        if (err < err_best) {
          err_best = err;
          std::copy(r, r + D * C, r_best.begin());
        }
        if (err < tol) break;
      }
    }
    if (it > max_iters) std::copy(r_best.begin(), r_best.end(), r);
    // outgoing velocity from the final impulses, v = v0 + M^-1 H^T r
    // (Ball2DSobogusInterface.cpp:142 with f = -M v0)
    for (size_t i = 0; i < (size_t)B.n * DOF; ++i) v[i] = v0[i];
    for (size_t ci = 0; ci < C; ++ci) apply((int)ci, r + D * ci, v);
    // GSIter as So-bogus reports it (scripts_include/sobogus_patch.patch:67-70): one past the
    // limit when the limit is reached
    *sweeps = it;
    return err_best;
  }
};

// ---------------------------------------------------------------------------
// Unconstrained maps.
// ---------------------------------------------------------------------------
// SymplecticEulerMap::flow, ball2d/SymplecticEulerMap.cpp:21-32 (force = m g from
// ball2d/Forces/Ball2DGravityForce.cpp:36-46; Minv = 1/m, ball2d/Ball2DState.cpp:54-65).
// VerletMap::flow, ball2d/VerletMap.cpp:15-41. 3D: translational part of
// SplitHamMap::flow, rigidbody3d/UnconstrainedMaps/SplitHamMap.cpp:43-71,147-181
// (angular velocity of an isotropic torque-free sphere is unchanged).
template <int D>
void umap_flow(int integrator, int n, const double* m, const double* g, double dt, const double* q0,
               const double* v0, double* q1, double* v1) {
  for (int i = 0; i < n; ++i) {
    const double minv = 1.0 / m[i];
    for (int a = 0; a < D; ++a) {
      const size_t k = (size_t)D * i + a;
      const double F = m[i] * g[a];
      if (D == 2 && integrator == 0) {
        v1[k] = v0[k] + (dt * minv) * F;
        q1[k] = q0[k] + dt * v1[k];
      } else if (D == 2) {
        double v = v0[k] + ((0.5 * dt) * minv) * F;
        q1[k] = q0[k] + dt * v;
        v1[k] = v + ((0.5 * dt) * minv) * F;
      } else {
        // p = M v0; p += dt/2 F0; q1 = q0 + (dt v0 + dt^2/2 Minv F); p += dt/2 F1; v1 = p / m
        double p = m[i] * v0[k];
        p = p + (0.5 * dt) * F;
        q1[k] = q0[k] + (dt * v0[k] + (((0.5 * dt) * dt) * minv) * F);
        p = p + (0.5 * dt) * F;
        v1[k] = p / m[i];
      }
    }
  }
}

}  // namespace

// ===========================================================================
// C API (ctypes).
// ===========================================================================
extern "C" {

typedef struct orc_params {
  double dt;
  double g[3];
  double mu;
  double CoR;
  double tol;
  uint32_t max_iters;
  uint32_t eval_every;
  int32_t integrator;
  int32_t order;
  int32_t warm_start;
  int32_t staggering;   // 0 geometric, 1 symplectic Euler, 2 stabilized
  int32_t stabilize;    // symplectic Euler: min(0, gap(q0)) / dt added to the normal forcing
  int32_t pad;
  double penetration_threshold;
} orc_params;

typedef struct orc_stats {
  uint64_t npairs;
  uint32_t ncontacts;
  uint32_t ncolours;
  uint32_t nrounds;
  uint32_t sweeps;
  double error;
  int32_t succeeded;
  int32_t pad;
} orc_stats;

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Threads of the parallel parts (residual evaluation, per-contact set-up); the sweep
// itself is sequential in the reference's natural order.
void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

// Grid broad phase (reference algorithm). Returns number of pairs, or -(needed) if cap too small.
int64_t orc_grid_pairs(int dim, uint32_t n, const double* mn, const double* mx, uint32_t* pairs, uint64_t cap) {
  std::vector<std::pair<uint32_t, uint32_t>> out;
  if (dim == 2)
    grid_pairs<2>(n, mn, mx, out);
  else
    grid_pairs<3>(n, mn, mx, out);
  if (out.size() > cap) return -(int64_t)out.size();
  for (size_t k = 0; k < out.size(); ++k) { pairs[2 * k] = out[k].first; pairs[2 * k + 1] = out[k].second; }
  return (int64_t)out.size();
}

int64_t orc_all_pairs(int dim, uint32_t n, const double* mn, const double* mx, uint32_t* pairs, uint64_t cap) {
  std::vector<std::pair<uint32_t, uint32_t>> out;
  if (dim == 2)
    all_pairs<2>(n, mn, mx, out);
  else
    all_pairs<3>(n, mn, mx, out);
  if (out.size() > cap) return -(int64_t)out.size();
  for (size_t k = 0; k < out.size(); ++k) { pairs[2 * k] = out[k].first; pairs[2 * k + 1] = out[k].second; }
  return (int64_t)out.size();
}

void orc_ccd_coeffs(const double* q0a, const double* q1a, double ra, const double* q0b, const double* q1b, double rb,
                    double* c) {
  ccd_coeffs(q0a, q1a, ra, q0b, q1b, rb, c);
}

int orc_ccd_happens(const double* c, double* t) { return ccd_happens(c, t) ? 1 : 0; }

void transport_impulse(int D, const double* n_old, const double* r_old, const double* B, double* f);

// Transported cached impulse in the local frame of basis B (column-major [n s t]).
void orc_transport(int dim, const double* n_old, const double* r_old, const double* B, double* rloc) {
  double f[3];
  transport_impulse(dim, n_old, r_old, B, f);
  for (int k = 0; k < dim; ++k) {
    double acc = B[k * dim] * f[0];
    for (int a = 1; a < dim; ++a) acc = acc + B[k * dim + a] * f[a];
    rloc[k] = acc;
  }
}

double orc_fb_err(int dim, double mu, const double* r, const double* u) {
  return dim == 2 ? fb_err<2>(mu, r, u) : fb_err<3>(mu, r, u);
}

void orc_fb_phi(int dim, double mu, const double* r, const double* u, double* phi) {
  if (dim == 2)
    fb_phi<2>(mu, r, u, phi);
  else
    fb_phi<3>(mu, r, u, phi);
}

void orc_local_solve(int dim, double mu, double wn, double wt, const double* bt, double* r) {
  if (dim == 2)
    local_solve<2>(mu, wn, wt, bt, r);
  else
    local_solve<3>(mu, wn, wt, bt, r);
}

void orc_umap_flow(int dim, int integrator, int n, const double* m, const double* g, double dt, const double* q0,
                   const double* v0, double* q1, double* v1) {
  if (dim == 2)
    umap_flow<2>(integrator, n, m, g, dt, q0, v0, q1, v1);
  else
    umap_flow<3>(integrator, n, m, g, dt, q0, v0, q1, v1);
}

// Active set + bases at (q0, q1, v0): Ball2DSim::computeActiveSet (ball2d/Ball2DSim.cpp:151-173,
// 553-608, 747-764) and RigidBody3DSim::computeActiveSet (rigidbody3d/RigidBody3DSim.cpp:250-262,
// 793-819, 1057-1260, 1414-1502) with computeContactBases (Ball2DSim.cpp:188-201,
// RigidBody3DSim.cpp:277-290).
static void active_set(int dim, int n, const double* q0, const double* q1, const double* v0, const double* w0,
                       const double* r, int nplanes, const double* px, const double* pn, std::vector<Contact>& cs,
                       uint64_t* npairs) {
  cs.clear();
  const int D = dim;
  std::vector<double> mn((size_t)D * n), mx((size_t)D * n);
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < D; ++a) {
      const size_t k = (size_t)D * i + a;
      if (D == 2) {
        mn[k] = std::min(q1[k], q0[k]) - r[i];
        mx[k] = std::max(q1[k], q0[k]) + r[i];
      } else {
        mn[k] = q1[k] - r[i];
        mx[k] = q1[k] + r[i];
      }
    }
  std::vector<std::pair<uint32_t, uint32_t>> pairs;
  if (D == 2)
    grid_pairs<2>(n, mn.data(), mx.data(), pairs);
  else
    grid_pairs<3>(n, mn.data(), mx.data(), pairs);
  if (npairs) *npairs = pairs.size();
  for (const auto& pr : pairs) {
    const int i = pr.first, j = pr.second;
    Contact c;
    std::memset(&c, 0, sizeof(c));
    c.a = i;
    c.b = j;
    c.plane = -1;
    if (D == 2) {
      double cc[3], t;
      ccd_coeffs(q0 + 2 * i, q1 + 2 * i, r[i], q0 + 2 * j, q1 + 2 * j, r[j], cc);
      if (!ccd_happens(cc, &t)) continue;
      double nn[2] = {q0[2 * i] - q0[2 * j], q0[2 * i + 1] - q0[2 * j + 1]};
      normalize<2>(nn);
      c.basis[0] = nn[0];
      c.basis[1] = nn[1];
      c.basis[2] = -nn[1];
      c.basis[3] = nn[0];
      c.ndotv = nn[0] * (v0[2 * i] - v0[2 * j]) + nn[1] * (v0[2 * i + 1] - v0[2 * j + 1]);
    } else {
      const double dx = q1[3 * i] - q1[3 * j], dy = q1[3 * i + 1] - q1[3 * j + 1], dz = q1[3 * i + 2] - q1[3 * j + 2];
      const double rs = r[i] + r[j];
      if (!((dx * dx + dy * dy + dz * dz) <= rs * rs)) continue;
      double nn[3] = {q0[3 * i] - q0[3 * j], q0[3 * i + 1] - q0[3 * j + 1], q0[3 * i + 2] - q0[3 * j + 2]};
      normalize<3>(nn);
      const double f = r[i] / (r[i] + r[j]);
      double p[3];
      for (int a = 0; a < 3; ++a) p[a] = q0[3 * i + a] + f * (q0[3 * j + a] - q0[3 * i + a]);
      for (int a = 0; a < 3; ++a) { c.pa[a] = p[a] - q0[3 * i + a]; c.pb[a] = p[a] - q0[3 * j + a]; }
      // SphereSphereConstraint::computeContactBasis (SphereSphereConstraint.cpp:273-300)
      double ca[3], cb[3], s[3];
      cross3(w0 + 3 * i, c.pa, ca);
      cross3(w0 + 3 * j, c.pb, cb);
      for (int a = 0; a < 3; ++a) s[a] = ((v0[3 * i + a] + ca[a]) - v0[3 * j + a]) - cb[a];
      double nxs[3];
      cross3(nn, s, nxs);
      if (dot3(nxs, nxs) < 1.0e-9) {
        orthogonal_vector(nn, s);
      } else {
        const double d = dot3(s, nn);
        for (int a = 0; a < 3; ++a) s[a] = s[a] - d * nn[a];
        normalize<3>(s);
      }
      for (int a = 0; a < 3; ++a) s[a] *= -1.0;
      double t[3];
      cross3(nn, s, t);
      normalize<3>(t);
      for (int a = 0; a < 3; ++a) { c.basis[a] = nn[a]; c.basis[3 + a] = s[a]; c.basis[6 + a] = t[a]; }
      c.ndotv = nn[0] * (v0[3 * i] - v0[3 * j]) + nn[1] * (v0[3 * i + 1] - v0[3 * j + 1]) +
                nn[2] * (v0[3 * i + 2] - v0[3 * j + 2]);
    }
    cs.push_back(c);
  }
  // planes, plane-major then body order
  for (int p = 0; p < nplanes; ++p) {
    const double* x = px + (size_t)D * p;
    double nn[3] = {pn[(size_t)D * p], pn[(size_t)D * p + 1], D == 3 ? pn[(size_t)D * p + 2] : 0.0};
    if (D == 2) normalize<2>(nn); else normalize<3>(nn);
    for (int i = 0; i < n; ++i) {
      double d;
      if (D == 2)
        d = nn[0] * (q1[2 * i] - x[0]) + nn[1] * (q1[2 * i + 1] - x[1]);
      else
        d = nn[0] * (q1[3 * i] - x[0]) + nn[1] * (q1[3 * i + 1] - x[1]) + nn[2] * (q1[3 * i + 2] - x[2]);
      if (!(d <= r[i])) continue;
      Contact c;
      std::memset(&c, 0, sizeof(c));
      c.a = i;
      c.b = -1;
      c.plane = p;
      if (D == 2) {
        // StaticPlane: t = (-n_y, n_x), ball2d/StaticGeometry/StaticPlane.cpp:10-20
        c.basis[0] = nn[0];
        c.basis[1] = nn[1];
        c.basis[2] = -nn[1];
        c.basis[3] = nn[0];
        c.ndotv = nn[0] * v0[2 * i] + nn[1] * v0[2 * i + 1];
      } else {
        // StaticPlaneSphereConstraint, StaticPlaneSphereConstraint.cpp:255-283; arm -r n
        for (int a = 0; a < 3; ++a) c.pa[a] = (-r[i]) * nn[a];
        double ca[3], s[3];
        cross3(w0 + 3 * i, c.pa, ca);
        for (int a = 0; a < 3; ++a) s[a] = (v0[3 * i + a] + ca[a]) - 0.0;
        double nxs[3];
        cross3(nn, s, nxs);
        if (dot3(nxs, nxs) < 1.0e-9) {
          orthogonal_vector(nn, s);
        } else {
          const double dd = dot3(s, nn);
          for (int a = 0; a < 3; ++a) s[a] = s[a] - dd * nn[a];
          normalize<3>(s);
        }
        for (int a = 0; a < 3; ++a) s[a] *= -1.0;
        double t[3];
        cross3(nn, s, t);
        normalize<3>(t);
        for (int a = 0; a < 3; ++a) { c.basis[a] = nn[a]; c.basis[3 + a] = s[a]; c.basis[6 + a] = t[a]; }
        c.ndotv = nn[0] * v0[3 * i] + nn[1] * v0[3 * i + 1] + nn[2] * v0[3 * i + 2];
      }
      cs.push_back(c);
    }
  }
}

// Previous world impulse r_old rotated by the minimal rotation taking the old normal
// n_old to the current normal B[0..D) (Eigen's Quaternion::FromTwoVectors and
// quaternion-vector product, as SymplecticEulerImpactFrictionMap.cpp:96-115 uses them);
// an all-zero cache entry gives zero. 2D vectors are rotated in the xy-plane.
void transport_impulse(int D, const double* n_old, const double* r_old, const double* B, double* f) {
  double a[3] = {0, 0, 0}, b[3] = {0, 0, 0}, r[3] = {0, 0, 0};
  bool zero = true;
  for (int k = 0; k < D; ++k) {
    a[k] = n_old[k];
    b[k] = B[k];
    r[k] = r_old[k];
    zero = zero && n_old[k] == 0.0 && r_old[k] == 0.0;
  }
  f[0] = f[1] = f[2] = 0.0;
  if (zero) return;
  const double na = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  const double nb = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  for (int k = 0; k < 3; ++k) {
    a[k] = a[k] / na;
    b[k] = b[k] / nb;
  }
  const double c = b[0] * a[0] + b[1] * a[1] + b[2] * a[2];
  double v[3], w;
  if (c < -1.0 + 1e-12) {
    // opposite normals: half turn about an axis orthogonal to the old normal
    double ax[3];
    orthogonal_vector(a, ax);
    const double w2 = (1.0 + std::max(c, -1.0)) * 0.5;
    w = std::sqrt(w2);
    const double sv = std::sqrt(1.0 - w2);
    for (int k = 0; k < 3; ++k) v[k] = ax[k] * sv;
  } else {
    const double ax[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double s = std::sqrt((1.0 + c) * 2.0);
    const double invs = 1.0 / s;
    for (int k = 0; k < 3; ++k) v[k] = ax[k] * invs;
    w = s * 0.5;
  }
  double uv[3] = {v[1] * r[2] - v[2] * r[1], v[2] * r[0] - v[0] * r[2], v[0] * r[1] - v[1] * r[0]};
  for (int k = 0; k < 3; ++k) uv[k] = uv[k] + uv[k];
  const double vu[3] = {v[1] * uv[2] - v[2] * uv[1], v[2] * uv[0] - v[0] * uv[2], v[0] * uv[1] - v[1] * uv[0]};
  for (int k = 0; k < 3; ++k) f[k] = (r[k] + w * uv[k]) + vu[k];
}

// Persistent oracle context: holds the 2D warm-start cache between steps
// (ball2d/ConstraintCache.cpp:20-123; GeometricImpactFrictionMap.cpp:174-264) and the
// last step's contacts and impulses for inspection.
typedef struct orc_ctx {
  int dim;
  // (i,j) / (plane,i) -> (alpha, beta, 0...) for NORMAL / NORMAL_AND_FRICTION, or
  // (old normal, world impulse) for the transported cache
  std::map<std::pair<int, int>, std::array<double, 6>> cache_bb;
  std::map<std::pair<int, int>, std::array<double, 6>> cache_pl;
  // layout of the cached entries: 0 (alpha, beta...) or 1 (old normal, world impulse); a
  // step that reads the other layout finds no entry (the reference's map has one mode
  // per simulation, so this never arises there)
  int cache_layout = 0;
  std::vector<Contact> cs;
  std::vector<int> colour;
  std::vector<double> alpha, beta, rworld;
  std::vector<double> v2;
  std::vector<double> hist;
} orc_ctx;

orc_ctx* orc_create(int dim) {
  orc_ctx* c = new orc_ctx;
  c->dim = dim;
  return c;
}
void orc_destroy(orc_ctx* c) { delete c; }
// A copy of a context (impulse cache and last step), for stepping the same state twice.
orc_ctx* orc_clone(const orc_ctx* c) { return new orc_ctx(*c); }

// Gap function at q (evaluateGapFunction): ball2d/Constraints/BallBallConstraint.cpp:76-86,
// BallStaticPlaneConstraint.cpp:58-63, rigidbody3d/Constraints/SphereSphereConstraint.cpp:28-31,
// StaticPlaneSphereConstraint.cpp:31-35.
static double gap(int D, const Contact& c, const double* q, const double* r, const double* px, const double* pn) {
  if (c.b >= 0) {
    double d2 = 0.0;
    for (int a = 0; a < D; ++a) {
      const double d = q[(size_t)D * c.a + a] - q[(size_t)D * c.b + a];
      d2 = (a == 0) ? d * d : d2 + d * d;
    }
    // 2D subtracts the radius sum, 3D each radius in turn
    return D == 2 ? std::sqrt(d2) - (r[c.a] + r[c.b]) : (std::sqrt(d2) - r[c.a]) - r[c.b];
  }
  double nn[3] = {pn[(size_t)D * c.plane], pn[(size_t)D * c.plane + 1], D == 3 ? pn[(size_t)D * c.plane + 2] : 0.0};
  if (D == 2) normalize<2>(nn); else normalize<3>(nn);
  double s = 0.0;
  for (int a = 0; a < D; ++a) {
    const double d = q[(size_t)D * c.a + a] - px[(size_t)D * c.plane + a];
    s = (a == 0) ? nn[0] * d : s + nn[a] * d;
  }
  return s - r[c.a];
}

// One step of an impact-friction map for circles/spheres, q, v (D*n) and w (3D: 3n
// angular velocity) updated in place:
//   staggering 0: GeometricImpactFrictionMap::flow
//     (scisim/ConstrainedMaps/GeometricImpactFrictionMap.cpp:267-442);
//   staggering 1: SymplecticEulerImpactFrictionMap::flow
//     (scisim/ConstrainedMaps/SymplecticEulerImpactFrictionMap.cpp:201-407): force increment
//     vdelta = dt M^-1 F(q0), predictor q0 + dt (v0 + vdelta), bases at (q0, v0), forcing
//     nrel = N^T vdelta (+ min(0, g(q0)) / dt with stabilize), drel = D^T vdelta, solve on v0,
//     v1 = v2 + vdelta, q1 = q0 + dt v1;
//   staggering 2: StabilizedImpactFrictionMap::flow
//     (scisim/ConstrainedMaps/StabilizedImpactFrictionMap.cpp:48-183): predictor (q1, v1)
//     = U(q0, v0), bases at (q0, v1), cold solve on v1 (no cache), v1 = v2, q1 = q0 + dt v1.
int orc_step(orc_ctx* ctx, int n, double* q, double* v, double* w, const double* r, const double* m,
             const double* I, int nplanes, const double* px, const double* pn, const orc_params* P,
             orc_stats* st) {
  const int D = ctx->dim;
  const int stag = P->staggering;
  std::memset(st, 0, sizeof(*st));
  // "Caching in 3 space not yet supported." (GeometricImpactFrictionMap.cpp:258-262); the
  // 3D cache of the symplectic Euler map is the transported one (warm_start 3)
  if (P->warm_start == 1 && D == 3) return -3;
  if (stag < 0 || stag > 2) return -4;
  // rigidbody3dutils/RigidBody3DSceneParser.cpp:1881-1888: the symplectic Euler staggering
  // needs CoR = 0; the stabilized map has no impulse cache (StabilizedImpactFrictionMap.cpp:107-109)
  if (stag == 1 && P->CoR != 0.0) return -5;
  if (stag == 2 && P->warm_start != 0) return -6;
  std::vector<double> q0(q, q + (size_t)D * n), v0(v, v + (size_t)D * n);
  std::vector<double> w0;
  if (D == 3) w0.assign(w, w + (size_t)3 * n);
  std::vector<double> q1((size_t)D * n), v1((size_t)D * n), vd;
  if (stag == 1) {
    // vdelta = dt * (M^-1 F), F = m g (SymplecticEulerImpactFrictionMap.cpp:222-229); the
    // predictor is linearInertialConfigurationUpdate(q0, v0 + vdelta) (:232)
    vd.resize((size_t)D * n);
    for (int i = 0; i < n; ++i)
      for (int a = 0; a < D; ++a) {
        const size_t k = (size_t)D * i + a;
        vd[k] = P->dt * ((1.0 / m[i]) * (m[i] * P->g[a]));
        v1[k] = v0[k] + vd[k];
        q1[k] = q0[k] + P->dt * v1[k];
      }
  } else if (D == 2) {
    // predictor (GeometricImpactFrictionMap.cpp:279, StabilizedImpactFrictionMap.cpp:60)
    umap_flow<2>(P->integrator, n, m, P->g, P->dt, q0.data(), v0.data(), q1.data(), v1.data());
  } else {
    umap_flow<3>(P->integrator, n, m, P->g, P->dt, q0.data(), v0.data(), q1.data(), v1.data());
  }
  // active set at (q0, q1) and bases at (q0, v0) (geometric :283,:304; symplectic Euler
  // :235,:279), or at (q0, v1) (stabilized :63,:85)
  const std::vector<double>& vb = (stag == 2) ? v1 : v0;
  std::vector<Contact>& cs = ctx->cs;
  active_set(D, n, q0.data(), q1.data(), vb.data(), D == 3 ? w0.data() : nullptr, r, nplanes, px, pn, cs,
             &st->npairs);
  const int C = (int)cs.size();
  st->ncontacts = C;
  ctx->alpha.assign(C, 0.0);
  ctx->beta.assign((size_t)(D - 1) * C, 0.0);
  ctx->rworld.assign((size_t)D * C, 0.0);
  ctx->colour.assign(C, 0);
  ctx->hist.clear();
  if (C == 0) {
    // empty active set: keep the predictor; the geometric and symplectic Euler maps clear
    // their cache (:286-298; SymplecticEulerImpactFrictionMap.cpp:239-253), the stabilized
    // map has none (:66-78)
    if (stag != 2) {
      ctx->cache_bb.clear();
      ctx->cache_pl.clear();
    }
    std::copy(q1.begin(), q1.end(), q);
    std::copy(v1.begin(), v1.end(), v);
    ctx->v2 = (stag == 0) ? v0 : v1;
    st->succeeded = 1;
    return 0;
  }
  if (stag == 1) {
    // nrel = N^T vdelta, drel = D^T vdelta (:296-307): sparse column dot products over the
    // rows of body a then body b (b > a), +n / -n, +t / -t; the angular rows of vdelta are
    // zero (no torque)
    for (Contact& c : cs) {
      for (int row = 0; row < D; ++row) {
        const double* e = c.basis + (size_t)D * row;  // n, then the tangents
        double acc = 0.0;
        for (int a = 0; a < D; ++a) acc = acc + e[a] * vd[(size_t)D * c.a + a];
        if (c.b >= 0)
          for (int a = 0; a < D; ++a) acc = acc + (-e[a]) * vd[(size_t)D * c.b + a];
        if (row == 0)
          c.nrel = acc;
        else
          c.drel[row - 1] = acc;
      }
      if (P->stabilize) {
        // g0 = min(0, g(q0)), zero below the penetration threshold, nrel += g0 / dt (:309-326)
        double g0 = std::min(0.0, gap(D, c, q0.data(), r, px, pn));
        if (std::fabs(g0) < P->penetration_threshold) g0 = 0.0;
        c.nrel = c.nrel + g0 / P->dt;
      }
    }
  }
  std::vector<double> mu(C, P->mu);
  // warm start (initializeImpulses, :174-220): the cached (alpha, beta) of the same
  // constraint, zero on a miss (ball2d/ConstraintCache.cpp:71-123); NORMAL restores alpha
  // only. NORMAL_AND_FRICTION exists in 2D only (the reference exits in 3D, :212-216).
  // The transported cache (warm_start 3) keeps the old normal and world impulse and
  // rotates the impulse with the normal (SymplecticEulerImpactFrictionMap.cpp:87-127).
  std::vector<double> rloc((size_t)D * C, 0.0);
  const bool caching = (P->warm_start == 1 && D == 2) || P->warm_start == 2 || P->warm_start == 3;
  const int layout = P->warm_start == 3 ? 1 : 0;
  if (caching && layout == ctx->cache_layout) {
    for (int ci = 0; ci < C; ++ci) {
      const auto key = cs[ci].b >= 0 ? std::make_pair(cs[ci].a, cs[ci].b) : std::make_pair(cs[ci].plane, cs[ci].a);
      const auto& cache = cs[ci].b >= 0 ? ctx->cache_bb : ctx->cache_pl;
      auto it = cache.find(key);
      if (it == cache.end()) continue;
      const double* B = cs[ci].basis;
      const double* cv = it->second.data();
      double rw[3] = {0, 0, 0};
      if (P->warm_start == 3) {
        transport_impulse(D, cv, cv + 3, B, rw);
      } else {
        // r = alpha n + beta t (Sobogus.cpp:301; beta = 0 for NORMAL)
        for (int a = 0; a < D; ++a) rw[a] = cv[0] * B[a];
        if (P->warm_start == 1)
          for (int a = 0; a < D; ++a) rw[a] = rw[a] + cv[1] * B[D + a];
      }
      // local coordinates E^T r
      for (int k = 0; k < D; ++k) {
        double acc = B[k * D] * rw[0];
        for (int a = 1; a < D; ++a) acc = acc + B[k * D + a] * rw[a];
        rloc[(size_t)D * ci + k] = acc;
      }
    }
  }
  // a step without caching leaves the cache untouched (the map only writes it when caching)
  if (caching) {
    ctx->cache_bb.clear();
    ctx->cache_pl.clear();
  }
  // order
  std::vector<int> order(C);
  for (int i = 0; i < C; ++i) order[i] = i;
  int nrounds = 0;
  const int ncol = colour_contacts(n, cs, ctx->colour, &nrounds);
  if (ncol < 0) return -2;
  st->ncolours = ncol;
  st->nrounds = nrounds;
  if (P->order == 1)
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return ctx->colour[x] < ctx->colour[y]; });
  // velocities in solver layout [lin | ang] per body
  const int DOF = (D == 2) ? 2 : 6;
  std::vector<double> vs((size_t)DOF * n), vout((size_t)DOF * n);
  // the solve starts from v0, or from the predictor velocity in the stabilized map (:117)
  for (int i = 0; i < n; ++i) {
    for (int a = 0; a < D; ++a) vs[(size_t)DOF * i + a] = vb[(size_t)D * i + a];
    if (D == 3)
      for (int a = 0; a < 3; ++a) vs[(size_t)DOF * i + 3 + a] = w0[(size_t)3 * i + a];
  }
  Bodies B{D, n, m, I};
  double err;
  uint32_t sweeps = 0;
  if (D == 2) {
    Solver<2> S(B, cs, mu, P->CoR);
    err = S.solve(order, vs.data(), rloc.data(), vout.data(), P->max_iters, P->eval_every, P->tol, &sweeps, &ctx->hist);
  } else {
    Solver<3> S(B, cs, mu, P->CoR);
    err = S.solve(order, vs.data(), rloc.data(), vout.data(), P->max_iters, P->eval_every, P->tol, &sweeps, &ctx->hist);
  }
  st->sweeps = sweeps;
  st->error = err;
  st->succeeded = err < P->tol ? 1 : 0;
  // world impulses r = E r_loc; alpha = r.n, beta = r.t (Sobogus.cpp:312-332, :528-541)
  for (int ci = 0; ci < C; ++ci) {
    const double* Bm = cs[ci].basis;
    double rw[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a)
      for (int col = 0; col < D; ++col) rw[a] += Bm[col * D + a] * rloc[(size_t)D * ci + col];
    for (int a = 0; a < D; ++a) ctx->rworld[(size_t)D * ci + a] = rw[a];
    double al = 0.0;
    for (int a = 0; a < D; ++a) al += rw[a] * Bm[a];
    ctx->alpha[ci] = al;
    for (int k = 1; k < D; ++k) {
      double be = 0.0;
      for (int a = 0; a < D; ++a) be += rw[a] * Bm[k * D + a];
      ctx->beta[(size_t)(D - 1) * ci + (k - 1)] = (P->mu == 0.0) ? 0.0 : be;
    }
  }
  // cache (cacheImpulses, :222-264): alpha (NORMAL), (alpha, beta) (NORMAL_AND_FRICTION),
  // or the normal and world impulse alpha n + sum_k beta_k t_k (transported)
  if (caching) ctx->cache_layout = layout;
  if (caching)
    for (int ci = 0; ci < C; ++ci) {
      std::array<double, 6> val{};
      if (P->warm_start == 3) {
        const double* B = cs[ci].basis;
        for (int a = 0; a < D; ++a) {
          val[a] = B[a];
          double w = ctx->alpha[ci] * B[a];
          for (int k = 1; k < D; ++k) w = w + ctx->beta[(size_t)(D - 1) * ci + k - 1] * B[k * D + a];
          val[3 + a] = w;
        }
      } else {
        val[0] = ctx->alpha[ci];
        val[1] = P->warm_start == 1 ? ctx->beta[ci] : 0.0;
      }
      if (cs[ci].b >= 0)
        ctx->cache_bb[std::make_pair(cs[ci].a, cs[ci].b)] = val;
      else
        ctx->cache_pl[std::make_pair(cs[ci].plane, cs[ci].a)] = val;
    }
  // v2 back to the reference layout; final flow from q0 with v2 (:441)
  std::vector<double> v2((size_t)D * n);
  for (int i = 0; i < n; ++i) {
    for (int a = 0; a < D; ++a) v2[(size_t)D * i + a] = vout[(size_t)DOF * i + a];
    if (D == 3)
      for (int a = 0; a < 3; ++a) w[(size_t)3 * i + a] = vout[(size_t)DOF * i + 3 + a];
  }
  if (stag == 0) {
    ctx->v2 = v2;
    if (D == 2)
      umap_flow<2>(P->integrator, n, m, P->g, P->dt, q0.data(), v2.data(), q, v);
    else
      umap_flow<3>(P->integrator, n, m, P->g, P->dt, q0.data(), v2.data(), q, v);
    return 0;
  }
  // symplectic Euler: v1 = v2 + vdelta (:368); both: q1 = q0 + dt v1
  // (linearInertialConfigurationUpdate, ball2d/Ball2DSim.cpp:92-98,
  // rigidbody3d/UnconstrainedMaps/IntegrationTools.cpp:36-51)
  for (size_t k = 0; k < (size_t)D * n; ++k) {
    const double vk = (stag == 1) ? v2[k] + vd[k] : v2[k];
    v[k] = vk;
    q[k] = q0[k] + P->dt * vk;
  }
  ctx->v2.assign(v, v + (size_t)D * n);
  return 0;
}

// nsteps consecutive steps (the CLI step loop, ball2dcli/ball2d_cli.cpp:441-476);
// st receives the last step's stats, max_sweeps the largest per-step sweep count.
int orc_run(orc_ctx* ctx, int n, double* q, double* v, double* w, const double* r, const double* m, const double* I,
            int nplanes, const double* px, const double* pn, const orc_params* P, int nsteps, orc_stats* st,
            uint32_t* max_sweeps, uint32_t* failures) {
  uint32_t ms = 0, nf = 0;
  for (int k = 0; k < nsteps; ++k) {
    const int rc = orc_step(ctx, n, q, v, w, r, m, I, nplanes, px, pn, P, st);
    if (rc != 0) return rc;
    ms = std::max(ms, st->sweeps);
    if (!st->succeeded) ++nf;
  }
  if (max_sweeps) *max_sweeps = ms;
  if (failures) *failures = nf;
  return 0;
}

// Last step's contacts, in the reference's active-set order.
int orc_get_contacts(const orc_ctx* ctx, uint32_t cap, int32_t* a, int32_t* b, int32_t* plane, double* basis,
                     double* alpha, double* beta, double* rworld, int32_t* colour) {
  const int D = ctx->dim;
  const uint32_t C = (uint32_t)ctx->cs.size();
  if (C > cap) return -(int)C;
  for (uint32_t ci = 0; ci < C; ++ci) {
    if (a) a[ci] = ctx->cs[ci].a;
    if (b) b[ci] = ctx->cs[ci].b;
    if (plane) plane[ci] = ctx->cs[ci].plane;
    if (basis)
      for (int k = 0; k < D * D; ++k) basis[(size_t)D * D * ci + k] = ctx->cs[ci].basis[k];
    if (alpha) alpha[ci] = ctx->alpha[ci];
    if (beta)
      for (int k = 0; k < D - 1; ++k) beta[(size_t)(D - 1) * ci + k] = ctx->beta[(size_t)(D - 1) * ci + k];
    if (rworld)
      for (int k = 0; k < D; ++k) rworld[(size_t)D * ci + k] = ctx->rworld[(size_t)D * ci + k];
    if (colour) colour[ci] = ctx->colour[ci];
  }
  return (int)C;
}

int orc_get_v2(const orc_ctx* ctx, double* v2, uint64_t len) {
  if (ctx->v2.size() > len) return -1;
  std::copy(ctx->v2.begin(), ctx->v2.end(), v2);
  return (int)ctx->v2.size();
}

int orc_get_hist(const orc_ctx* ctx, double* h, uint32_t cap) {
  const uint32_t k = std::min<uint32_t>(cap, (uint32_t)ctx->hist.size());
  for (uint32_t i = 0; i < k; ++i) h[i] = ctx->hist[i];
  return (int)ctx->hist.size();
}

// Stand-alone friction solve on an explicit contact set (FrictionSolver::solve shape,
// scisim/ConstrainedMaps/FrictionSolver.h:28). bases: D*D per contact (column-major
// [n s t]); arms (3D): 3 per contact for each body; bodies a/b (b=-1: static);
// v0: solver layout [lin | ang] per body. r: in/out local impulses (warm start in).
double orc_friction_solve(int dim, int n, const double* m, const double* I, int C, const int32_t* a, const int32_t* b,
                          const double* basis, const double* arma, const double* armb, const double* mu,
                          const double* ndotv, double CoR, const double* v0, double* r, double* vout,
                          uint32_t max_iters, uint32_t eval_every, double tol, int order_mode, uint32_t* sweeps) {
  const int D = dim;
  std::vector<Contact> cs(C);
  for (int ci = 0; ci < C; ++ci) {
    std::memset(&cs[ci], 0, sizeof(Contact));
    cs[ci].a = a[ci];
    cs[ci].b = b[ci];
    cs[ci].plane = b[ci] < 0 ? 0 : -1;
    for (int k = 0; k < D * D; ++k) cs[ci].basis[k] = basis[(size_t)D * D * ci + k];
    if (D == 3)
      for (int k = 0; k < 3; ++k) {
        cs[ci].pa[k] = arma[3 * ci + k];
        cs[ci].pb[k] = armb[3 * ci + k];
      }
    cs[ci].ndotv = ndotv ? ndotv[ci] : 0.0;
  }
  std::vector<double> muv(mu, mu + C);
  std::vector<int> order(C);
  for (int i = 0; i < C; ++i) order[i] = i;
  if (order_mode == 1) {
    std::vector<int> col;
    if (colour_contacts(n, cs, col, nullptr) < 0) return -1.0;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return col[x] < col[y]; });
  }
  Bodies B{D, n, m, I};
  if (D == 2) {
    Solver<2> S(B, cs, muv, CoR);
    return S.solve(order, v0, r, vout, max_iters, eval_every, tol, sweeps, nullptr);
  }
  Solver<3> S(B, cs, muv, CoR);
  return S.solve(order, v0, r, vout, max_iters, eval_every, tol, sweeps, nullptr);
}

// Colouring only (for bit-exact colour parity with the GPU).
int orc_colour(int nbodies, int C, const int32_t* a, const int32_t* b, int32_t* colour, int32_t* nrounds) {
  std::vector<Contact> cs(C);
  for (int ci = 0; ci < C; ++ci) {
    std::memset(&cs[ci], 0, sizeof(Contact));
    cs[ci].a = a[ci];
    cs[ci].b = b[ci];
  }
  std::vector<int> col;
  int nr = 0;
  const int nc = colour_contacts(nbodies, cs, col, &nr);
  for (int ci = 0; ci < C; ++ci) colour[ci] = col[ci];
  if (nrounds) *nrounds = nr;
  return nc;
}

}  // extern "C"
