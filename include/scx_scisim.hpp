// scx_scisim.hpp -- C++ host interface over the scx C ABI that mirrors the plug-in
// points of the reference's per-step contact path, so that host code written against
// them (SpatialGridDetector, UnconstrainedMap, FrictionSolver/Sobogus,
// GeometricImpactFrictionMap) can switch to the MI355X pipeline with the same names,
// argument meanings and error behaviour. Eigen types are replaced by std::vector<double>
// in the reference's layouts (2D q = [x0 y0 x1 y1 ...]; 3D q, v = the first 3n entries
// of the reference's [centres | rotations] / [linear | angular] vectors).
//
// Errors: the reference prints and continues on non-convergence
// (GeometricImpactFrictionMap.cpp:346-350) and exits on impossible paths; this layer
// reports non-convergence through solve_succeeded / error like the reference and throws
// scx_scisim::Error for failures of the C ABI (invalid input, device errors).
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "scx.h"

namespace scx_scisim {

class Error : public std::runtime_error {
 public:
  Error(const std::string& what, int code) : std::runtime_error(what), m_code(code) {}
  int code() const { return m_code; }

 private:
  int m_code;
};

// One scx context (device + stream) for circles (dim 2) or spheres (dim 3).
class Device {
 public:
  Device(int dim, uint32_t max_bodies, uint32_t max_planes = 8, int device = 0) : m_dim(dim) {
    scx_config cfg{};
    cfg.dim = dim;
    cfg.device = device;
    cfg.max_bodies = max_bodies;
    cfg.max_planes = max_planes;
    scx_ctx* c = nullptr;
    const int rc = scx_create(&cfg, &c);
    if (rc != SCX_OK) {
      const std::string msg = c ? scx_last_error(c) : "invalid configuration";
      if (c) scx_destroy(c);
      throw Error("scx_create: " + msg, rc);
    }
    m_ctx.reset(c);
  }
  scx_ctx* get() const { return m_ctx.get(); }
  int dim() const { return m_dim; }
  void check(int rc, const char* what) const {
    if (rc != SCX_OK) throw Error(std::string(what) + ": " + scx_last_error(m_ctx.get()), rc);
  }

 private:
  struct Deleter {
    void operator()(scx_ctx* c) const { scx_destroy(c); }
  };
  std::unique_ptr<scx_ctx, Deleter> m_ctx;
  int m_dim;
};

// AABB (ball2d/SpatialGridDetector.h:14-35, rigidbody3d/SpatialGridDetector.h:12-33).
template <int D>
class AABB final {
 public:
  AABB() = default;
  AABB(const std::array<double, D>& mn, const std::array<double, D>& mx) : m_min(mn), m_max(mx) {}
  // no separating axis with a strict max < min (ball2d/SpatialGridDetector.cpp:15-26)
  bool overlaps(const AABB& other) const {
    for (int a = 0; a < D; ++a)
      if (m_max[a] < other.m_min[a] || other.m_max[a] < m_min[a]) return false;
    return true;
  }
  std::array<double, D>& min() { return m_min; }
  std::array<double, D>& max() { return m_max; }
  const std::array<double, D>& min() const { return m_min; }
  const std::array<double, D>& max() const { return m_max; }

 private:
  std::array<double, D> m_min{};
  std::array<double, D> m_max{};
};

namespace SpatialGridDetector {

// getPotentialOverlaps (ball2d/SpatialGridDetector.h:39): every overlapping pair, i < j.
template <int D>
void getPotentialOverlaps(const Device& dev, const std::vector<AABB<D>>& aabbs,
                          std::set<std::pair<unsigned, unsigned>>& overlaps) {
  if (dev.dim() != D) throw Error("SpatialGridDetector: device dimension mismatch", SCX_E_ARG);
  const uint32_t n = static_cast<uint32_t>(aabbs.size());
  std::vector<double> mn((size_t)D * n), mx((size_t)D * n);
  for (uint32_t i = 0; i < n; ++i)
    for (int a = 0; a < D; ++a) {
      mn[(size_t)D * i + a] = aabbs[i].min()[a];
      mx[(size_t)D * i + a] = aabbs[i].max()[a];
    }
  uint64_t cap = 8ull * n + 16, np = 0;
  std::vector<uint32_t> pairs;
  for (;;) {
    pairs.resize(2 * cap);
    const int rc = scx_potential_overlaps(dev.get(), n, mn.data(), mx.data(), cap, pairs.data(), &np);
    if (rc == SCX_E_CAPACITY && np > cap) {
      cap = np;
      continue;
    }
    dev.check(rc, "scx_potential_overlaps");
    break;
  }
  for (uint64_t k = 0; k < np; ++k) overlaps.insert(std::make_pair(pairs[2 * k], pairs[2 * k + 1]));
}

// getPotentialOverlapsAllPairs (:40), on the host.
template <int D>
void getPotentialOverlapsAllPairs(const std::vector<AABB<D>>& aabbs,
                                  std::set<std::pair<unsigned, unsigned>>& overlaps) {
  for (size_t i = 0; i < aabbs.size(); ++i)
    for (size_t j = i + 1; j < aabbs.size(); ++j)
      if (aabbs[i].overlaps(aabbs[j])) overlaps.insert(std::make_pair((unsigned)i, (unsigned)j));
}

}  // namespace SpatialGridDetector

// The FlowableSystem facts the path uses: per-body mass and uniform gravity
// (ball2d/Forces/Ball2DGravityForce.cpp:36-46, rigidbody3d/Forces/NearEarthGravityForce.cpp:39-52).
struct GravitySystem {
  int dim = 2;
  std::vector<double> m;
  std::array<double, 3> g{{0.0, 0.0, 0.0}};
};

// UnconstrainedMap (scisim/UnconstrainedMaps/UnconstrainedMap.h:33).
class UnconstrainedMap {
 public:
  virtual ~UnconstrainedMap() = default;
  virtual void flow(const Device& dev, const std::vector<double>& q0, const std::vector<double>& v0,
                    const GravitySystem& fsys, unsigned iteration, double dt, std::vector<double>& q1,
                    std::vector<double>& v1) const {
    (void)iteration;
    q1.resize(q0.size());
    v1.resize(v0.size());
    dev.check(scx_unconstrained_flow(dev.get(), integrator(), (uint32_t)fsys.m.size(), fsys.m.data(), fsys.g.data(),
                                     dt, q0.data(), v0.data(), q1.data(), v1.data()),
              "scx_unconstrained_flow");
  }
  virtual std::string name() const = 0;
  virtual int32_t integrator() const = 0;
};

// SymplecticEulerMap (ball2d/SymplecticEulerMap.cpp:21-32).
class SymplecticEulerMap final : public UnconstrainedMap {
 public:
  std::string name() const override { return "symplectic_euler"; }
  int32_t integrator() const override { return SCX_INTEGRATOR_SYMPLECTIC_EULER; }
};

// VerletMap (ball2d/VerletMap.cpp:15-41); in 3D the translational part of SplitHamMap
// (rigidbody3d/UnconstrainedMaps/SplitHamMap.cpp:43-71,147-181).
class VerletMap final : public UnconstrainedMap {
 public:
  std::string name() const override { return "verlet"; }
  int32_t integrator() const override { return SCX_INTEGRATOR_VERLET; }
};

// An explicit active set (the data Sobogus::solve extracts from its Constraint objects).
struct ContactSet {
  std::vector<int32_t> body0, body1;  // body1 = -1: static geometry
  std::vector<double> bases;          // dim*dim per contact, column-major [n s t]
  std::vector<double> arm0, arm1;     // 3D: contact point - body centre, 3 per contact
  size_t size() const { return body0.size(); }
};

// FrictionSolver / Sobogus (scisim/ConstrainedMaps/FrictionSolver.h:28, Sobogus.h:76-104).
class Sobogus {
 public:
  explicit Sobogus(unsigned eval_every = 1, int32_t precision = SCX_PRECISION_FP64)
      : m_eval_every(eval_every), m_precision(precision) {}

  // Sobogus::solve (Sobogus.cpp:880-1088): alpha/beta are the warm start in and the
  // impulses out; v2 = v0 + M^-1 H^T r; f = friction-only generalized impulse.
  void solve(const Device& dev, unsigned iteration, double dt, const GravitySystem& fsys,
             const std::vector<double>& inertia, const std::vector<double>& CoR, const std::vector<double>& mu,
             const std::vector<double>& v0, const std::vector<double>& omega0, const ContactSet& active_set,
             const std::vector<double>& nrel_extra, unsigned max_iters, double tol, std::vector<double>& f,
             std::vector<double>& alpha, std::vector<double>& beta, std::vector<double>& v2,
             std::vector<double>& omega2, bool& solve_succeeded, double& error) const {
    (void)iteration;
    (void)dt;
    const int D = dev.dim();
    const size_t n = fsys.m.size(), C = active_set.size();
    alpha.resize(C, 0.0);
    beta.resize((size_t)(D - 1) * C, 0.0);
    v2.resize((size_t)D * n);
    omega2.resize(D == 3 ? 3 * n : 0);
    f.resize((D == 2 ? 2 : 6) * n);
    scx_contact_set cs{};
    cs.ncontacts = (uint32_t)C;
    cs.body0 = active_set.body0.data();
    cs.body1 = active_set.body1.data();
    cs.bases = active_set.bases.data();
    cs.arm0 = D == 3 ? active_set.arm0.data() : nullptr;
    cs.arm1 = D == 3 ? active_set.arm1.data() : nullptr;
    cs.mu = mu.data();
    cs.CoR = CoR.empty() ? nullptr : CoR.data();
    cs.nrel_extra = nrel_extra.empty() ? nullptr : nrel_extra.data();
    int32_t ok = 0;
    uint32_t sweeps = 0;
    dev.check(scx_friction_solve(dev.get(), (uint32_t)n, fsys.m.data(), inertia.empty() ? nullptr : inertia.data(),
                                 v0.data(), omega0.empty() ? nullptr : omega0.data(), &cs, tol, max_iters,
                                 m_eval_every, m_precision, alpha.data(), beta.data(), v2.data(),
                                 D == 3 ? omega2.data() : nullptr, f.data(), &ok, &error, &sweeps),
              "scx_friction_solve");
    solve_succeeded = ok != 0;
    m_last_sweeps = sweeps;
  }
  unsigned numFrictionImpulsesPerNormal(unsigned ambient_space_dimensions) const { return ambient_space_dimensions - 1; }
  std::string name() const { return "sobogus"; }
  unsigned evalEvery() const { return m_eval_every; }
  int32_t precision() const { return m_precision; }
  unsigned lastSweeps() const { return m_last_sweeps; }

 private:
  unsigned m_eval_every;
  int32_t m_precision;
  mutable unsigned m_last_sweeps = 0;
};

enum class ImpulsesToCache { NONE, NORMAL_AND_FRICTION };

// A circle/sphere scene: what Ball2DSim / RigidBody3DSim hold for this path.
struct BallScene {
  int dim = 2;
  std::vector<double> r, m, inertia;        // inertia: 3D, empty = 2/5 m r^2
  std::vector<double> omega;                // 3D angular velocity (3n), empty = 0
  std::vector<double> plane_x, plane_n;     // dim per static plane
};

// GeometricImpactFrictionMap::flow (scisim/ConstrainedMaps/GeometricImpactFrictionMap.cpp:267-442)
// for circles/spheres: predictor, active set, bases, solve on v0, final flow from q0 with v2.
class GeometricImpactFrictionMap {
 public:
  GeometricImpactFrictionMap(double abs_tol, unsigned max_iters, ImpulsesToCache cache = ImpulsesToCache::NONE)
      : m_tol(abs_tol), m_max_iters(max_iters), m_cache(cache) {}

  void flow(const Device& dev, BallScene& scene, const UnconstrainedMap& umap, const Sobogus& solver,
            unsigned iteration, double dt, const std::array<double, 3>& g, double CoR, double mu,
            const std::vector<double>& q0, const std::vector<double>& v0, std::vector<double>& q1,
            std::vector<double>& v1) {
    const int D = dev.dim();
    const uint32_t n = (uint32_t)scene.r.size();
    dev.check(scx_set_planes(dev.get(), (uint32_t)(scene.plane_x.size() / D), scene.plane_x.data(),
                             scene.plane_n.data()),
              "scx_set_planes");
    dev.check(scx_set_bodies(dev.get(), n, q0.data(), v0.data(), scene.omega.empty() ? nullptr : scene.omega.data(),
                             scene.r.data(), scene.m.data(), scene.inertia.empty() ? nullptr : scene.inertia.data()),
              "scx_set_bodies");
    scx_step_params p{};
    p.dt = dt;
    p.g[0] = g[0];
    p.g[1] = g[1];
    p.g[2] = g[2];
    p.mu = mu;
    p.CoR = CoR;
    p.tol = m_tol;
    p.max_iters = m_max_iters;
    p.eval_every = solver.evalEvery();
    p.integrator = umap.integrator();
    p.precision = solver.precision();
    p.warm_start = SCX_WARM_NONE;
    p.iteration = iteration;
    dev.check(scx_step(dev.get(), &p, &m_stats), "scx_step");
    if (!m_stats.succeeded && m_stats.ncontacts > 0) m_warnings++;
    q1.resize(q0.size());
    v1.resize(v0.size());
    if (D == 3 && scene.omega.size() != 3 * (size_t)n) scene.omega.assign(3 * (size_t)n, 0.0);
    dev.check(scx_get_state(dev.get(), q1.data(), v1.data(), D == 3 ? scene.omega.data() : nullptr),
              "scx_get_state");
  }
  void resetCachedData() {}
  std::string name() const { return "geometric_impact_friction_map"; }
  const scx_step_stats& lastStats() const { return m_stats; }
  unsigned warnings() const { return m_warnings; }

 private:
  double m_tol;
  unsigned m_max_iters;
  ImpulsesToCache m_cache;
  scx_step_stats m_stats{};
  unsigned m_warnings = 0;
};

}  // namespace scx_scisim
