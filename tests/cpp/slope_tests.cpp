// slope_tests.cpp -- a ball on an inclined static plane, stepped with the C++ host
// interface (GeometricImpactFrictionMap + VerletMap + Sobogus), against the closed form
// the reference's checker uses (assets/ball2d/shell_scripts/sliding_particle_test.py:38-72,
// static_particle_test.py): slides with a = g (-sin + mu cos) above the friction angle,
// stays put below it.
// Arguments: plane_x plane_y normal_x normal_y radius mass dt steps mu tol_x tol_y tol_vx tol_vy
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "scx_scisim.hpp"

int main(int argc, char** argv) {
  if (argc != 14) {
    std::cerr << "Usage: " << argv[0] << " px py nx ny r m dt steps mu tx ty tvx tvy" << std::endl;
    return EXIT_FAILURE;
  }
  double a[14];
  for (int i = 1; i < 14; ++i) a[i] = std::atof(argv[i]);
  const double dt = a[7], mu = a[9];
  const int steps = std::atoi(argv[8]);
  try {
    scx_scisim::Device dev(2, 1, 4);
    scx_scisim::BallScene scene;
    scene.dim = 2;
    scene.r = {a[5]};
    scene.m = {a[6]};
    scene.plane_x = {a[1], a[2]};
    scene.plane_n = {a[3], a[4]};
    scx_scisim::VerletMap umap;
    scx_scisim::Sobogus solver(1);
    scx_scisim::GeometricImpactFrictionMap ifmap(1.0e-12, 5000);
    std::vector<double> q{0.0, 0.0}, v{0.0, 0.0}, q1, v1;
    for (int it = 1; it <= steps; ++it) {
      ifmap.flow(dev, scene, umap, solver, it, dt, {0.0, -10.0, 0.0}, 0.0, mu, q, v, q1, v1);
      q.swap(q1);
      v.swap(v1);
    }
    const double t = steps * dt, g = 10.0;
    double nx = a[3], ny = a[4];
    const double nn = std::sqrt(nx * nx + ny * ny);
    nx /= nn;
    ny /= nn;
    const double theta = std::fabs(std::atan2(nx, ny));
    double dx = 0, dy = 0, vx = 0, vy = 0;
    const double acc = -g * std::sin(theta) + g * mu * std::cos(theta);
    if (acc < 0.0) {
      const double dpl = 0.5 * t * t * acc, vpl = t * acc;
      dx = dpl * std::cos(theta) * (nx > 0 ? -1.0 : 1.0);
      dy = dpl * std::sin(theta);
      vx = vpl * std::cos(theta) * (nx > 0 ? -1.0 : 1.0);
      vy = vpl * std::sin(theta);
    }
    const double err[4] = {std::fabs(q[0] - dx), std::fabs(q[1] - dy), std::fabs(v[0] - vx), std::fabs(v[1] - vy)};
    bool ok = true;
    for (int k = 0; k < 4; ++k) ok = ok && err[k] <= a[10 + k];
    std::printf("t=%g q=(%.6e, %.6e) v=(%.6e, %.6e) expected q=(%.6e, %.6e) v=(%.6e, %.6e) warnings=%u %s\n", t, q[0],
                q[1], v[0], v[1], dx, dy, vx, vy, ifmap.warnings(), ok ? "ok" : "FAIL");
    return ok ? EXIT_SUCCESS : EXIT_FAILURE;
  } catch (const scx_scisim::Error& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
}
