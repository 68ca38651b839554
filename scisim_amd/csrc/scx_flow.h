// scx_flow.h -- per-component unconstrained update under gravity (force = m g):
// SymplecticEulerMap (ball2d/SymplecticEulerMap.cpp:21-32), VerletMap
// (ball2d/VerletMap.cpp:15-41) and the translational part of SplitHamMap
// (rigidbody3d/UnconstrainedMaps/SplitHamMap.cpp:43-71,147-181), with the
// reference's operation order and no FMA contraction.
#pragma once

#include <hip/hip_runtime.h>

namespace scx {

// Flow modes beyond SCX_INTEGRATOR_SYMPLECTIC_EULER (0) and SCX_INTEGRATOR_VERLET (1):
// the predictor of the symplectic Euler staggering, v1 = v0 + dt (M^-1 F), q1 = q0 + dt v1
// (SymplecticEulerImpactFrictionMap.cpp:222-232), and the linear configuration update
// q1 = q0 + dt v0, v1 = v0 (linearInertialConfigurationUpdate, ball2d/Ball2DSim.cpp:92-98,
// rigidbody3d/UnconstrainedMaps/IntegrationTools.cpp:36-51).
constexpr int kFlowForceIncrement = 2;
constexpr int kFlowLinear = 3;

template <int D>
__host__ __device__ __forceinline__ void flow_component(int integrator, double m, double g, double dt, double q0,
                                                        double v0, double* q1, double* v1) {
#pragma clang fp contract(off)
  const double minv = 1.0 / m;
  const double F = m * g;
  if (integrator == kFlowForceIncrement) {
    const double v = v0 + dt * (minv * F);
    *v1 = v;
    *q1 = q0 + dt * v;
  } else if (integrator == kFlowLinear) {
    *v1 = v0;
    *q1 = q0 + dt * v0;
  } else if (D == 2 && integrator == 0) {
    const double v = v0 + (dt * minv) * F;
    *v1 = v;
    *q1 = q0 + dt * v;
  } else if (D == 2) {
    const double v = v0 + ((0.5 * dt) * minv) * F;
    *q1 = q0 + dt * v;
    *v1 = v + ((0.5 * dt) * minv) * F;
  } else {
    double p = m * v0;
    p = p + (0.5 * dt) * F;
    *q1 = q0 + (dt * v0 + (((0.5 * dt) * dt) * minv) * F);
    p = p + (0.5 * dt) * F;
    *v1 = p / m;
  }
}

}  // namespace scx
