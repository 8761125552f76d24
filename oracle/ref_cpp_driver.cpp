// C entry points around the reference's own C++ forward solver
// (include/morphompm/growth_solver.h, FP32, forward only), compiled from the
// reference sources by oracle/Makefile into oracle/_ref/libref_growth.so.
//
// TEST INFRASTRUCTURE ONLY: a second, independent forward checker.  Its
// semantics differ from python/morphompm/transfer.py in documented ways
// (singular values clamped >= 1e-6, grid mass mask 1e-9, zero-weight nodes
// skipped), so it is compared with the loose tolerances of the reference's
// own numpy <-> C++ parity gate (scripts/parity.py:48).
#include <vector>

#include "morphompm/growth_solver.h"

using morphompm::GrowthSolver;
using morphompm::math::Matrix3;
using morphompm::math::Vec3;

static Matrix3 load(const float* a) {
  return Matrix3(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8]);
}
static void store(const Matrix3& M, float* a) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) a[3 * r + c] = M.m[r][c];
}

extern "C" {

// Hencky Kirchhoff stress of the reference solver (growth_solver.h:82-99).
void ref_kirchhoff(const float* F, const float* Fg, float E, float nu, float* tau,
                   float* eps_norm) {
  GrowthSolver s(8, 0.1f);
  s.set_material(E, nu);
  float en = 0.0f;
  store(s.kirchhoff_stress(load(F), load(Fg), &en), tau);
  if (eps_norm) *eps_norm = en;
}

// Runs `steps` forward steps of n particles (positions x[3n], growth Fg[9n],
// zero gravity) and writes the final positions and F.
void ref_run(int N, float dx, float E, float nu, float damping, float density, float sp, int n,
             const float* x, const float* Fg, int steps, float dt, float* x_out, float* F_out) {
  GrowthSolver s(N, dx);
  s.set_material(E, nu);
  s.set_gravity(Vec3(0.0f, 0.0f, 0.0f));
  s.set_damping(damping);
  const float pmass = density * sp * sp * sp;
  for (int p = 0; p < n; ++p) s.add_particle(Vec3(x[3 * p], x[3 * p + 1], x[3 * p + 2]), pmass, density);
  for (int p = 0; p < n; ++p) s.particles()[p].Fg = load(Fg + 9 * p);
  for (int k = 0; k < steps; ++k) s.step(dt);
  for (int p = 0; p < n; ++p) {
    const auto& q = s.particles()[p];
    x_out[3 * p] = q.x.x;
    x_out[3 * p + 1] = q.x.y;
    x_out[3 * p + 2] = q.x.z;
    store(q.F, F_out + 9 * p);
  }
}

}  // extern "C"
