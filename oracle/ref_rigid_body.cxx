// Test infrastructure: an extern "C" face over the reference's own rigid-body
// and quaternion code, compiled by oracle/Makefile from the reference tree
// (rigid_body.cxx, quaternion.h, particle_rotation_solver.h, variable.h) into
// oracle/_ref/libref.so.  The CPU tests use it to check the oracle's
// orientation integrators and the spm_particle layout; nothing in the product
// links or loads it.
#include <cstddef>

#include "variable.h"
// particle_rotation_solver.h names this global (input.h:97) in an inline
// solver that is never called here, so a declaration suffices.
extern double *IMOI;
#include "particle_rotation_solver.h"

extern "C" {

// MD_solver_orientation_Euler / _AB2 (particle_rotation_solver.h:15-39) on a
// Particle in the reference's own layout.
void ref_orientation_euler(Particle *p, double dt) { MD_solver_orientation_Euler(*p, dt); }
void ref_orientation_ab2(Particle *p, double dt) { MD_solver_orientation_AB2(*p, dt); }

// rigid_body_rotation with a quaternion (rigid_body.cxx:10-33); q = (s, v0, v1, v2).
// transform 0 = BODY2SPACE, 1 = SPACE2BODY.
void ref_rigid_body_rotation(double *rotated, const double *original, const double *q4, int transform) {
  quaternion q;
  qtn_init(q, q4);
  rigid_body_rotation(rotated, original, q, transform == 0 ? BODY2SPACE : SPACE2BODY);
}

// qdot (rigid_body.cxx:48-64); frame 0 = BODY_FRAME, 1 = SPACE_FRAME.
void ref_qdot(double *dq4, const double *q4, const double *omega, int frame) {
  quaternion q, dq;
  qtn_init(q, q4);
  qdot(dq, q, omega, frame == 0 ? BODY_FRAME : SPACE_FRAME);
  dq4[0] = qtn_q0(dq);
  dq4[1] = qtn_q1(dq);
  dq4[2] = qtn_q2(dq);
  dq4[3] = qtn_q3(dq);
}

// qtn_normalize (quaternion.h:237-245).
void ref_qtn_normalize(double *q4) {
  quaternion q;
  qtn_init(q, q4);
  qtn_normalize(q);
  q4[0] = qtn_q0(q);
  q4[1] = qtn_q1(q);
  q4[2] = qtn_q2(q);
  q4[3] = qtn_q3(q);
}

// Layout of the reference Particle (variable.h:27-75): sizeof, then the byte
// offset of every member in declaration order.  Returns the number written.
int ref_particle_layout(size_t *out, int cap) {
  const size_t v[] = {
      sizeof(Particle),
      offsetof(Particle, spec), offsetof(Particle, x), offsetof(Particle, x_previous),
      offsetof(Particle, x_nopbc), offsetof(Particle, v), offsetof(Particle, v_old),
      offsetof(Particle, v_slip), offsetof(Particle, f_hydro), offsetof(Particle, f_hydro_previous),
      offsetof(Particle, f_hydro1), offsetof(Particle, f_slip), offsetof(Particle, f_slip_previous),
      offsetof(Particle, fr), offsetof(Particle, fr_previous), offsetof(Particle, omega),
      offsetof(Particle, omega_old), offsetof(Particle, omega_slip), offsetof(Particle, torque_hydro),
      offsetof(Particle, torque_hydro_previous), offsetof(Particle, torque_hydro1),
      offsetof(Particle, torque_slip), offsetof(Particle, torque_slip_previous),
      offsetof(Particle, momentum_depend_fr), offsetof(Particle, QR), offsetof(Particle, QR_old),
      offsetof(Particle, q), offsetof(Particle, q_old), offsetof(Particle, mass),
      offsetof(Particle, mass_center), offsetof(Particle, inertia), offsetof(Particle, surface_mass),
      offsetof(Particle, surface_mass_center), offsetof(Particle, surface_inertia),
      offsetof(Particle, surface_dv), offsetof(Particle, surface_dw)};
  int n = (int)(sizeof(v) / sizeof(v[0]));
  if (n > cap) n = cap;
  for (int i = 0; i < n; i++) out[i] = v[i];
  return n;
}

}  // extern "C"
