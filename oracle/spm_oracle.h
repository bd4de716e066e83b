/*
 * spm_oracle.h -- CPU FP64 restatement of KAPSEL's SPM time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the parity checker and the CPU
 * baseline timed by bench.py; the product (kapsel_old_amd/libspm.so) never
 * links, loads or calls it.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may use it.
 *
 * Each function restates the reference routine cited beside it.  The entry
 * points mirror include/spm.h one for one (prefix oracle_ instead of spm_),
 * so a parity test drives both through the same sequence of calls.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   FFT (A2a_k/A_k2a)      -- Ooura rdft3d/rdft3dsort compiled from the
 *                             reference (oracle/_ref), numpy.fft, brute DFT.
 *   MT19937 / Gauss2       -- mt19937ar.c compiled from the reference,
 *                             the C++ std::mt19937 known answer, Marsaglia.
 *   2/3 truncation         -- de-aliasing of products vs exact convolution.
 *   zeta<->u, projection   -- curl/divergence identities vs numpy spectra.
 *   advection              -- curl(u x omega) from numpy spectral derivatives,
 *                             Beltrami (ABC) flow gives zero.
 *   slaved Euler           -- exact viscous decay of a shear mode.
 *   stencil / profile      -- NP_domain counts measured on the reference,
 *                             phi(a) = 1/2, phi = 1 inside, 0 outside.
 *   make_phi / f_hydro     -- closed forms (uniform flow, rigid rotation).
 *   LJ (linked cells)      -- all-pairs brute force, Newton's third law,
 *                             the reference's 2x double count at lc = 2.
 *   integrators            -- AB2 exact on linear velocity, quaternion
 *                             rotation about a fixed axis.
 *   full step              -- invariants only (momentum balance, rigid
 *                             motion inside particles); the coupled
 *                             multi-step trajectory is parity unpinned.
 */
#ifndef SPM_ORACLE_H
#define SPM_ORACLE_H

#include <stdint.h>
#include "../include/spm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ctx oracle_ctx;

oracle_ctx *oracle_create(const spm_params *prm);
void oracle_destroy(oracle_ctx *c);
int oracle_stencil_size(const oracle_ctx *c);
int oracle_stencil(const oracle_ctx *c, int *out_xyz); /* 3*S ints */

int oracle_put_particles(oracle_ctx *c, const spm_particle *p, int n);
int oracle_get_particles(oracle_ctx *c, spm_particle *p, int n);
int oracle_put_field(oracle_ctx *c, int field, const double *host);
int oracle_get_field(oracle_ctx *c, int field, double *host);
int oracle_put_uk_dc(oracle_ctx *c, const double uk_dc[3]);
int oracle_get_uk_dc(oracle_ctx *c, double uk_dc[3]);
int oracle_put_time(oracle_ctx *c, int ts, double time);
int oracle_get_time(oracle_ctx *c, int *ts, double *time);
int oracle_put_rng(oracle_ctx *c, const uint32_t mt[624], int mti);
int oracle_get_rng(oracle_ctx *c, uint32_t mt[624], int *mti);

int oracle_init_fluid(oracle_ctx *c);
int oracle_ns_solver_slaved_euler(oracle_ctx *c);
int oracle_zeta_k2u(oracle_ctx *c);
int oracle_md_solver_position(oracle_ctx *c);
int oracle_calc_f_hydro(oracle_ctx *c);
int oracle_force(oracle_ctx *c);
int oracle_md_solver_velocity(oracle_ctx *c);
int oracle_add_random_force_thermostat(oracle_ctx *c);
int oracle_make_phi_u_particle(oracle_ctx *c);
int oracle_make_phi_particle(oracle_ctx *c);
int oracle_make_f_particle_dt_sole(oracle_ctx *c);
int oracle_u2zeta_k(oracle_ctx *c);
int oracle_step(oracle_ctx *c, int nsteps);

/* Electrolyte (SW_EQ == Electrolyte), mirroring spm_set_electrolyte .. */
int oracle_set_electrolyte(oracle_ctx *c, const spm_electrolyte *el);
int oracle_init_rho_ion(oracle_ctx *c);
int oracle_init_rho_ion_pb(oracle_ctx *c, double tol, int *iterations);
int oracle_put_concentration(oracle_ctx *c, int species, const double *host);
int oracle_get_concentration(oracle_ctx *c, int species, double *host);
int oracle_put_total_solute(oracle_ctx *c, const double *total);
int oracle_get_total_solute(oracle_ctx *c, double *total);
int oracle_nssolute_solver_euler(oracle_ctx *c);
int oracle_rescale_solute(oracle_ctx *c);
int oracle_make_coulomb_force(oracle_ctx *c);
void oracle_make_phi_qq(oracle_ctx *c, double *phi, double *surface);
void oracle_make_surface_normal(oracle_ctx *c, double *const normal[3]);

/* Lees-Edwards shear, fluid only (Time_evolution_hydro_OBL) */
int oracle_set_lees_edwards(oracle_ctx *c, double shear_rate);
int oracle_put_oblique(oracle_ctx *c, double degree_oblique, double shear_rate_eff);
int oracle_get_oblique(oracle_ctx *c, double *degree_oblique, double *shear_rate_eff);
/* spherical particles under shear; the Cartesian velocity is u0..u2 */
int oracle_make_phi_u_particle_obl(oracle_ctx *c);
int oracle_calc_f_hydro_obl(oracle_ctx *c);
int oracle_md_solver_position_obl(oracle_ctx *c);
int oracle_md_solver_velocity_obl(oracle_ctx *c);
int oracle_force_obl(oracle_ctx *c);
int oracle_le_shear_stress(oracle_ctx *c, double out[7]);
int oracle_u2u_oblique(oracle_ctx *c);
int oracle_u_oblique2u(oracle_ctx *c);
int oracle_remap_obl(oracle_ctx *c); /* the integer-strain remap of the real-space u */
double oracle_distance0_obl(oracle_ctx *c, const double *x1, const double *x2, double *x12);
int oracle_ns_solver_shear_obl(oracle_ctx *c);
int oracle_zeta_k2u_k_obl(oracle_ctx *c);
int oracle_u_k2zeta_k_obl(oracle_ctx *c);

/* Building blocks, on caller buffers of NX*NY*NZ_ doubles. */
void oracle_A2a_k(oracle_ctx *c, double *a);
void oracle_A_k2a(oracle_ctx *c, double *a);
void oracle_truncate(oracle_ctx *c, double *a);
void oracle_zeta_k2u_k(oracle_ctx *c, double *const zeta[2], const double uk_dc[3], double *const u[3]);
void oracle_u_k2zeta_k(oracle_ctx *c, double *const u[3], double *const zeta[2], double uk_dc[3]);
void oracle_u2advection_k(oracle_ctx *c, double *const u[3], double *const adv[2]);
void oracle_solenoidal_uk(oracle_ctx *c, double *const u[3]);
double oracle_phi(const oracle_ctx *c, double x);
/* Linked-cell LJ of md_force.cxx:23-134 and the all-pairs variant of
 * md_force.cxx:136-179, on positions x[3*n]; forces added into fr[3*n]. */
void oracle_lj_lnk(oracle_ctx *c, const double *x, int n, double *fr);
void oracle_lj_allpairs(oracle_ctx *c, const double *x, int n, double *fr);
/* MT19937 and the polar Gaussian pair generator (fluct.h:48-60). */
uint32_t oracle_genrand_int32(oracle_ctx *c);
double oracle_genrand_real3(oracle_ctx *c);
void oracle_gauss2(oracle_ctx *c, double out[2]);

void oracle_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
