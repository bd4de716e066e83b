/*
 * spm_udf.h -- the subset of KAPSEL's UDF input the SPM step needs.
 *
 * KAPSEL reads its parameters from a UDF file through the GOURMET
 * UDFManager (input.cxx:438-1813, schema Udf/define.udf).  That library is
 * not part of the reference, so this is a reader of its own for the same
 * format: a schema (\begin{global_def} / \begin{def}: classes, records,
 * `select` alternatives, arrays `name[]`, int/double/float/string) and a
 * data section whose values follow the schema in order (a record's values
 * in braces, an array's in brackets, every alternative record of a select
 * present after the selecting string).  \include{"file"} is resolved
 * relative to the including file.
 *
 * spm_udf_params maps a parsed input file to spm_params / spm_electrolyte
 * the way input.cxx does for constitutive_eq Navier_Stokes and Electrolyte
 * with spherical particles.  All functions return SPM_OK or a status of
 * include/spm.h; spm_udf_last_error describes the last failure.
 */
#ifndef SPM_UDF_H
#define SPM_UDF_H

#include "spm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct spm_udf spm_udf;

/* Parses `path` (schema and data, following \include). */
int spm_udf_open(const char *path, spm_udf **out);
void spm_udf_close(spm_udf *udf);
const char *spm_udf_last_error(void);

/* Values by location, as UDFManager::get(Location, T&): dotted member
 * names with [i] for array elements, e.g. "mesh.NPX",
 * "object_type.spherical_particle.Particle_spec[0].MASS_RATIO". */
int spm_udf_get_double(const spm_udf *udf, const char *location, double *value);
int spm_udf_get_int(const spm_udf *udf, const char *location, int *value);
int spm_udf_get_string(const spm_udf *udf, const char *location, char *buf, int buflen);
/* Element count of an array location ("...Particle_spec[]" or without []). */
int spm_udf_size(const spm_udf *udf, const char *location, int *n);
/* Every number a location with [] steps reaches, in order, as
 * UDFManager's get("Particles[].R.x") returns a list: *n is set to the
 * count; SPM_ERR_INVALID if it exceeds cap.  cap 0 with out NULL counts. */
int spm_udf_get_doubles(const spm_udf *udf, const char *location, double *out, int cap, int *n);

/* Records (\begin{record} blocks, one per snapshot, as UDFManager's
 * totalRecord / jump): after spm_udf_jump(udf, i) the getters read record
 * i's entries (E, t, Particles[] ...) before the data section; -1 returns
 * to the data section alone. */
int spm_udf_num_records(const spm_udf *udf, int *n);
int spm_udf_jump(spm_udf *udf, int record);

/* Output file (the -O file of the reference): the input's text with every
 * \include inlined, so the file stands alone. */
int spm_udf_create_output(const char *in_path, const char *out_path);
/* Appends one record as Output_udf writes it (avs_output.cxx:343-396):
 * E = 1.0, t = ts, Particles[i] = {R, R_raw, v, q, omega, f_hydro,
 * torque_hydro, f_r, torque_r, f_slip, torque_slip} from x, x_nopbc, v, q,
 * omega, f_hydro_previous, torque_hydro_previous, fr_previous, 0,
 * f_slip_previous, torque_slip_previous, each number with 17 significant
 * digits so that it reads back exactly. */
int spm_udf_append_record(const char *out_path, int ts, const spm_particle *p, int n);

/* Run control read with the parameters (input.cxx:1384-1656). */
enum spm_udf_equation { SPM_UDF_NAVIER_STOKES = 0, SPM_UDF_LEES_EDWARDS = 2, SPM_UDF_ELECTROLYTE = 3 };
enum spm_udf_distribution {
  SPM_UDF_NONE = 0, SPM_UDF_UNIFORM_RANDOM = 1, SPM_UDF_RANDOM_WALK = 2, SPM_UDF_FCC = 3, SPM_UDF_BCC = 4,
  SPM_UDF_USER_SPECIFY = 5
};
typedef struct spm_udf_run {
  int equation;          /* constitutive_eq.type                           */
  int auto_dt;           /* time_increment.type == "auto"                  */
  double dt_factor;      /* time_increment.auto.factor (Axel)              */
  int gts, num_snap;     /* output.GTS, output.Num_snap; MSTEP = GTS*Num_snap */
  int distribution;      /* switch.INIT_distribution.type                  */
  int poisson_boltzmann; /* Electrolyte INIT_profile == "Poisson_Boltzmann" */
  double shear_rate;     /* Shear_Navier_Stokes_Lees_Edwards External_field Shear_rate */
} spm_udf_run;

/* input.cxx:438-1813 (subset) and the derived time step of
 * Set_global_parameters (input.cxx:316-381).  `el` may be NULL for
 * Navier_Stokes input; for Electrolyte input it is filled.
 * Shear_Navier_Stokes_Lees_Edwards input gives run->shear_rate (its AC
 * branch reads Shear_rate as a steady shear, input.cxx:646-653).  Rejects (SPM_ERR_UNSUPPORTED) the
 * equations, object types and switches the library does not run:
 * Shear_Navier_Stokes (sheared walls), chains, rigid bodies, janus
 * particles, pinning. */
int spm_udf_params(const spm_udf *udf, spm_params *prm, spm_electrolyte *el, spm_udf_run *run);

/* The particles Init_Particle sets up for INIT_distribution user_specify
 * (init_particle.cxx:226-288 and 318-393): x from Particles[i].R, species by
 * Particle_spec counts in order, v and omega zero (the reference reads them
 * and then sets both to 0.0*RA()), q from Particles[i].q normalized when
 * ROTATION is ON and INIT_orientation is user_specify and (1,0,0,0)
 * otherwise, q_old = q; every other field zero.  n must be the array's
 * length.  INIT_orientation random with ROTATION ON is SPM_ERR_UNSUPPORTED. */
int spm_udf_user_particles(const spm_udf *udf, spm_particle *p, int n);
/* Init_Particle (init_particle.cxx:12-393) for spherical particles: as
 * spm_udf_user_particles for user_specify; uniform_random (positions drawn
 * uniformly in the box, redrawn while closer than 1.05 SIGMA to an earlier
 * one under periodic images) and BCC (the reference's lattice) place the
 * particles; INIT_orientation random with ROTATION ON draws q by
 * RA_on_sphere4D.  Positions and orientations come from glibc's rand()
 * sequence from its default state, as in the reference.  FCC and
 * random_walk (init_particle.cxx:50-224) place an FCC lattice, then move
 * every particle once by a random step of 1e-3 RADIUS and once down its LJ
 * force by 0.05 RADIUS (Random_Walk / Steepest_descent, fluct.h:66-120,
 * with the linked-cell LJ of md_force.cxx:23-134, whose double count of
 * pairs under 3 cells per axis is kept because it decides the positions);
 * random_walk repeats the pair random_walk.iteration times with a step of
 * 0.025 RADIUS.  The random steps come from MT19937 seeded with 0 (the
 * reference's MT_seed(GIVEN_SEED, 0) precedes Init_Particle).  Boxes under
 * 16 points per axis are SPM_ERR_UNSUPPORTED for these two.  n is the
 * Particle_number total. */
int spm_udf_init_particles(const spm_udf *udf, spm_particle *p, int n);
/* As spm_udf_init_particles, and the MT19937 state the initialisation
 * leaves (the seed-0 state unless FCC or random_walk drew from it): the
 * reference's thermostat goes on from it, so a caller hands it to
 * spm_put_rng. */
int spm_udf_init_particles_rng(const spm_udf *udf, spm_particle *p, int n, uint32_t mt[SPM_MT_N], int *mti);

/* Restart data (resume.CONTINUE.Saved_Data of Udf/define.udf).  The caller
 * sets the sizes and owns the buffers: zeta0/zeta1 hold nx*ny*2*(nz/2+1)
 * values in the field layout, concentration n_ion times as many (ion
 * spectra, n_ion 0 for Navier_Stokes), particles n_particles records. */
typedef struct spm_restart {
  int nx, ny, nz;
  int n_particles;
  int n_ion;
  int ts;                /* jikan.ts   */
  double time;           /* jikan.time */
  double uk_dc[3];
  double degree_oblique; /* oblique.degree_oblique (0 without shear) */
  double *zeta0, *zeta1;
  double *concentration;
  spm_particle *particles;
} spm_restart;

/* Save_Restart_udf (resume.cxx:10-78, Save_Particle_udf): writes the input
 * file's text (includes inlined, data section re-serialized) with
 * resume.Calculation = "CONTINUE" and the saved data, every number with 17
 * significant digits.  A file of text: for small boxes; spm_save_state is
 * the exact, compact checkpoint. */
int spm_udf_write_restart(const char *in_path, const char *out_path, const spm_restart *r);
/* Force_restore_parameters (resume.cxx:274-364, Read_Particle_udf): fills
 * r from resume.CONTINUE.Saved_Data of a file whose resume.Calculation is
 * "CONTINUE".  Per particle: the 17 saved vectors, q and q_old normalized,
 * x_previous = x_nopbc = R, v_slip, omega_slip and the mass / inertia
 * temporaries zero; other members keep the caller's values.  The sizes in r
 * must match the file's arrays; on any error r is left unchanged. */
int spm_udf_read_restart(const spm_udf *udf, spm_restart *r);

#ifdef __cplusplus
}
#endif

#endif /* SPM_UDF_H */
