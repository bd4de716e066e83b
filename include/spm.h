/*
 * spm.h -- C ABI of the MI355X Smoothed Profile Method (SPM) time step.
 *
 * The library replaces KAPSEL's CPU time step `Time_evolution_hydro`
 * (sp_3d_ns.cxx:61-212) and the solver calls it is built from:
 *   NS_solver_slavedEuler          fluid_solver.cxx:328 (decl. fluid_solver.h:80)
 *   Make_phi_u_particle            make_phi.cxx:401     (decl. make_phi.h:70)
 *   Calc_f_hydro_correct_precision md_force.cxx:263     (decl. md_force.h:60)
 *   Force                          particle_solver.h:97
 * The reference keeps all of this in globals (input.cxx:13-271); here the
 * state lives in an opaque context that owns device memory on one GPU.
 *
 * Layouts at the boundary are the reference's:
 *   - every field is NX*NY*NZ_ doubles, NZ_ = 2*(NZ/2+1), index
 *     (i*NY + j)*NZ_ + k (e.g. operate_omega.cxx:37);
 *   - a real-space field uses k < NZ (the two padding slots are ignored);
 *   - a spectrum holds NZ/2+1 interleaved complex values per (i,j) row,
 *     X(k) = sum_x a(x) exp(+i k.x) (Ooura sign, fftsg3d.c:109-120), and
 *     the inverse scales by 1/(NX*NY*NZ) overall (fft_wrapper.h:552);
 *   - particles are the reference `Particle` struct (variable.h:27-75),
 *     1000 bytes each, mirrored as `spm_particle` below.
 *
 * Every entry point returns an int status (SPM_OK = 0) and never exits the
 * process (the reference calls exit(1), macro.h:194-196).  Calls on one
 * context are ordered on the context's HIP stream; results copied out by
 * spm_get_* are complete when the call returns.  A context is not
 * thread-safe; separate contexts are independent.
 */
#ifndef SPM_H
#define SPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPM_DIM 3
#define SPM_MAX_SPECIES 8
#define SPM_MT_N 624

/* ---- status codes ---------------------------------------------------- */
enum spm_status {
  SPM_OK = 0,
  SPM_ERR_INVALID = 1,       /* bad argument or parameter                */
  SPM_ERR_OOM = 2,           /* device or host allocation failed         */
  SPM_ERR_HIP = 3,           /* a HIP runtime call failed                */
  SPM_ERR_NONFINITE = 4,     /* NaN/Inf found by spm_check_finite        */
  SPM_ERR_UNSUPPORTED = 5,   /* option outside what this build implements*/
  SPM_ERR_NO_DEVICE = 6      /* no usable GPU                            */
};

/* ---- field identifiers for spm_get_field / spm_put_field -------------- */
enum spm_field {
  SPM_FIELD_ZETA0 = 0, SPM_FIELD_ZETA1 = 1,        /* reduced vorticity (k-space) */
  SPM_FIELD_U0 = 2, SPM_FIELD_U1 = 3, SPM_FIELD_U2 = 4,    /* fluid velocity    */
  SPM_FIELD_UP0 = 5, SPM_FIELD_UP1 = 6, SPM_FIELD_UP2 = 7, /* phi*(v+w x r)     */
  SPM_FIELD_PHI = 8,                                       /* smooth profile    */
  SPM_FIELD_F0 = 9, SPM_FIELD_F1 = 10, SPM_FIELD_F2 = 11,  /* SPM body force    */
  SPM_NUM_FIELDS = 12
};

/* ---- LJ switches (input.cxx:1451-1480) -------------------------------- */
enum spm_lj_truncate { SPM_LJ_NONE = -1, SPM_LJ_FULL = 0, SPM_LJ_WCA = 1 };
enum spm_lj_powers { SPM_LJ_12_6 = 0, SPM_LJ_24_12 = 1, SPM_LJ_36_18 = 2 };

/* ---- arithmetic of the field path ------------------------------------- */
enum spm_precision { SPM_FP64 = 0, SPM_FP32 = 1 };

/*
 * Run parameters: the subset of the reference's UDF input
 * (Udf/define.udf:14-225, read in input.cxx:438-1813) that the
 * Navier_Stokes + spherical_particle time step uses.  Derived constants
 * follow Set_global_parameters (input.cxx:274-433).
 */
typedef struct spm_params {
  int nx, ny, nz;            /* grid; each a power of two, 8..1024, or   *
                              * 48, 80, 96, 160, 192 (literal schedule,   *
                              * no Lees-Edwards shear)                    */
  double dx;                 /* DX                                        */
  double rho, eta;           /* RHO, ETA (NU = ETA/RHO)                   */
  double kbt;                /* kBT; thermostat runs when > 0             */
  double alpha_v, alpha_o;   /* thermostat factors (input.cxx:518-519)    */
  double dt;                 /* DT (time_increment manual)                */
  double a, a_xi;            /* radius and interface width, in units of DX */
  int n_particles;           /* Particle_Number                           */
  int n_species;             /* Component_Number, 1..SPM_MAX_SPECIES      */
  double mass_ratio[SPM_MAX_SPECIES];  /* MASS_RATIOS                     */
  int rotation;              /* ROTATION switch                           */
  int lj_truncate;           /* enum spm_lj_truncate                      */
  int lj_powers;             /* enum spm_lj_powers                        */
  double epsilon;            /* EPSILON                                   */
  double g;                  /* gravity G                                 */
  int g_direction;           /* 0,1,2 for -X,-Y,-Z                        */
  int fix_cell[SPM_DIM];     /* FIX_CELLxyz                               */
  unsigned long seed;        /* MT19937 seed (reference: 0, sp_3d_ns.cxx:477) */
  int precision;             /* enum spm_precision                        */
  int device;                /* HIP device ordinal                        */
} spm_params;

/* The reference quaternion (quaternion.h:22-25). */
typedef struct spm_quaternion {
  double s;
  double v[SPM_DIM];
} spm_quaternion;

/* Byte-for-byte mirror of the reference `Particle` (variable.h:27-75). */
typedef struct spm_particle {
  int spec;
  double x[SPM_DIM];
  double x_previous[SPM_DIM];
  double x_nopbc[SPM_DIM];
  double v[SPM_DIM];
  double v_old[SPM_DIM];
  double v_slip[SPM_DIM];
  double f_hydro[SPM_DIM];
  double f_hydro_previous[SPM_DIM];
  double f_hydro1[SPM_DIM];
  double f_slip[SPM_DIM];
  double f_slip_previous[SPM_DIM];
  double fr[SPM_DIM];
  double fr_previous[SPM_DIM];
  double omega[SPM_DIM];
  double omega_old[SPM_DIM];
  double omega_slip[SPM_DIM];
  double torque_hydro[SPM_DIM];
  double torque_hydro_previous[SPM_DIM];
  double torque_hydro1[SPM_DIM];
  double torque_slip[SPM_DIM];
  double torque_slip_previous[SPM_DIM];
  double momentum_depend_fr[SPM_DIM];
  double QR[SPM_DIM][SPM_DIM];
  double QR_old[SPM_DIM][SPM_DIM];
  spm_quaternion q;
  spm_quaternion q_old;
  double mass;
  double mass_center[SPM_DIM];
  double inertia[SPM_DIM][SPM_DIM];
  double surface_mass;
  double surface_mass_center[SPM_DIM];
  double surface_inertia[SPM_DIM][SPM_DIM];
  double surface_dv[SPM_DIM];
  double surface_dw[SPM_DIM];
} spm_particle;

typedef struct spm_ctx spm_ctx;

/* ---- lifetime ---------------------------------------------------------- */
/* Allocates the context and its device buffers; fields start at zero,
 * particles at rest at the origin, ts = 0, the RNG seeded with prm->seed
 * (MT_seed, fluct.h:34-46). Replaces Gourmet_file_io + Init_fft +
 * Mem_alloc_var (input.cxx:438, fft_wrapper.cxx:92, sp_3d_ns.cxx:392). */
int spm_create(const spm_params *prm, spm_ctx **out);
int spm_destroy(spm_ctx *ctx);
const char *spm_strerror(int status);
/* Text of the last error on this context ("" if none). */
const char *spm_last_error(const spm_ctx *ctx);
/* Number of points in the integration stencil (NP_domain, profile.cxx:25-74). */
int spm_stencil_size(const spm_ctx *ctx);

/* ---- state transfer (host <-> device, reference layouts) --------------- */
int spm_put_particles(spm_ctx *ctx, const spm_particle *p, int n);
int spm_get_particles(spm_ctx *ctx, spm_particle *p, int n);
int spm_put_field(spm_ctx *ctx, int field, const double *host);
int spm_get_field(spm_ctx *ctx, int field, double *host);
int spm_put_uk_dc(spm_ctx *ctx, const double uk_dc[SPM_DIM]);
int spm_get_uk_dc(spm_ctx *ctx, double uk_dc[SPM_DIM]);
/* CTime jikan.ts / jikan.time (variable.h:18-25). */
int spm_put_time(spm_ctx *ctx, int ts, double time);
int spm_get_time(spm_ctx *ctx, int *ts, double *time);
/* MT19937 state (mt19937ar.c:60-61): 624 words and the index mti. */
int spm_put_rng(spm_ctx *ctx, const uint32_t mt[SPM_MT_N], int mti);
int spm_get_rng(spm_ctx *ctx, uint32_t mt[SPM_MT_N], int *mti);

/* ---- the solver call surface ------------------------------------------ */
/* Init_zeta_k (init_fluid.cxx:10) followed by the initial projection
 * block of main (sp_3d_ns.cxx:508-524). */
int spm_init_fluid(spm_ctx *ctx);
/* NS_solver_slavedEuler (fluid_solver.cxx:328-354) over the four 2/3-rule
 * blocks of Time_evolution_noparticle (sp_3d_ns.cxx:15-20). */
int spm_ns_solver_slaved_euler(spm_ctx *ctx);
/* FIX_CELL, truncation and Zeta_k2u of Time_evolution_hydro
 * (sp_3d_ns.cxx:67-79): zeta -> real-space u. */
int spm_zeta_k2u(spm_ctx *ctx);
/* MD_solver_position_Euler at ts == 0 else MD_solver_position_AB2
 * (particle_solver.cxx:36-88; sp_3d_ns.cxx:81-87). */
int spm_md_solver_position(spm_ctx *ctx);
/* Reset_phi_u + Calc_f_hydro_correct_precision(p, u) (md_force.cxx:263-389). */
int spm_calc_f_hydro(spm_ctx *ctx);
/* Force(p) (particle_solver.h:97-109): LJ/WCA + gravity, added to fr. */
int spm_force(spm_ctx *ctx);
/* MD_solver_velocity_Euler at ts == 0 else MD_solver_velocity_AB2_hydro
 * (particle_solver.cxx:106-147, 247-289); calls Force first. */
int spm_md_solver_velocity(spm_ctx *ctx);
/* Add_random_force_thermostat (fluct.cxx:10-86); no-op when kbt <= 0. */
int spm_add_random_force_thermostat(spm_ctx *ctx);
/* Reset_phi_u + Make_phi_u_particle (make_phi.cxx:401-409). */
int spm_make_phi_u_particle(spm_ctx *ctx);
/* Make_f_particle_dt_sole + Add_f_particle (f_particle.h:71-99). */
int spm_make_f_particle_dt_sole(spm_ctx *ctx);
/* U2zeta_k (operate_omega.h:186-189): real-space u -> zeta, uk_dc. */
int spm_u2zeta_k(spm_ctx *ctx);
/* nsteps x { Time_evolution_hydro (sp_3d_ns.cxx:61-212);
 *            jikan.time += dt; jikan.ts++ }   (sp_3d_ns.cxx:593-594). */
int spm_step(spm_ctx *ctx, int nsteps);
/* Waits for all work queued on the context. */
int spm_synchronize(spm_ctx *ctx);

/* How spm_step runs a step (no counterpart in the reference):
 *   SPM_SCHEDULE_FUSED (default) -- pointwise work fused into the FFT passes;
 *     after spm_step, zeta, uk_dc, u, phi, up and the particles read as the
 *     reference's end-of-step state (u, phi and up, which the next step does
 *     not read, are formed by the first call that reads or replaces them),
 *     the f field is scratch; power-of-two grids only (a
 *     grid with a mixed length runs LITERAL, and asking for FUSED there is
 *     SPM_ERR_UNSUPPORTED);
 *   SPM_SCHEDULE_LITERAL -- one kernel sequence per reference routine, every
 *     field (f included) as the reference leaves it. */
enum spm_schedule { SPM_SCHEDULE_FUSED = 0, SPM_SCHEDULE_LITERAL = 1 };
int spm_set_schedule(spm_ctx *ctx, int schedule);
/* Test hook (no counterpart in the reference): runs one stage of the fused
 * schedule so tests can compare it with the literal routines --
 *   1  NS_solver_slavedEuler
 *   2  FIX_CELL + truncation + Zeta_k2u (sp_3d_ns.cxx:67-79)
 *   3  Make_phi_u_particle .. U2zeta_k on the current u and particles
 *      (leaves u and the f field as scratch)
 *   4  rebuild u (the spectrum U2zeta_k leaves, operate_omega.h:186-189),
 *      phi and up from zeta and the particles. */
int spm_fused_stage(spm_ctx *ctx, int stage);
/* Scans u and zeta for NaN/Inf; SPM_ERR_NONFINITE if any is found. */
int spm_check_finite(spm_ctx *ctx);

/* ---- transforms of a caller's field, without a context -----------------
 * A2a_k (direction +1, fft_wrapper.h:432-527) and A_k2a (direction -1,
 * fft_wrapper.h:534-631) take a bare `double *` in the reference and
 * allocate a work copy on every call (:437,461,540,566).  spm_fft3d
 * transforms, in place, one FP64 field that lives in device memory in the
 * reference layout [NX][NY][NZ/2+1 complex] (real data in k < NZ): +1 is
 * the forward transform with the Ooura sign exp(+i k.x), -1 the inverse
 * with the 1/(NX NY NZ) scale and rdft3dsort's reading of the kz = 0 and
 * kz = NZ/2 planes.  Sizes are powers of two in [8, 1024] or one of the
 * mixed lengths 48, 80, 96, 160, 192 (factors 2, 3, 5; the reference's
 * Ooura transform has powers of two only).  `workspace`
 * (device memory of at least spm_fft_workspace_size() bytes, rewritten by
 * every call) holds the roots of unity; with workspace == NULL the call
 * allocates one, waits for `stream` and frees it.  `stream` is a
 * hipStream_t (NULL: the default stream); with a workspace the call only
 * queues work on it.  Errors: SPM_ERR_INVALID for bad sizes, a null field
 * or a short workspace (text in spm_last_error(NULL)). */
size_t spm_fft_workspace_size(int nx, int ny, int nz);
int spm_fft3d(int nx, int ny, int nz, int direction, double *field, void *workspace, size_t workspace_bytes,
              void *stream);
/* No counterpart in the reference: spm_fft3d for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same transform and results;
 * the root tables go to a 65536-byte static device array of the library
 * (one per GPU, held until the process ends) instead of a workspace, so the
 * call allocates nothing and only queues kernels on `stream`.  Sizes are
 * those of spm_fft3d (the tables then fit the array); others get
 * SPM_ERR_INVALID.  Every call on a GPU rewrites that GPU's array: calls
 * that overlap on one GPU (two streams or two threads) race. */
int spm_fft3d_noalloc(int nx, int ny, int nz, int direction, double *field, void *stream);

/* ---- checkpoint / resume ------------------------------------------------
 * Replace Save_Restart_udf / Force_restore_parameters (resume.cxx:10-364)
 * for exact resumption: the file holds zeta0/zeta1 in the context's own
 * precision, uk_dc, every particle record, ts and time, the MT19937 state
 * (which the reference's restart file does not keep) and, for an
 * electrolyte, the ion spectra and their totals, and under Lees-Edwards
 * shear degree_oblique and Shear_rate_eff, with a hash over all of it.  A
 * context created with the same spm_params (and spm_set_electrolyte or
 * spm_set_lees_edwards) that loads the file continues bit for bit like the
 * one that saved it.  spm_load_state rejects a file from another grid,
 * precision, particle count, ion count, time step or shear rate, or a
 * damaged one, and then leaves the context unchanged.  spm_save_state writes PATH.tmp and renames it. */
int spm_save_state(spm_ctx *ctx, const char *path);
int spm_load_state(spm_ctx *ctx, const char *path);

/* ---- measurement (no counterpart in the reference) --------------------
 * With timing on, every kernel launch is bracketed by HIP events on the
 * context's stream and attributed to a kernel family together with the
 * algorithmic HBM bytes it moves (each element it must read or write, once).
 * spm_timing_read waits for the stream and returns the totals since the
 * last read (and clears them). */
enum spm_kernel_family {
  SPM_K_FFT_Z_R2C = 0,   /* z pass, real -> half spectrum               */
  SPM_K_FFT_Z_C2R = 1,   /* z pass, half spectrum -> real               */
  SPM_K_FFT_Y = 2,       /* y pass, complex lines                       */
  SPM_K_FFT_X = 3,       /* x pass, complex lines                       */
  SPM_K_SPECTRAL = 4,    /* pointwise k-space kernels                   */
  SPM_K_REALSPACE = 5,   /* pointwise real-space kernels                */
  SPM_K_MAKE_PHI = 6,    /* phi / up scatter (owner-computes gather)    */
  SPM_K_F_HYDRO = 7,     /* hydrodynamic force / torque reduction       */
  SPM_K_PARTICLE = 8,    /* lists, LJ, integrators, thermostat          */
  SPM_NUM_KERNEL_FAMILIES = 9
};
int spm_timing_enable(spm_ctx *ctx, int on);
int spm_timing_read(spm_ctx *ctx, double ms[SPM_NUM_KERNEL_FAMILIES], double bytes[SPM_NUM_KERNEL_FAMILIES],
                    long launches[SPM_NUM_KERNEL_FAMILIES]);

/* ---- charged colloids in an electrolyte (SW_EQ == Electrolyte) ---------
 * The reference reads these from constitutive_eq.Electrolyte of the UDF
 * (input.cxx:660-790) into globals (input.cxx:238-258).  Valencies and
 * Onsager coefficients are indexed by ion species: with n_spec == 1,
 * [0] is the counterion (Valency_counterion, Onsager_coeff_counterion);
 * with n_spec == 2 (Add_salt), [0] the positive and [1] the negative ion. */
typedef struct spm_electrolyte {
  int n_spec;                                /* N_spec, 1 or 2                 */
  double valency[2];                         /* Valency_*                      */
  double onsager[2];                         /* Onsager_coeff_*                */
  double dielectric;                         /* Dielectric_cst                 */
  double elementary_charge;                  /* Elementary_charge (1 in KAPSEL) */
  double debye_length;                       /* Debye_length (n_spec == 2)     */
  double surface_charge[SPM_MAX_SPECIES];    /* Surface_charge per particle species */
  int external_field;                        /* External_field                 */
  int ac;                                    /* AC                             */
  double e_ext[3];                           /* E_ext                          */
  double angular_frequency;                  /* Angular_Frequency = 2 pi Frequency */
} spm_electrolyte;

/* Switches the context to the electrolyte equations (Time_evolution_hydro's
 * SW_EQ == Electrolyte branches, sp_3d_ns.cxx:46-57,89-121,183) and
 * allocates the ion fields.  Call once, after spm_create.  Electrolyte
 * steps run the literal schedule.  In an SPM_FP32 context the transport
 * fields (ion spectra and right-hand sides, fluxes, velocity, vorticity,
 * phi, f) are FP32 while the Poisson solve (charge density, potential and
 * its gradient in Fourier space, the particles' charge profile, surface
 * normals) and every grid sum stay FP64. */
int spm_set_electrolyte(spm_ctx *ctx, const spm_electrolyte *el);
/* Init_rho_ion (operate_electrolyte.cxx:695-815) with the uniform initial
 * distribution (Poisson_Boltzmann off): Set_uniform_ion_charge_density_salt
 * / _nosalt, then Count_solute_each.  Needs the particles. */
int spm_init_rho_ion(spm_ctx *ctx);
/* Init_rho_ion with Poisson_Boltzmann on (operate_electrolyte.cxx:
 * 786-805) for salt (n_spec 2) and more than one particle: the uniform
 * start, then Set_steadystate_ion_density (:374-412): the ions relax by
 * Ion_diffusion_solver_Euler with the fluid at rest until the free energy
 * of Calc_free_energy_PB changes by less than 1e-4, at most 9999 times
 * (`iterations`, may be NULL, receives the count).  u, phi, up and f are
 * working memory meanwhile, as in the reference.  The iterative profile of
 * a single particle or of salt-free ions is SPM_ERR_UNSUPPORTED. */
int spm_init_rho_ion_pb(spm_ctx *ctx, int *iterations);
/* Concentration[species] spectrum (reference layout, FP64 on the host) in
 * and out. */
int spm_put_concentration(spm_ctx *ctx, int species, const double *host);
int spm_get_concentration(spm_ctx *ctx, int species, double *host);
/* Total_solute[0 .. n_spec) in and out. */
int spm_put_total_solute(spm_ctx *ctx, const double *total);
int spm_get_total_solute(spm_ctx *ctx, double *total);
/* NSsolute_solver_Euler (fluid_solver.cxx:306-326): explicit Euler step of
 * zeta and the ion spectra with Rhs_NS_Nernst_Planck (:243-280). */
int spm_nssolute_solver_euler(spm_ctx *ctx);
/* Rescale_solute (solute_rhs.cxx:97-126): ion spectra rescaled to
 * Total_solute over the fluid volume. */
int spm_rescale_solute(spm_ctx *ctx);
/* Make_Coulomb_force_x_on_fluid (operate_electrolyte.cxx:90-177): the
 * Coulomb body force on the fluid into the f field (real space). */
int spm_make_coulomb_force(spm_ctx *ctx);

/* ---- Lees-Edwards shear in oblique coordinates ----------------------------
 * SW_EQ == Shear_Navier_Stokes_Lees_Edwards with free spherical particles
 * (or none): spm_step runs Time_evolution_hydro_OBL (sp_3d_ns.cxx:214-390).  The grid shears
 * with the flow; degree_oblique (the strain since the last remap, a global
 * of the reference, input.cxx:271) grows by Shear_rate * dt each step and
 * the grid is remapped when it reaches 1.  Shear_rate_eff (sp_3d_ns.h:
 * 174-196) is measured from the velocity at the end of each step and drives
 * the next step's shear term.  zeta holds the oblique components
 * (U_k2zeta_k_OBL, fft_wrapper.cxx:341-406) and uk_dc the contravariant
 * mean flow.  FP64, literal schedule. */
/* Switches the context to the sheared equations; Shear_rate is
 * constitutive_eq.Shear_Navier_Stokes_Lees_Edwards.External_field.DC.Shear_rate
 * (input.cxx:600-646).  Call once, after spm_create. */
int spm_set_lees_edwards(spm_ctx *ctx, double shear_rate);
/* degree_oblique and Shear_rate_eff in and out (the restart's
 * resume.CONTINUE.Saved_Data.oblique, resume.cxx:335-339). */
int spm_put_oblique(spm_ctx *ctx, double degree_oblique, double shear_rate_eff);
int spm_get_oblique(spm_ctx *ctx, double *degree_oblique, double *shear_rate_eff);
/* NS_solver_slavedEuler_Shear_OBL (fluid_solver.cxx:398-429) with
 * Zeta_k2advection_k_OBL (operate_omega.cxx:115-173). */
int spm_ns_solver_slaved_euler_shear_obl(spm_ctx *ctx);
/* Zeta_k2u_k_OBL (fft_wrapper.cxx:513-584): the contravariant velocity
 * spectrum of zeta into u0..u2. */
int spm_zeta_k2u_k_obl(spm_ctx *ctx);
/* Solenoidal_uk_OBL (operate_omega.cxx:222-262), contra2co (fft_wrapper.h:
 * 327-361) and U_k2zeta_k_OBL of the contravariant spectrum in u0..u2 into
 * zeta and uk_dc (u0..u2 are left holding the covariant spectrum). */
int spm_u_k2zeta_k_obl(spm_ctx *ctx);
/* The particle routines of the sheared step, on the context's particles
 * with the current degree_oblique and Shear_rate_eff; the Cartesian fluid
 * velocity they read is u0..u2 (the step passes U_oblique2u of u).
 * Make_phi_u_particle_OBL (make_phi.cxx:299-385, after Reset_phi_u) into
 * phi, up0..up2. */
int spm_make_phi_u_particle_obl(spm_ctx *ctx);
/* Calc_f_hydro_correct_precision_OBL (md_force.cxx:391-470, 497-506; the
 * Hydro_force stress fields are not formed). */
int spm_calc_f_hydro_obl(spm_ctx *ctx);
/* MD_solver_position_Euler_OBL (ts == 0) / _AB2_OBL with PBC_OBL
 * (particle_solver.cxx:292-346, periodic_boundary.h:22-44). */
int spm_md_solver_position_obl(spm_ctx *ctx);
/* MD_solver_velocity_Euler_OBL (ts == 0) / _AB2_hydro_OBL with Force_OBL:
 * all-pairs LJ with Distance0_OBL and gravity (particle_solver.cxx:348-443,
 * particle_solver.h:111-127, md_force.cxx:136-179). */
int spm_md_solver_velocity_obl(spm_ctx *ctx);

/* Mean_shear_stress(SHOW, ...) of a sheared context (sp_3d_ns.h:209-358,
 * the Lees-Edwards DC line, printed by the reference after every step,
 * sp_3d_ns.cxx:598-601): the stresses of the last step from the sums of its
 * Hydro_force and Hydro_force_new fields (md_force.cxx:432-620 summed as
 * Calc_hydro_stress does, fluid_solver.h:195-221) and the LJ pair virial
 * of its Force_OBL (particle_solver.h:111-127).  Without particles the
 * stresses are zero and the viscosity is ETA; with Shear_rate 0 the
 * viscosity divides by zero as the reference's does. */
typedef struct spm_le_stress {
  double time;             /* jikan.time                                  */
  double shear_rate;       /* Shear_rate (the imposed rate)               */
  double strain;           /* Shear_strain_realized                       */
  double lj_dev_stress;    /* dev_shear_stress_lj                         */
  double hydro_stress;     /* hydro_stress[1][0] (Hydro_force)            */
  double hydro_stress_new; /* hydro_stress_new[1][0] (Hydro_force_new)    */
  double viscosity;        /* (hydro_stress_new + lj_dev_stress) / rate + ETA */
} spm_le_stress;
int spm_le_shear_stress(spm_ctx *ctx, spm_le_stress *out);

#ifdef __cplusplus
}
#endif

#endif /* SPM_H */
