// spm_api.cpp -- the C ABI of include/spm.h: context lifetime, state
// transfer in the reference layouts, and the solver call surface that
// sequences the HIP kernels exactly as Time_evolution_hydro does
// (sp_3d_ns.cxx:61-212).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "mt19937.h"
#include "spm_internal.h"
#include "../../include/spm_closure.h"

using namespace spm;

namespace spm {

static size_t next_event(spm_ctx *c) {
  if (c->ev_used == c->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return (size_t)-1;
    c->ev_pool.push_back(e);
  }
  return c->ev_used++;
}

void tbeg(spm_ctx *c) {
  if (!c->timing) return;
  const size_t e = next_event(c);
  if (e == (size_t)-1) return;
  (void)hipEventRecord(c->ev_pool[e], c->stream);
  c->launches.push_back({-1, e, e, 0.0});
}

void tend(spm_ctx *c, int family, double bytes) {
  if (!c->timing || c->launches.empty() || c->launches.back().family != -1) return;
  const size_t e = next_event(c);
  if (e == (size_t)-1) return;
  (void)hipEventRecord(c->ev_pool[e], c->stream);
  c->launches.back().family = family;
  c->launches.back().e1 = e;
  c->launches.back().bytes = bytes;
}

// Folds the recorded launches into per-family totals and frees the events
// for reuse (the stream must be idle).
static int fold_timing(spm_ctx *c) {
  for (const auto &l : c->launches) {
    if (l.family < 0) continue;
    float ms = 0.f;
    SPM_HIP(c, hipEventElapsedTime(&ms, c->ev_pool[l.e0], c->ev_pool[l.e1]));
    c->fam_ms[l.family] += ms;
    c->fam_bytes[l.family] += l.bytes;
    c->fam_count[l.family] += 1;
  }
  c->launches.clear();
  c->ev_used = 0;
  return 0;
}

}  // namespace spm

namespace {

// The thermal-noise stream (MT19937 + polar Gaussian pairs, mt19937.h) is
// drawn on the host: it is sequential with rejection and uses the host
// libm's log/sqrt, so drawing it here keeps it identical to the reference's.
static_assert(SPM_MT_N == spm::kMtN, "the state size of include/spm.h");
void mt_seed(spm_ctx *c, unsigned long s) { spm::mt_init(c->mt, c->mti, (uint32_t)(s & 0xffffffffUL)); }

void gauss2(spm_ctx *c, double *out) { spm::mt_gauss2(c->mt, c->mti, out); }

bool is_pow2(int n, int lo, int hi) { return n >= lo && n <= hi && (n & (n - 1)) == 0; }
// the fused passes and the sheared solver take power-of-two grids; others
// run the literal schedule
bool pow2_grid(const spm_ctx *c) { return grid_pow2(c->grid.nx) && grid_pow2(c->grid.ny) && grid_pow2(c->grid.nz); }
// The fused x passes address each field through a buffer resource, whose
// byte counts and offsets are 32 bits: fields under 4 GiB.
bool fused_fits(const spm_ctx *c) {
  const size_t elem = c->prm.precision == SPM_FP32 ? sizeof(float) : sizeof(double);
  return c->nfield * elem < ((size_t)1 << 32) - 64;
}

double *field_ptr(spm_ctx *c, int f) {
  switch (f) {
    case SPM_FIELD_ZETA0: return c->zeta[0];
    case SPM_FIELD_ZETA1: return c->zeta[1];
    case SPM_FIELD_U0: return c->u[0];
    case SPM_FIELD_U1: return c->u[1];
    case SPM_FIELD_U2: return c->u[2];
    case SPM_FIELD_UP0: return c->up[0];
    case SPM_FIELD_UP1: return c->up[1];
    case SPM_FIELD_UP2: return c->up[2];
    case SPM_FIELD_PHI: return c->phi;
    case SPM_FIELD_F0: return c->f[0];
    case SPM_FIELD_F1: return c->f[1];
    case SPM_FIELD_F2: return c->f[2];
    default: return nullptr;
  }
}

int fix_mask(const spm_ctx *c) {
  return (c->prm.fix_cell[0] ? 1 : 0) | (c->prm.fix_cell[1] ? 2 : 0) | (c->prm.fix_cell[2] ? 4 : 0);
}

#define RC(x)                  \
  do {                         \
    int rc_ = (x);             \
    if (rc_) return rc_;       \
  } while (0)

// ---------------------------------------------------------------------
// the step pieces
// ---------------------------------------------------------------------
// NS_solver_slavedEuler (fluid_solver.cxx:328-354) with
// Zeta_k2advection_k (operate_omega.cxx:102-113).
int ns_solver(spm_ctx *c) {
  double *const u3[3] = {c->u[0], c->u[1], c->u[2]};
  double *const p5[5] = {c->u[0], c->u[1], c->u[2], c->work[0], c->work[1]};
  RC(k_truncate(c, c->zeta[0]));
  RC(k_truncate(c, c->zeta[1]));
  RC(k_zeta_k2u_k(c, c->zeta[0], c->zeta[1], c->u[0], c->u[1], c->u[2], 0));
  RC(fft_c2r_many(c, u3, 3));
  RC(k_products(c, c->u[0], c->u[1], c->u[2], c->work[0], c->work[1]));
  RC(fft_r2c_many(c, p5, 5));
  RC(k_advection_contract(c, c->u[0], c->u[1], c->u[2], c->work[0], c->work[1], c->work[0], c->work[1]));
  const double nu = c->prm.eta * (1. / c->prm.rho);
  RC(k_slaved_euler(c, c->zeta[0], c->zeta[1], c->work[0], c->work[1], nu, c->prm.dt));
  return 0;
}

// sp_3d_ns.cxx:67-79.
int zeta_k2u(spm_ctx *c, int mask) {
  double *const u3[3] = {c->u[0], c->u[1], c->u[2]};
  RC(k_truncate(c, c->zeta[0]));
  RC(k_truncate(c, c->zeta[1]));
  RC(k_zeta_k2u_k(c, c->zeta[0], c->zeta[1], c->u[0], c->u[1], c->u[2], mask));
  RC(fft_c2r_many(c, u3, 3));
  return 0;
}

// Make_f_particle_dt_sole + Add_f_particle (f_particle.h:71-99).
int make_f_sole_add(spm_ctx *c) {
  double *const f3[3] = {c->f[0], c->f[1], c->f[2]};
  RC(k_make_f(c));
  RC(fft_r2c_many(c, f3, 3));
  RC(k_solenoidal_trunc(c, c->f[0], c->f[1], c->f[2]));
  RC(fft_c2r_many(c, f3, 3));
  RC(k_add_f(c));
  return 0;
}

// U2zeta_k (operate_omega.h:186-189).
int u2zeta_k(spm_ctx *c) {
  double *const u3[3] = {c->u[0], c->u[1], c->u[2]};
  RC(fft_r2c_many(c, u3, 3));
  RC(k_u_k2zeta_k(c, c->u[0], c->u[1], c->u[2], c->zeta[0], c->zeta[1]));
  return 0;
}

int reset_phi_u(spm_ctx *c) {
  RC(k_zero(c, c->phi, c->nfield));
  for (int d = 0; d < 3; ++d) RC(k_zero(c, c->up[d], c->nfield));
  return 0;
}

// Add_random_force_thermostat (fluct.cxx:10-36): draw 6 normals per
// particle on the host, ship them, apply on the device.  Two pinned
// staging slots let the host draw step n+1 while step n is in flight.
int thermostat(spm_ctx *c) {
  if (!(c->prm.kbt > 0.) || c->np == 0) return 0;
  if (c->pool_mode) {  // drawn ahead into the pool, in the same order
    RC(p_noise_from_pool(c));
    return p_thermostat_apply(c, c->noise_dev);
  }
  const int s = c->noise_slot;
  SPM_HIP(c, hipEventSynchronize(c->noise_done[s]));
  double *h = c->noise_host[s];
  for (int n = 0; n < c->np; ++n) {
    gauss2(c, h + 6 * n);
    gauss2(c, h + 6 * n + 2);
    gauss2(c, h + 6 * n + 4);
  }
  SPM_HIP(c, hipMemcpyAsync(c->noise_dev, h, sizeof(double) * 6 * c->np, hipMemcpyHostToDevice, c->stream));
  SPM_HIP(c, hipEventRecord(c->noise_done[s], c->stream));
  c->noise_slot ^= 1;
  return p_thermostat_apply(c, c->noise_dev);
}

// The particle part of the step (sp_3d_ns.cxx:81-95,183-185): positions,
// hydrodynamic force and torque (plus the attached closure's correction),
// velocities, thermostat.  With prescribed motion the particles keep their
// positions and velocities and only the force is formed.
int particle_motion(spm_ctx *c) {
  const bool free_motion = c->motion == SPM_MOTION_FREE;
  if (free_motion) RC(p_position(c, c->ts == 0));
  RC(p_calc_f_hydro(c));
  if (c->closure) RC(closure_apply(c));
  if (free_motion) {
    RC(p_velocity(c, c->ts == 0));
    RC(thermostat(c));
  }
  return 0;
}

// Time_evolution_hydro (sp_3d_ns.cxx:61-212), Navier_Stokes with free
// spherical particles, no slip.
int time_evolution_hydro(spm_ctx *c) {
  RC(ns_solver(c));
  RC(zeta_k2u(c, fix_mask(c)));
  RC(particle_motion(c));
  RC(p_make_phi(c, 1));
  RC(make_f_sole_add(c));
  RC(u2zeta_k(c));
  return 0;
}

// f = up - phi u (make_phi mode f) and its z transform
static int make_f_z(spm_ctx *c, double *const f3[3]) {
  RC(p_make_f(c));
  return fft_pass_z_r2c(c, f3, 3);
}

// The same step as fused FFT passes (fused.hip).  Leaves zeta, uk_dc and the
// particles as the literal schedule does; u (a spectrum at the end of a
// step), phi and up are rebuilt by materialize() when a spm_step call
// returns; f is scratch.
// The particle work of a free-motion step that reads only the previous
// step's velocities (MD_solver_position, the tile lists at the new
// positions, the LJ cell lists and forces), started on the context's second
// stream after everything queued on its stream so far.  The fluid passes
// that follow on the stream touch no particle data; fused_step joins the
// two before Calc_f_hydro.
static int particle_head_side(spm_ctx *c) {
  SPM_HIP(c, hipEventRecord(c->ev_fork, c->stream));
  SPM_HIP(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
  hipStream_t main = c->stream;
  c->stream = c->side;  // the kernels and their timing events go to the side stream
  int rc = p_position(c, c->ts == 0);
  if (!rc) rc = p_build_tiles(c);
  if (!rc) rc = p_force(c);
  c->stream = main;
  if (rc) return rc;
  SPM_HIP(c, hipEventRecord(c->ev_join, c->side));
  return 0;
}

// the rest of particle_motion once particle_head_side's work has joined
static int particle_tail(spm_ctx *c) {
  SPM_HIP(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
  RC(p_calc_f_hydro_listed(c));
  if (c->closure) RC(closure_apply(c));
  RC(p_velocity_forced(c, c->ts == 0));
  return thermostat(c);
}

int fused_step(spm_ctx *c) {
  double *const u3[3] = {c->u[0], c->u[1], c->u[2]};
  double *const p5[5] = {c->u[0], c->u[1], c->u[2], c->work[0], c->work[1]};
  double *const f3[3] = {c->f[0], c->f[1], c->f[2]};
  const double nu = c->prm.eta * (1. / c->prm.rho);
  const bool overlap = c->motion == SPM_MOTION_FREE && c->np > 0;
  if (overlap) RC(particle_head_side(c));
  // NS_solver_slavedEuler
  RC(fx_zeta2u(c, 0));
  RC(fft_pass_y(c, u3, 3, -1));
  RC(fz_advect(c));
  RC(fft_pass_y(c, p5, 5, +1));
  // FIX_CELL, truncation, Zeta_k2u (the solver's last x pass and the first
  // of Zeta_k2u share a launch)
  RC(fx_update_zeta2u(c, nu, c->prm.dt, fix_mask(c)));
  RC(fft_pass_y(c, u3, 3, -1));
  RC(fft_pass_z_c2r(c, u3, 3));
  // particles (p_calc_f_hydro also builds the tile lists at the new positions)
  RC(overlap ? particle_tail(c) : particle_motion(c));
  // Make_phi_u_particle + Make_f_particle_dt_sole + Add_f_particle + U2zeta_k
  // (with no particles the tile lists stay empty and f = 0)
  RC(make_f_z(c, f3));
  RC(fft_pass_y(c, f3, 3, +1));
  RC(fx_project(c));
  return 0;
}

// u, phi and up as the literal schedule leaves them.  The reference's step
// ends with U2zeta_k, which leaves u transformed in place: u holds the
// spectrum of u + f, i.e. Zeta_k2u_k of the new zeta (the field is
// solenoidal and band-limited, so the two agree to rounding).
int materialize(spm_ctx *c) {
  RC(k_zeta_k2u_k(c, c->zeta[0], c->zeta[1], c->u[0], c->u[1], c->u[2], 0));
  RC(p_make_phi(c, 1));
  return 0;
}

// One stage of fused_step, for stage-by-stage tests against the literal
// routines: 1 NS_solver_slavedEuler, 2 FIX_CELL + Zeta_k2u, 3 the particle
// coupling tail (Make_phi_u_particle .. U2zeta_k, u and particles given),
// 4 materialize.
int fused_stage(spm_ctx *c, int stage) {
  double *const u3[3] = {c->u[0], c->u[1], c->u[2]};
  double *const p5[5] = {c->u[0], c->u[1], c->u[2], c->work[0], c->work[1]};
  double *const f3[3] = {c->f[0], c->f[1], c->f[2]};
  const double nu = c->prm.eta * (1. / c->prm.rho);
  if (stage == 1 || stage == 3) {  // as at the start of a fused step_n call
    RC(k_truncate(c, c->zeta[0]));
    RC(k_truncate(c, c->zeta[1]));
  }
  switch (stage) {
    case 1:
      RC(fx_zeta2u(c, 0));
      RC(fft_pass_y(c, u3, 3, -1));
      RC(fz_advect(c));
      RC(fft_pass_y(c, p5, 5, +1));
      RC(fx_update(c, nu, c->prm.dt));
      return 0;
    case 2:
      RC(fx_zeta2u(c, fix_mask(c)));
      RC(fft_pass_y(c, u3, 3, -1));
      RC(fft_pass_z_c2r(c, u3, 3));
      return 0;
    case 3:
      RC(p_build_tiles(c));
      RC(make_f_z(c, f3));
      RC(fft_pass_y(c, f3, 3, +1));
      RC(fx_project(c));
      return 0;
    case 4:
      return materialize(c);
  }
  return SPM_ERR_INVALID;
}

void free_ctx(spm_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (double *p : {c->zeta[0], c->zeta[1], c->u[0], c->u[1], c->u[2], c->up[0], c->up[1], c->up[2], c->phi, c->f[0],
                    c->f[1], c->f[2], c->work[0], c->work[1], c->work[2], c->work[3], c->work[4], c->uk_dc, c->pbuf,
                    c->pres, c->noise_dev, c->pool_dev})
    if (p) (void)hipFree(p);
  if (c->pool_counter) (void)hipFree(c->pool_counter);
  if (c->pool_host) (void)hipHostFree(c->pool_host);
  for (double *p : {c->conc[0], c->conc[1], c->conc_rhs[0], c->conc_rhs[1], c->surface_normal[0],
                    c->surface_normal[1], c->surface_normal[2], c->f_ns1[0], c->f_ns1[1], c->red, c->le_red,
                    c->le_stress, c->le_strain, c->el_q, c->el_p, c->el_pq, c->el_g[0], c->el_g[1], c->el_g[2]})
    if (p) (void)hipFree(p);
  for (int *p : {c->spec, c->stencil, c->pxint, c->tile_count, c->tile_start, c->tile_cursor, c->tile_list,
                 c->cell_count, c->cell_start, c->cell_cursor, c->cell_list, c->pcell, c->flag, c->tile_count_obl,
                 c->tile_start_obl, c->tile_cursor_obl, c->tile_list_obl})
    if (p) (void)hipFree(p);
  for (double2 *p : {c->tw_x, c->tw_y, c->tw_z, c->tw_rz})
    if (p) (void)hipFree(p);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  for (int s = 0; s < 2; ++s) {
    if (c->noise_host[s]) (void)hipHostFree(c->noise_host[s]);
    if (c->noise_done[s]) (void)hipEventDestroy(c->noise_done[s]);
  }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

template <typename P>
int dalloc(spm_ctx *c, P **p, size_t count) {
  SPM_HIP(c, hipMalloc((void **)p, sizeof(P) * (count > 0 ? count : 1)));
  SPM_HIP(c, hipMemsetAsync(*p, 0, sizeof(P) * (count > 0 ? count : 1), c->stream));
  return 0;
}

// Row pitch (elements) of the fields on the device.  The reference layout
// pads each z row to 2 (NZ/2 + 1) elements (2064 bytes at NZ = 256 in FP64),
// so consecutive rows start 16 bytes further from a 128-byte line and every
// row segment of a y-pass tile straddles two lines.  Measured on the MI355X
// (256^3, FP64, load + store through LDS), 128-byte aligned rows move the
// y-pass tile pattern at 5.2 TB/s and the x-pass pattern at 4.7 TB/s against
// 3.0 and 3.1 TB/s with the reference pitch.  Rows of 512 bytes or more
// start on 128-byte lines (on 64-byte lines where 128 would add more than an
// eighth to the field); shorter rows keep the reference pitch.
// SPM_ROW_ALIGN=<bytes> in the environment overrides the alignment (16 or
// less gives the reference pitch).
int row_pitch(int nz, size_t elem) {
  const long used = 2L * (nz / 2 + 1);
  long align = 0;
  if (const char *e = getenv("SPM_ROW_ALIGN")) {
    align = atol(e);
  } else if (used * (long)elem >= 512) {
    align = 128;
    const long padded = (used * (long)elem + 127) / 128 * 128;
    if (padded * 8 > used * (long)elem * 9) align = 64;
  }
  long a = align / (long)elem;
  if (a < 2) return (int)used;
  a += a & 1;  // whole complex elements
  return (int)((used + a - 1) / a * a);
}

int setup(spm_ctx *c) {
  const spm_params &p = c->prm;
  {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess && n > 0)
      c->n_cu = n;
  }
  Grid &g = c->grid;
  g.nx = p.nx;
  g.ny = p.ny;
  g.nz = p.nz;
  g.hnz = p.nz / 2 + 1;
  g.nz_ = row_pitch(p.nz, p.precision == SPM_FP32 ? sizeof(float) : sizeof(double));
  g.hp = g.nz_ / 2;
  g.trn_x = (p.nx + 2) / 3;
  g.trn_y = (p.ny + 2) / 3;
  g.trn_z = (p.nz + 2) / 3;
  g.dx = p.dx;
  g.lx = p.nx * p.dx;
  g.ly = p.ny * p.dx;
  g.lz = p.nz * p.dx;
  const double PI2 = 2.0 * M_PI;
  g.wave_x = PI2 / g.lx;
  g.wave_y = PI2 / g.ly;
  g.wave_z = PI2 / g.lz;
  c->nfield = (size_t)g.nx * g.ny * g.nz_;
  if (!pow2_grid(c) || !fused_fits(c)) c->schedule = SPM_SCHEDULE_LITERAL;

  PModel &m = c->pm;
  m.np = p.n_particles;
  m.rotation = p.rotation ? 1 : 0;
  m.dx = p.dx;
  m.radius = p.a * p.dx;
  m.hxi = (p.a_xi * p.dx) * 0.5;
  m.max_radius = m.radius + m.hxi + 1.7321 * p.dx;
  m.nh = (int)p.a + (int)p.a_xi + 2;
  m.rho = p.rho;
  m.dx3 = p.dx * p.dx * p.dx;
  const double R = m.radius;
  for (int s = 0; s < SPM_MAX_SPECIES; ++s) {
    const double mr = s < p.n_species ? p.mass_ratio[s] : 1.0;
    const double mass = (mr * p.rho) * 4. / 3. * M_PI * (R * R * R);
    const double moi = 2. / 5. * mass * (R * R);
    m.mass_ratio[s] = mr;
    m.imass[s] = 1. / mass;
    m.imoi[s] = 1. / moi;
  }
  m.lj_dia = 2.0 * R;  // SIGMA (input.cxx:1340, 400-403)
  double a_r_cutoff = 0.0;
  if (p.lj_truncate > 0)
    a_r_cutoff = p.lj_powers == 0 ? std::pow(2., 1. / 6.) : p.lj_powers == 1 ? std::pow(2., 1. / 12.)
                                                                           : std::pow(2., 1. / 18.);
  else if (p.lj_truncate == 0) {
    const int nmax = std::max(p.nx, std::max(p.ny, p.nz));  // Nmin as set at input.cxx:1406
    a_r_cutoff = std::min(nmax * p.dx * .5 / m.lj_dia, 2.5);
  }
  m.lj_cutoff = a_r_cutoff * m.lj_dia;
  c->a_r_cutoff = a_r_cutoff;
  m.lj_on = (p.lj_truncate >= 0 && m.lj_cutoff > 0.0) ? 1 : 0;
  m.epsilon = p.epsilon;
  m.lj_powers = p.lj_powers;
  m.g_on_fluid = p.g * p.rho * 4. / 3. * M_PI * (R * R) * R;
  m.g_direction = p.g_direction;
  c->np = m.np;

  // thermostat amplitudes (fluct.cxx:11-18)
  const double zeta_drag = 6. * M_PI * p.eta * R;
  const double zeta_drag_rot = 8. * M_PI * p.eta * (R * R * R);
  c->noise_v = 1.439 * std::sqrt(zeta_drag * p.dt * p.kbt * p.alpha_v);
  c->noise_o = 1.22 * std::sqrt(zeta_drag_rot * p.dt * p.kbt * p.alpha_o);

  SPM_HIP(c, hipSetDevice(c->device));
  SPM_HIP(c, hipStreamCreate(&c->stream));
  SPM_HIP(c, hipStreamCreate(&c->side));
  SPM_HIP(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  SPM_HIP(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));

  for (int d = 0; d < 2; ++d) RC(dalloc(c, &c->zeta[d], c->nfield));
  for (int d = 0; d < 3; ++d) {
    RC(dalloc(c, &c->u[d], c->nfield));
    RC(dalloc(c, &c->up[d], c->nfield));
    RC(dalloc(c, &c->f[d], c->nfield));
  }
  RC(dalloc(c, &c->phi, c->nfield));
  for (int d = 0; d < 2; ++d) RC(dalloc(c, &c->work[d], c->nfield));
  RC(dalloc(c, &c->uk_dc, 3));
  RC(dalloc(c, &c->flag, 1));
  RC(fft_init_tables(c));

  // Particle_domain (profile.cxx:25-74)
  {
    std::vector<int> st;
    const int nh = m.nh;
    for (int i = -nh; i <= nh; i++)
      for (int j = -nh; j <= nh; j++)
        for (int k = -nh; k <= nh; k++) {
          const double dmy = std::sqrt(((double)i * p.dx) * ((double)i * p.dx) + ((double)j * p.dx) * ((double)j * p.dx) +
                                       ((double)k * p.dx) * ((double)k * p.dx));
          if (dmy < m.max_radius) {
            st.push_back(i);
            st.push_back(j);
            st.push_back(k);
          }
        }
    c->n_stencil = (int)st.size() / 3;
    RC(dalloc(c, &c->stencil, st.size()));
    SPM_HIP(c, hipMemcpy(c->stencil, st.data(), sizeof(int) * st.size(), hipMemcpyHostToDevice));
  }

  // particles
  const int np = c->np;
  RC(dalloc(c, &c->pbuf, (size_t)P_NSLOTS * np));
  RC(dalloc(c, &c->spec, np));
  RC(dalloc(c, &c->pxint, 3 * (size_t)np));
  RC(dalloc(c, &c->pres, 3 * (size_t)np));
  RC(dalloc(c, &c->pcell, np));
  c->pstage.assign((size_t)P_NSLOTS * (np > 0 ? np : 1), 0.0);
  {
    std::vector<double> init((size_t)P_NSLOTS * np, 0.0);
    for (int n = 0; n < np; ++n) {
      init[(size_t)P_Q * np + n] = 1.0;
      init[(size_t)P_Q_OLD * np + n] = 1.0;
    }
    if (np) SPM_HIP(c, hipMemcpy(c->pbuf, init.data(), sizeof(double) * init.size(), hipMemcpyHostToDevice));
  }

  // tiles for make_phi
  c->tiles[0] = g.nx / kPhiTile;
  c->tiles[1] = g.ny / kPhiTile;
  c->tiles[2] = g.nz / kPhiTile;
  c->ntiles = c->tiles[0] * c->tiles[1] * c->tiles[2];
  RC(dalloc(c, &c->tile_count, c->ntiles));
  RC(dalloc(c, &c->tile_start, c->ntiles));
  RC(dalloc(c, &c->tile_cursor, c->ntiles));
  {
    long span = 1;
    for (int d = 0; d < 3; ++d) span *= std::min(c->tiles[d], (2 * m.nh + kPhiTile) / kPhiTile + 1);
    c->tile_list_cap = (int)std::max<long>(1, span * np);
    RC(dalloc(c, &c->tile_list, c->tile_list_cap));
  }

  // LJ cells: at least one cutoff wide
  if (m.lj_on) {
    const double L[3] = {g.lx, g.ly, g.lz};
    for (int d = 0; d < 3; ++d) {
      int nc = (int)std::floor(L[d] / m.lj_cutoff);
      c->cells[d] = std::max(1, std::min(nc, 128));
    }
  } else {
    c->cells[0] = c->cells[1] = c->cells[2] = 1;
  }
  c->ncells = c->cells[0] * c->cells[1] * c->cells[2];
  RC(dalloc(c, &c->cell_count, c->ncells));
  RC(dalloc(c, &c->cell_start, c->ncells));
  RC(dalloc(c, &c->cell_cursor, c->ncells));
  RC(dalloc(c, &c->cell_list, np));

  // thermostat staging
  for (int s = 0; s < 2; ++s) {
    SPM_HIP(c, hipHostMalloc((void **)&c->noise_host[s], sizeof(double) * 6 * (np > 0 ? np : 1), hipHostMallocDefault));
    SPM_HIP(c, hipEventCreateWithFlags(&c->noise_done[s], hipEventDisableTiming));
    SPM_HIP(c, hipEventRecord(c->noise_done[s], c->stream));
  }
  RC(dalloc(c, &c->noise_dev, 6 * (size_t)np));
  mt_seed(c, p.seed);
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

// ---------------------------------------------------------------------
// electrolyte (SW_EQ == Electrolyte), literal schedule.  Transport fields
// (ion spectra, fluxes, velocity, vorticity, phi, f) of the context's field
// type; the Poisson side (charge density, potential and its gradient, the
// particles' charge profile, surface normals, the sums) in FP64.  An FP64
// context keeps the Poisson fields in its own working fields, an FP32 one in
// the FP64 buffers el_* (every field buffer holds nfield doubles).
// ---------------------------------------------------------------------
inline bool el32(const spm_ctx *c) { return c->prm.precision == SPM_FP32; }
// result slots of c->red (after the partial sums)
enum { SLOT_SURF = 0, SLOT_SCALE0 = 1, SLOT_VOLUME = 3, SLOT_COUNT = 4, SLOT_FE = 6 };

// Make_phi_qq_particle (operate_electrolyte.cxx:179-242) into phi, surface.
int make_phi_qq(spm_ctx *c, double *phi, double *surface) {
  RC(p_make_phi_qq(c, phi, surface));
  if (c->np == 0) return 0;
  // sum |q| over the particles, in particle order (host copy of the species)
  double abs_total = 0.;
  for (int n = 0; n < c->np; ++n) abs_total += std::fabs(c->qtab.q[c->spec_host[n]]);
  RC(e_reduce(c, 0, surface, nullptr, SLOT_SURF, 1, abs_total));
  RC(e_scale_by(c, surface, (long)c->nfield, SLOT_SURF, true));
  return 0;
}

// A_k2a_out (fft_wrapper.h:639-651), fields of the context's type.
int k2a_out(spm_ctx *c, const double *a_k, double *a_x) {
  const size_t eb = el32(c) ? sizeof(float) : sizeof(double);
  SPM_HIP(c, hipMemcpyAsync(a_x, a_k, eb * c->nfield, hipMemcpyDeviceToDevice, c->stream));
  return fft_c2r(c, a_x);
}

// Conc_k2charge_field (operate_electrolyte.cxx:45-74): charge and phi_p
// FP64, dmy of the field type.
int conc_k2charge(spm_ctx *c, double *charge, double *phi_p, double *dmy) {
  RC(make_phi_qq(c, phi_p, charge));
  for (int n = 0; n < c->el.n_spec; ++n) {
    RC(k2a_out(c, c->conc[n], dmy));
    RC(e_charge_add(c, charge, dmy, phi_p, c->valency_e[n]));
  }
  return 0;
}

// Rhs_NS_Nernst_Planck + Rhs_NS_solute (fluid_solver.cxx:181-280) with the
// working memory NSsolute_solver_Euler passes: u, solute_flux = up,
// grad_potential = f, rhs_ns = f_ns1, rhs_solute = conc_rhs.
int rhs_nernst_planck(spm_ctx *c) {
  double *const u3[3] = {c->u[0], c->u[1], c->u[2]};
  double *const fl[3] = {c->up[0], c->up[1], c->up[2]};
  double *const gp[3] = {c->f[0], c->f[1], c->f[2]};
  double *const p5[5] = {c->u[0], c->u[1], c->u[2], c->work[0], c->work[1]};
  // potential gradient in real space (:263-274), FP64: in f (FP64
  // context) or el_g, then copied to f as the field type (FP32 context),
  // where the step's Calc_f_hydro reads it
  double *const g64[3] = {c->el_g[0], c->el_g[1], c->el_g[2]};
  double *const *grad = el32(c) ? g64 : gp;
  double *charge = el32(c) ? c->el_q : c->u[0];
  RC(conc_k2charge(c, charge, el32(c) ? c->el_pq : c->u[1], c->u[2]));
  RC(el32(c) ? fft_r2c_f64(c, charge) : fft_r2c(c, charge));
  double cst[3];
  {
    double amp = (double)((long)c->grid.nx * c->grid.ny * c->grid.nz);
    if (c->el.ac) amp = std::sin(c->el.angular_frequency * c->time);
    for (int d = 0; d < 3; ++d) cst[d] = c->el.e_ext[d] * amp;
  }
  RC(e_poisson_grad(c, charge, grad, 1, cst));
  if (el32(c)) {
    RC(fft_c2r_many_f64(c, grad, 3));
    RC(e_to_t(c, grad, gp));
  } else {
    RC(fft_c2r_many(c, grad, 3));
  }
  // Rhs_NS_solute (:181-227)
  RC(zeta_k2u(c, 0));
  RC(p_surface_normal(c));  // tile lists are current from make_phi_qq
  for (int n = 0; n < c->el.n_spec; ++n) {
    RC(k_truncate(c, c->conc[n]));
    RC(e_grad_scaled(c, c->conc[n], c->prm.kbt * c->el.onsager[n], fl));
    RC(fft_c2r_many(c, fl, 3));
    RC(k2a_out(c, c->conc[n], c->conc_rhs[n]));
    RC(e_np_flux(c, fl, c->conc_rhs[n], grad, c->surface_normal, u3, c->el.onsager[n], c->valency_e[n]));
    RC(fft_r2c_many(c, fl, 3));
    RC(e_divergence(c, fl, c->conc_rhs[n]));
  }
  // Rhs_solvent: U2advection_k (operate_omega.cxx:14-100), u as scratch
  RC(k_products(c, c->u[0], c->u[1], c->u[2], c->work[0], c->work[1]));
  RC(fft_r2c_many(c, p5, 5));
  RC(k_advection_contract(c, c->u[0], c->u[1], c->u[2], c->work[0], c->work[1], c->f_ns1[0], c->f_ns1[1]));
  return 0;
}

// NSsolute_solver_Euler (fluid_solver.cxx:306-326); uk_dc gains dt * 0.
int nssolute_solver(spm_ctx *c) {
  RC(rhs_nernst_planck(c));
  const double nu = c->prm.eta * (1. / c->prm.rho);
  return e_nssolute_update(c, nu, c->prm.dt);
}

// Rescale_solute (solute_rhs.cxx:97-126) with phi and up[0] as working memory.
int rescale_solute(spm_ctx *c) {
  RC(p_make_phi(c, 0));
  for (int n = 0; n < c->el.n_spec; ++n) {
    RC(k2a_out(c, c->conc[n], c->up[0]));
    RC(e_reduce(c, 1, c->up[0], c->phi, SLOT_SCALE0 + n, 1, c->total_solute[n]));
    RC(e_scale_by(c, c->conc[n], (long)c->nfield, SLOT_SCALE0 + n));
  }
  return 0;
}

// Make_Coulomb_force_x_on_fluid (operate_electrolyte.cxx:90-177), free
// particles, into f.  FP64 context: charge in up[0], potential in up[1],
// gradient in f; FP32 context: el_q, el_p, el_g.
int coulomb_force(spm_ctx *c) {
  double *const f3[3] = {c->f[0], c->f[1], c->f[2]};
  double *const g64[3] = {c->el_g[0], c->el_g[1], c->el_g[2]};
  double *const *grad = el32(c) ? g64 : f3;
  double *charge = el32(c) ? c->el_q : c->up[0], *pot = el32(c) ? c->el_p : c->up[1];
  RC(conc_k2charge(c, charge, el32(c) ? c->el_pq : c->f[0], c->f[1]));
  SPM_HIP(c, hipMemcpyAsync(pot, charge, sizeof(double) * c->nfield, hipMemcpyDeviceToDevice, c->stream));
  RC(el32(c) ? fft_r2c_f64(c, pot) : fft_r2c(c, pot));
  const double zero[3] = {0.0, 0.0, 0.0};
  RC(e_poisson_grad(c, pot, grad, 0, zero));
  RC(el32(c) ? fft_c2r_many_f64(c, grad, 3) : fft_c2r_many(c, grad, 3));
  double ext[3];
  for (int d = 0; d < 3; ++d) {
    ext[d] = c->el.e_ext[d];
    if (c->el.ac) ext[d] *= std::sin(c->el.angular_frequency * c->time);
  }
  return e_coulomb(c, charge, grad, f3, ext, c->el.external_field);
}

// Solenoidal_u (operate_omega.h:96-103) on u.
int solenoidal_u(spm_ctx *c) {
  double *const u3[3] = {c->u[0], c->u[1], c->u[2]};
  RC(fft_r2c_many(c, u3, 3));
  RC(k_solenoidal_trunc(c, c->u[0], c->u[1], c->u[2]));
  RC(fft_c2r_many(c, u3, 3));
  return 0;
}

// Time_evolution_hydro, SW_EQ == Electrolyte (sp_3d_ns.cxx:46-57,61-212).
int electrolyte_step_impl(spm_ctx *c) {
  double *const u3[3] = {c->u[0], c->u[1], c->u[2]};
  double *const f3[3] = {c->f[0], c->f[1], c->f[2]};
  RC(nssolute_solver(c));
  RC(zeta_k2u(c, fix_mask(c)));
  RC(p_position(c, c->ts == 0));
  RC(rescale_solute(c));
  // Calc_f_hydro on the f field (the potential gradient left by the
  // solver) into f_hydro1 / torque_hydro1 (:97-106)
  RC(p_calc_f_hydro_field(c, f3));
  RC(p_copy_hydro1(c));
  RC(coulomb_force(c));
  RC(e_add_scaled(c, u3, f3, c->prm.dt * (1. / c->prm.rho)));
  RC(solenoidal_u(c));
  RC(p_calc_f_hydro(c));
  RC(p_velocity(c, c->ts == 0));
  // no thermostat in the electrolyte path (:183)
  RC(p_make_phi(c, 1));
  RC(make_f_sole_add(c));
  RC(u2zeta_k(c));
  return 0;
}

// Init_rho_ion (operate_electrolyte.cxx:695-815), uniform initial
// distribution.  Reads the fluid volume and ion counts back to the host.
int init_rho_ion(spm_ctx *c) {
  const spm_electrolyte &E = c->el;
  double surface_ion_number = 0.;
  for (int s = 0; s < c->prm.n_species; ++s) {
    if (E.surface_charge[s] == 0.) {
      c->err = "Surface_charge must be nonzero for every particle species";
      return SPM_ERR_INVALID;
    }
    surface_ion_number -= E.surface_charge[s] * c->spec_count[s];
  }
  RC(p_make_phi(c, 0));
  RC(e_reduce(c, 2, nullptr, c->phi, SLOT_VOLUME, 0, 0.0));
  double volume = 0.;  // sum (1 - phi) dx^3
  RC(e_read_slot(c, SLOT_VOLUME, &volume));
  double dens[2] = {0.0, 0.0};
  if (E.n_spec == 1) {
    if (c->prm.n_species == 1 && E.valency[0] * E.surface_charge[0] >= 0.) {
      c->err = "Valency_counterion must have the opposite sign of Surface_charge";
      return SPM_ERR_INVALID;
    }
    dens[0] = (surface_ion_number / E.valency[0]) / volume;
  } else {
    if (E.valency[0] <= 0. || E.valency[1] >= 0.) {
      c->err = "Valency_positive_ion must be > 0 and Valency_negative_ion < 0";
      return SPM_ERR_INVALID;
    }
    double pos = 0., neg = 0.;
    for (int s = 0; s < c->prm.n_species; ++s) {
      const double q = E.surface_charge[s], npart = (double)c->spec_count[s];
      if (q < 0.)
        pos += std::fabs(q * npart / E.valency[0]);
      else if (q > 0.)
        neg += std::fabs(q * npart / E.valency[1]);
    }
    const double rho_inf = c->prm.kbt * E.dielectric / ((E.elementary_charge * E.debye_length) *
                                                        (E.elementary_charge * E.debye_length));
    const double rho_pos = rho_inf / (E.valency[0] * (E.valency[0] - E.valency[1]));
    const double rho_neg = -rho_inf / (E.valency[1] * (E.valency[0] - E.valency[1]));
    dens[0] = pos / volume + rho_pos;
    dens[1] = neg / volume + rho_neg;
  }
  for (int n = 0; n < E.n_spec; ++n) {
    RC(e_fill_real(c, c->conc[n], dens[n]));
    RC(fft_r2c(c, c->conc[n]));
  }
  // Count_solute_each (solute_rhs.h:143-156), phi and up[0] as working memory
  RC(p_make_phi(c, 0));
  for (int n = 0; n < E.n_spec; ++n) {
    RC(k2a_out(c, c->conc[n], c->up[0]));
    RC(e_reduce(c, 1, c->up[0], c->phi, SLOT_COUNT + n, 0, 0.0));
    RC(e_read_slot(c, SLOT_COUNT + n, &c->total_solute[n]));
  }
  return 0;
}

// Calc_free_energy_PB (operate_electrolyte.cxx:302-371): the ideal-gas
// term kBT sum (1 - phi) C (log C - 1) dx^3 over the species and the
// electrostatic term sum rho_e (Phi / 2 - E.x) dx^3, Phi the periodic
// Poisson potential of the charge density (phi, up[0..2] as working memory).
int free_energy_pb(spm_ctx *c, double *F) {
  double *charge = el32(c) ? c->el_q : c->up[1], *val = c->up[2];
  RC(p_make_phi(c, 0));
  for (int n = 0; n < c->el.n_spec; ++n) {
    RC(k2a_out(c, c->conc[n], val));
    RC(e_reduce(c, 3, val, c->phi, SLOT_FE + n, 0, 0.0));
  }
  RC(conc_k2charge(c, charge, el32(c) ? c->el_pq : c->up[0], val));
  // A2a_k_out, Charge_field_k2Coulomb_potential_k_PBC, A_k2a (FP64)
  if (el32(c)) val = c->el_p;
  SPM_HIP(c, hipMemcpyAsync(val, charge, sizeof(double) * c->nfield, hipMemcpyDeviceToDevice, c->stream));
  RC(el32(c) ? fft_r2c_f64(c, val) : fft_r2c(c, val));
  RC(e_coulomb_potential(c, val));
  RC(el32(c) ? fft_c2r_many_f64(c, &val, 1) : fft_c2r(c, val));
  double ext[3];
  for (int d = 0; d < 3; ++d) {
    ext[d] = c->el.e_ext[d];
    if (c->el.ac) ext[d] *= std::sin(c->el.angular_frequency * c->time);
  }
  RC(e_half_potential(c, val, ext, c->el.external_field));
  RC(e_reduce(c, 4, charge, val, SLOT_FE + 2, 0, 0.0));
  double fe[3] = {0.0, 0.0, 0.0};
  for (int q = 0; q < 3; ++q)
    if (q < c->el.n_spec || q == 2) RC(e_read_slot(c, SLOT_FE + q, &fe[q]));
  *F = (fe[0] + fe[1]) * c->prm.kbt + fe[2];  // e_reduce has applied dx^3
  return 0;
}

// Set_steadystate_ion_density (operate_electrolyte.cxx:374-412): with the
// fluid at rest (u = 0, zeta = 0, uk_dc = 0) Ion_diffusion_solver_Euler
// (fluid_solver.cxx:281-304: the Nernst-Planck right-hand side, the Euler
// update of the ion spectra on the kept modes, Rescale_solute) until the
// free energy changes by less than TOL = 1e-4 (operate_electrolyte.h:21),
// at most 9999 times.  The zero vorticity is held in f_ns1 (the solvent
// right-hand side the step does not need here, rewritten after zeta is
// read).
int steady_ion_density(spm_ctx *c, int *iterations) {
  double f_prev = 0.0;
  RC(free_energy_pb(c, &f_prev));
  double *dc0 = nullptr;
  SPM_HIP(c, hipMalloc((void **)&dc0, 3 * sizeof(double)));
  SPM_HIP(c, hipMemsetAsync(dc0, 0, 3 * sizeof(double), c->stream));
  double *const zsave[2] = {c->zeta[0], c->zeta[1]};
  double *const dcsave = c->uk_dc;
  const double nu = c->prm.eta * (1. / c->prm.rho);
  int m = 1, rc = 0;
  for (; m < 10000 && !rc; ++m) {
    rc = k_zero(c, c->f_ns1[0], c->nfield);
    if (!rc) rc = k_zero(c, c->f_ns1[1], c->nfield);
    c->zeta[0] = c->f_ns1[0];
    c->zeta[1] = c->f_ns1[1];
    c->uk_dc = dc0;
    if (!rc) rc = rhs_nernst_planck(c);
    if (!rc) rc = e_nssolute_update(c, nu, c->prm.dt, false);
    c->zeta[0] = zsave[0];
    c->zeta[1] = zsave[1];
    c->uk_dc = dcsave;
    if (!rc) rc = rescale_solute(c);
    double f = 0.0;
    if (!rc) rc = free_energy_pb(c, &f);
    if (!rc && std::fabs(f - f_prev) < 1e-4) break;
    f_prev = f;
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(dc0);
  if (iterations) *iterations = m < 10000 ? m : 9999;
  return rc;
}

// FP32 transport: the relaxation stops when the free energy changes by less
// than 1e-4, a change that FP64 sums over FP32 ion fields resolve only to
// their rounding (the test then ends at a different, noise-chosen
// iteration).  So an FP32 context relaxes the ions in an FP64 twin built
// from the same parameters, particles and time, and takes the relaxed
// spectra rounded to FP32; its own Count_solute_each follows.
int pb_relax_in_f64(spm_ctx *c, int *iterations) {
  spm_params p = c->prm;
  p.precision = SPM_FP64;
  spm_ctx *t = nullptr;
  if (int rc = spm_create(&p, &t)) {
    c->err = std::string("FP64 context for the Poisson-Boltzmann start: ") + spm_last_error(nullptr);
    return rc;
  }
  int rc = 0;
  std::vector<spm_particle> pp(c->np > 0 ? c->np : 1);
  std::vector<double> host(ref_count(c->grid));
  double total[2] = {0.0, 0.0};
  if (!rc) rc = spm_get_particles(c, pp.data(), c->np);
  if (!rc) rc = spm_put_particles(t, pp.data(), c->np);
  if (!rc) rc = spm_put_time(t, c->ts, c->time);
  if (!rc) rc = spm_init_fluid(t);
  if (!rc) rc = spm_set_electrolyte(t, &c->el);
  if (!rc) rc = spm_init_rho_ion_pb(t, iterations);
  for (int n = 0; n < c->el.n_spec && !rc; ++n) {
    rc = spm_get_concentration(t, n, host.data());
    if (!rc) rc = spm_put_concentration(c, n, host.data());
  }
  if (!rc) rc = spm_get_total_solute(t, total);
  if (rc && c->err.empty()) c->err = std::string("FP64 twin of the Poisson-Boltzmann start: ") + spm_last_error(t);
  spm_destroy(t);
  if (!rc)
    for (int n = 0; n < 2; ++n) c->total_solute[n] = total[n];
  return rc;
}

int init_rho_ion_pb(spm_ctx *c, int *iterations) {
  if (c->el.n_spec != 2 || c->np < 2) {
    c->err = "the Poisson-Boltzmann start is built for salt with more than one particle (uniform start, then "
             "Set_steadystate_ion_density); the iterative profile of one particle or salt-free ions is not";
    return SPM_ERR_UNSUPPORTED;
  }
  if (el32(c)) {
    RC(pb_relax_in_f64(c, iterations));
  } else {
    RC(init_rho_ion(c));  // Set_uniform_ion_charge_density_salt, Total_solute
    RC(steady_ion_density(c, iterations));
  }
  // Count_solute_each (solute_rhs.h:143-156)
  RC(p_make_phi(c, 0));
  for (int n = 0; n < c->el.n_spec; ++n) {
    RC(k2a_out(c, c->conc[n], c->up[0]));
    RC(e_reduce(c, 1, c->up[0], c->phi, SLOT_COUNT + n, 0, 0.0));
    RC(e_read_slot(c, SLOT_COUNT + n, &c->total_solute[n]));
  }
  return 0;
}

int set_electrolyte(spm_ctx *c, const spm_electrolyte *el) {
  if (!el || el->n_spec < 1 || el->n_spec > 2 || !(el->dielectric > 0.) || !(el->elementary_charge > 0.)) {
    c->err = "invalid electrolyte parameters";
    return SPM_ERR_INVALID;
  }
  if (el->n_spec == 2 && !(el->debye_length > 0.)) {
    c->err = "Debye_length must be > 0 with salt";
    return SPM_ERR_INVALID;
  }
  if (c->electro) {
    c->err = "the context already runs the electrolyte equations";
    return SPM_ERR_INVALID;
  }
  c->el = *el;
  for (int n = 0; n < 2; ++n) {
    RC(dalloc(c, &c->conc[n], c->nfield));
    RC(dalloc(c, &c->conc_rhs[n], c->nfield));
    RC(dalloc(c, &c->f_ns1[n], c->nfield));
  }
  for (int d = 0; d < 3; ++d) RC(dalloc(c, &c->surface_normal[d], c->nfield));
  if (el32(c)) {  // the FP64 Poisson side of FP32 transport
    RC(dalloc(c, &c->el_q, c->nfield));
    RC(dalloc(c, &c->el_p, c->nfield));
    RC(dalloc(c, &c->el_pq, c->nfield));
    for (int d = 0; d < 3; ++d) RC(dalloc(c, &c->el_g[d], c->nfield));
  }
  RC(dalloc(c, &c->red, (size_t)e_red_slots()));
  for (int s = 0; s < SPM_MAX_SPECIES; ++s) c->qtab.q[s] = el->surface_charge[s] * el->elementary_charge;
  for (int n = 0; n < 2; ++n) c->valency_e[n] = el->valency[n] * el->elementary_charge;
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  c->electro = true;
  return 0;
}

// SW_EQ == Shear_Navier_Stokes_Lees_Edwards with no particles
// (Time_evolution_hydro_OBL, sp_3d_ns.cxx:214-390); the shear rate comes
// from the UDF's constitutive_eq.Shear_Navier_Stokes_Lees_Edwards.
// External_field.DC.Shear_rate (input.cxx:600-646); degree_oblique is a
// zero-initialised global (input.cxx:271).
int set_lees_edwards(spm_ctx *c, double shear_rate) {
  // the strain grows and is remapped at 1 (sp_3d_ns.cxx:280-305): a
  // negative rate would take degree_oblique below 0, which the reference
  // never remaps
  if (!std::isfinite(shear_rate) || shear_rate < 0.0) {
    c->err = "the shear rate must be finite and >= 0";
    return SPM_ERR_INVALID;
  }
  // one remap per step keeps degree_oblique in [0, 1) only while the strain
  // of a step stays below 1
  if (!(shear_rate * c->prm.dt < 1.0)) {
    c->err = "Shear_rate * dt must be below 1 (one grid remap per step)";
    return SPM_ERR_INVALID;
  }
  if (c->le || c->electro) {
    c->err = "the context already runs another set of equations";
    return SPM_ERR_INVALID;
  }
  if (c->prm.precision != SPM_FP64) {
    c->err = "the Lees-Edwards path runs in FP64 only";
    return SPM_ERR_UNSUPPORTED;
  }
  if (!pow2_grid(c)) {
    c->err = "the Lees-Edwards path needs power-of-two grids";
    return SPM_ERR_UNSUPPORTED;
  }
  RC(dalloc(c, &c->le_red, (size_t)le_red_slots()));  // zeroes Shear_rate_eff
  RC(dalloc(c, &c->le_stress, (size_t)(kStressSlots + 1) * (c->np > 0 ? c->np : 1)));
  RC(dalloc(c, &c->le_strain, 1));  // Shear_strain_realized = 0 (input.cxx:612)
  // tile lists of the phi gather, images across the y faces included
  RC(dalloc(c, &c->tile_count_obl, (size_t)c->ntiles));
  RC(dalloc(c, &c->tile_start_obl, (size_t)c->ntiles));
  RC(dalloc(c, &c->tile_cursor_obl, (size_t)c->ntiles));
  c->tile_list_obl_cap = std::max<long>(1, 3L * c->np * obl_tile_span(c, 0) * obl_tile_span(c, 1) * obl_tile_span(c, 2));
  RC(dalloc(c, &c->tile_list_obl, (size_t)c->tile_list_obl_cap));
  // under shear the LJ diameter is min((SIGMA + XI) / 2^(1/p), SIGMA) with
  // the LJ_powers exponent p (input.cxx:389-403); R_cutoff = A_R_cutoff LJ_dia
  {
    PModel &m = c->pm;
    const double sigma = m.lj_dia, xi = c->prm.a_xi * c->prm.dx;
    const double dummy_pow = c->prm.lj_powers == 0 ? std::pow(2., 1. / 6.)
                             : c->prm.lj_powers == 1 ? std::pow(2., 1. / 12.) : std::pow(2., 1. / 18.);
    m.lj_dia = std::min((sigma + xi) / dummy_pow, sigma);
    m.lj_cutoff = c->a_r_cutoff * m.lj_dia;
  }
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  c->le_rate = shear_rate;
  c->le_deg = 0.0;
  c->le = true;
  return 0;
}

// degree_oblique lives on the host (it selects the remap), Shear_rate_eff on
// the device (the step measures it and the next step's advection reads it).
int put_oblique(spm_ctx *c, double deg, double sre) {
  if (!(deg >= 0.0 && deg < 1.0) || !std::isfinite(sre)) {
    c->err = "degree_oblique must lie in [0, 1) and Shear_rate_eff be finite";
    return SPM_ERR_INVALID;
  }
  SPM_HIP(c, hipMemcpyAsync(c->le_red + le_red_slots() - 2, &sre, sizeof(double), hipMemcpyHostToDevice, c->stream));
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  c->le_deg = deg;
  return 0;
}

int get_oblique(spm_ctx *c, double *deg, double *sre) {
  SPM_HIP(c, hipMemcpyAsync(sre, c->le_red + le_red_slots() - 2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  *deg = c->le_deg;
  return 0;
}

// Mean_shear_stress (sp_3d_ns.h:209-358) from the moments f_hydro_obl left
// per particle.  Per particle n (ry its points' unwrapped y, phi the
// profile, f0 the x force density, V = sum phi, It = 2/3 sum phi |r|^2):
//   sum Hydro_force = sum_n [rho_p sum f0 ry - (mdfr_x / V) sum ry phi]
//                     - (sum_n sum f0) / (sum_n V / RHO) sum_n sum ry phi
//   sum Hydro_force_new = sum_n [rho_p sum f0 ry - RHO (sum f0 / V) sum ry phi
//                     - RHO / It (T_y sum ry phi r_z - T_z sum ry phi r_y)]
// (md_force.cxx:497-617; the raw force and torque sums), each times
// -Ivolume DX^3 / dt (Calc_hydro_stress); the LJ part is the pair virial
// times Ivolume.
int le_shear_stress(spm_ctx *c, spm_le_stress *out) {
  const int np = c->np;
  std::vector<double> st((size_t)(kStressSlots + 1) * (np > 0 ? np : 1), 0.0);
  double strain = 0.0;
  if (np > 0)
    SPM_HIP(c, hipMemcpyAsync(st.data(), c->le_stress, sizeof(double) * st.size(), hipMemcpyDeviceToHost,
                              c->stream));
  SPM_HIP(c, hipMemcpyAsync(&strain, c->le_strain, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  const Grid &g = c->grid;
  const double rho = c->prm.rho;
  double hf = 0.0, hfn = 0.0, sum_force = 0.0, sum_volume = 0.0, sum_b = 0.0, vir = 0.0;
  for (int n = 0; n < np; ++n) {
    const double *o = &st[(size_t)kStressSlots * n];
    const double rhop = c->pm.mass_ratio[c->spec_host[n]] * rho;
    hf += rhop * o[0] - (o[9] / o[4]) * o[1];
    hfn += rhop * o[0] + (-rho * o[6] / o[4]) * o[1] + (-rho / ((2.0 / 3.0) * o[5])) * (o[7] * o[2] - o[8] * o[3]);
    sum_force += o[6];
    sum_volume += o[4];
    sum_b += o[1];
    vir += st[(size_t)kStressSlots * np + n];
  }
  if (np > 0) hf -= sum_force / (sum_volume / rho) * sum_b;
  const double ivol = 1.0 / (g.lx * g.ly * g.lz);
  const double dmy = ivol * (g.dx * g.dx * g.dx) / c->prm.dt;
  out->time = c->time;
  out->shear_rate = c->le_rate;
  out->strain = strain;
  out->lj_dev_stress = vir * ivol;
  out->hydro_stress = -hf * dmy;
  out->hydro_stress_new = -hfn * dmy;
  out->viscosity = (out->hydro_stress_new + out->lj_dev_stress) / c->le_rate + c->prm.eta;
  return 0;
}

// particle hooks: the Cartesian velocity is u0..u2, Shear_rate_eff the
// device value spm_put_oblique or the last step left
const double *le_sre(const spm_ctx *c) { return c->le_red + le_red_slots() - 2; }
int le_make_phi_api(spm_ctx *c) {
  RC(p_build_tiles(c));
  return p_make_phi_obl(c, le_sre(c));
}
int le_f_hydro_api(spm_ctx *c) {
  double *const u3[3] = {c->u[0], c->u[1], c->u[2]};
  return p_calc_f_hydro_obl(c, u3, le_sre(c));
}
int le_position_api(spm_ctx *c) { return p_position_obl(c, c->ts == 0, le_sre(c)); }
int le_velocity_api(spm_ctx *c) { return p_velocity_obl(c, c->ts == 0); }

int le_zeta_k2u_k_api(spm_ctx *c) {
  double *const u3[3] = {c->u[0], c->u[1], c->u[2]};
  return le_zeta2u_k(c, u3, 0);
}

}  // namespace

namespace spm {
// text of the last error of a call without a context (spm_create, the
// stateless transforms), per thread; spm_last_error(NULL) returns it
std::string &global_error() {
  static thread_local std::string s;
  return s;
}
int electrolyte_step(spm_ctx *c) { return electrolyte_step_impl(c); }

// Host arrays hold a field in the reference layout, rows of 2 (NZ/2 + 1)
// elements back to back; on the device rows are nz_ elements apart.  Queued
// on the context's stream (host memory must stay valid until it syncs).
int rows_to_device(spm_ctx *c, void *dev, const void *host, size_t elem) {
  const Grid &g = c->grid;
  const size_t w = 2 * (size_t)g.hnz * elem;
  if (g.nz_ == 2 * g.hnz)
    SPM_HIP(c, hipMemcpyAsync(dev, host, w * g.nx * g.ny, hipMemcpyHostToDevice, c->stream));
  else
    SPM_HIP(c, hipMemcpy2DAsync(dev, (size_t)g.nz_ * elem, host, w, w, (size_t)g.nx * g.ny, hipMemcpyHostToDevice,
                                c->stream));
  return 0;
}
int rows_to_host(spm_ctx *c, void *host, const void *dev, size_t elem) {
  const Grid &g = c->grid;
  const size_t w = 2 * (size_t)g.hnz * elem;
  if (g.nz_ == 2 * g.hnz)
    SPM_HIP(c, hipMemcpyAsync(host, dev, w * g.nx * g.ny, hipMemcpyDeviceToHost, c->stream));
  else
    SPM_HIP(c, hipMemcpy2DAsync(host, w, dev, (size_t)g.nz_ * elem, w, (size_t)g.nx * g.ny, hipMemcpyDeviceToHost,
                                c->stream));
  return 0;
}
int host_thermostat(spm_ctx *c) { return thermostat(c); }
}  // namespace spm

// =====================================================================
// C ABI
// =====================================================================
// ---------------------------------------------------------------------
// pieces of spm_step for ensembles (ensemble.cpp), which capture one step
// of every box into a graph and replay it
// ---------------------------------------------------------------------
namespace spm {

int settle(spm_ctx *c) {
  if (!c->stale) return 0;
  c->stale = false;
  return materialize(c);
}

// phase 0: what spm_step does before its steps; 1: one step; 2: what it
// does after them (for fused boxes: marking u, phi and up stale).  No host
// synchronisation on any phase.
int box_step_phase(spm_ctx *c, int phase) {
  const bool fused = c->schedule == SPM_SCHEDULE_FUSED;
  if (phase == 0) {
    if (fused) {
      RC(k_truncate(c, c->zeta[0]));
      RC(k_truncate(c, c->zeta[1]));
    }
    return 0;
  }
  if (phase == 1) return fused ? fused_step(c) : time_evolution_hydro(c);
  if (fused) c->stale = true;
  return 0;
}

bool box_uses_noise(const spm_ctx *c) {
  return c->prm.kbt > 0. && c->np > 0 && c->motion == SPM_MOTION_FREE;
}

// device pool and pinned staging for `steps` steps of noise
int pool_setup(spm_ctx *c, int steps) {
  if (!box_uses_noise(c) || c->pool_steps >= steps) return 0;
  if (c->pool_dev) (void)hipFree(c->pool_dev);
  if (c->pool_host) (void)hipHostFree(c->pool_host);
  c->pool_dev = nullptr;
  c->pool_host = nullptr;
  const size_t n = (size_t)steps * 6 * c->np;
  RC(dalloc(c, &c->pool_dev, n));
  if (!c->pool_counter) RC(dalloc(c, &c->pool_counter, 1));
  SPM_HIP(c, hipHostMalloc((void **)&c->pool_host, sizeof(double) * n, hipHostMallocDefault));
  c->pool_steps = steps;
  return 0;
}

// Draws the next k steps of noise (the order thermostat() draws them in,
// step by step) into pool slots slot0 .. slot0 + k - 1 and queues their
// copy on c->stream.  The caller makes sure no queued work still reads
// those slots and no earlier copy still reads their staging.
int pool_fill(spm_ctx *c, int k, int slot0) {
  if (!box_uses_noise(c)) return 0;
  if (slot0 < 0 || slot0 + k > c->pool_steps) {
    c->err = "noise pool smaller than the steps asked for";
    return SPM_ERR_INVALID;
  }
  const size_t n6 = (size_t)6 * c->np;
  double *h = c->pool_host + (size_t)slot0 * n6;
  for (int s = 0; s < k; ++s)
    for (int n = 0; n < c->np; ++n) {
      double *d = h + (size_t)s * n6 + (size_t)n * 6;
      gauss2(c, d);
      gauss2(c, d + 2);
      gauss2(c, d + 4);
    }
  SPM_HIP(c, hipMemcpyAsync(c->pool_dev + (size_t)slot0 * n6, h, sizeof(double) * n6 * k, hipMemcpyHostToDevice,
                            c->stream));
  return 0;
}

// the thermostat's next read is pool slot 0
int pool_reset(spm_ctx *c) {
  if (!box_uses_noise(c)) return 0;
  return k_zero_int(c, c->pool_counter, 1);
}

}  // namespace spm

extern "C" {

const char *spm_strerror(int status) {
  switch (status) {
    case SPM_OK: return "ok";
    case SPM_ERR_INVALID: return "invalid argument";
    case SPM_ERR_OOM: return "out of memory";
    case SPM_ERR_HIP: return "HIP runtime error";
    case SPM_ERR_NONFINITE: return "non-finite value in fields";
    case SPM_ERR_UNSUPPORTED: return "unsupported option";
    case SPM_ERR_NO_DEVICE: return "no GPU device";
    default: return "unknown status";
  }
}

const char *spm_last_error(const spm_ctx *ctx) { return ctx ? ctx->err.c_str() : spm::global_error().c_str(); }

int spm_create(const spm_params *prm, spm_ctx **out) {
  if (!prm || !out) return SPM_ERR_INVALID;
  *out = nullptr;
  spm::global_error().clear();
  const spm_params &p = *prm;
  if (!grid_size_ok(p.nx) || !grid_size_ok(p.ny) || !grid_size_ok(p.nz)) {
    spm::global_error() = "grid sizes must be powers of two in [8, 1024] or 48, 80, 96, 160, 192";
    return SPM_ERR_UNSUPPORTED;
  }
  if (p.n_particles < 0 || p.n_species < 1 || p.n_species > SPM_MAX_SPECIES || !(p.dt > 0) || !(p.dx > 0) ||
      !(p.rho > 0) || !(p.eta > 0) || !(p.a > 0) || !(p.a_xi > 0) || p.g_direction < 0 || p.g_direction > 2) {
    spm::global_error() = "invalid parameter";
    return SPM_ERR_INVALID;
  }
  for (int s = 0; s < p.n_species; ++s)
    if (!(p.mass_ratio[s] > 0)) {
      spm::global_error() = "mass ratios must be positive";
      return SPM_ERR_INVALID;
    }
  if (p.precision != SPM_FP64 && p.precision != SPM_FP32) {
    spm::global_error() = "precision must be SPM_FP64 or SPM_FP32";
    return SPM_ERR_INVALID;
  }
  const int nh = (int)p.a + (int)p.a_xi + 2;
  if (p.n_particles > 0 && (p.nx < 2 * nh + 1 || p.ny < 2 * nh + 1 || p.nz < 2 * nh + 1)) {
    spm::global_error() = "box narrower than a particle stencil (2*(a+xi+2)+1 points)";
    return SPM_ERR_UNSUPPORTED;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    spm::global_error() = "no HIP device";
    return SPM_ERR_NO_DEVICE;
  }
  if (p.device < 0 || p.device >= ndev) {
    spm::global_error() = "device ordinal out of range";
    return SPM_ERR_INVALID;
  }
  spm_ctx *c = new (std::nothrow) spm_ctx();
  if (!c) return SPM_ERR_OOM;
  c->prm = p;
  c->device = p.device;
  int rc = setup(c);
  if (rc) {
    spm::global_error() = c->err;
    free_ctx(c);
    return rc == SPM_ERR_HIP && spm::global_error().find("out of memory") != std::string::npos ? SPM_ERR_OOM : rc;
  }
  *out = c;
  return SPM_OK;
}

int spm_destroy(spm_ctx *ctx) {
  free_ctx(ctx);
  return SPM_OK;
}

int spm_stencil_size(const spm_ctx *ctx) { return ctx ? ctx->n_stencil : -1; }

// Fields cross the boundary as doubles; an FP32 context converts.
int spm_put_field(spm_ctx *c, int field, const double *host) {
  if (!c || !host) return SPM_ERR_INVALID;
  SPM_HIP(c, hipSetDevice(c->device));
  RC(spm::settle(c));
  double *p = field_ptr(c, field);
  if (!p) return SPM_ERR_INVALID;
  SPM_HIP(c, hipSetDevice(c->device));
  if (c->prm.precision == SPM_FP32) {
    std::vector<float> tmp(host, host + spm::ref_count(c->grid));
    RC(spm::rows_to_device(c, p, tmp.data(), sizeof(float)));
    SPM_HIP(c, hipStreamSynchronize(c->stream));
    return SPM_OK;
  }
  RC(spm::rows_to_device(c, p, host, sizeof(double)));
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  return SPM_OK;
}

int spm_get_field(spm_ctx *c, int field, double *host) {
  if (!c || !host) return SPM_ERR_INVALID;
  SPM_HIP(c, hipSetDevice(c->device));
  RC(spm::settle(c));
  double *p = field_ptr(c, field);
  if (!p) return SPM_ERR_INVALID;
  SPM_HIP(c, hipSetDevice(c->device));
  if (c->prm.precision == SPM_FP32) {
    const size_t n = spm::ref_count(c->grid);
    std::vector<float> tmp(n);
    RC(spm::rows_to_host(c, tmp.data(), p, sizeof(float)));
    SPM_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; ++i) host[i] = tmp[i];
    return SPM_OK;
  }
  RC(spm::rows_to_host(c, host, p, sizeof(double)));
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  return SPM_OK;
}

int spm_put_particles(spm_ctx *c, const spm_particle *p, int n) {
  if (!c || (!p && n > 0) || n != c->np) return SPM_ERR_INVALID;
  SPM_HIP(c, hipSetDevice(c->device));
  RC(spm::settle(c));  // phi and up of the particles they were formed from
  const double L[3] = {c->grid.lx, c->grid.ly, c->grid.lz};
  std::vector<int> spec(n > 0 ? n : 1);
  for (int i = 0; i < n; ++i) {
    for (int d = 0; d < 3; ++d)
      if (!(p[i].x[d] >= 0.0 && p[i].x[d] < L[d])) {
        c->err = "particle position outside [0, L)";
        return SPM_ERR_INVALID;
      }
    if (p[i].spec < 0 || p[i].spec >= c->prm.n_species) {
      c->err = "particle species out of range";
      return SPM_ERR_INVALID;
    }
    spec[i] = p[i].spec;
    const double *src = reinterpret_cast<const double *>(reinterpret_cast<const char *>(&p[i]) + 8);
    for (int s = 0; s < P_NSLOTS; ++s) c->pstage[(size_t)s * n + i] = src[s];
  }
  c->spec_host.assign(spec.begin(), spec.begin() + n);
  for (int s = 0; s < SPM_MAX_SPECIES; ++s) c->spec_count[s] = 0;
  for (int i = 0; i < n; ++i) c->spec_count[spec[i]]++;
  if (n == 0) return SPM_OK;
  SPM_HIP(c, hipSetDevice(c->device));
  SPM_HIP(c, hipMemcpyAsync(c->pbuf, c->pstage.data(), sizeof(double) * P_NSLOTS * n, hipMemcpyHostToDevice, c->stream));
  SPM_HIP(c, hipMemcpyAsync(c->spec, spec.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  return SPM_OK;
}

int spm_get_particles(spm_ctx *c, spm_particle *p, int n) {
  if (!c || (!p && n > 0) || n != c->np) return SPM_ERR_INVALID;
  if (n == 0) return SPM_OK;
  std::vector<int> spec(n);
  SPM_HIP(c, hipSetDevice(c->device));
  SPM_HIP(c, hipMemcpyAsync(c->pstage.data(), c->pbuf, sizeof(double) * P_NSLOTS * n, hipMemcpyDeviceToHost, c->stream));
  SPM_HIP(c, hipMemcpyAsync(spec.data(), c->spec, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; ++i) {
    std::memset(&p[i], 0, sizeof(spm_particle));
    p[i].spec = spec[i];
    double *dst = reinterpret_cast<double *>(reinterpret_cast<char *>(&p[i]) + 8);
    for (int s = 0; s < P_NSLOTS; ++s) dst[s] = c->pstage[(size_t)s * n + i];
  }
  return SPM_OK;
}

int spm_put_uk_dc(spm_ctx *c, const double uk_dc[3]) {
  if (!c || !uk_dc) return SPM_ERR_INVALID;
  SPM_HIP(c, hipSetDevice(c->device));
  RC(spm::settle(c));  // u's mean mode is formed from the old uk_dc
  SPM_HIP(c, hipMemcpyAsync(c->uk_dc, uk_dc, sizeof(double) * 3, hipMemcpyHostToDevice, c->stream));
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  return SPM_OK;
}

int spm_get_uk_dc(spm_ctx *c, double uk_dc[3]) {
  if (!c || !uk_dc) return SPM_ERR_INVALID;
  SPM_HIP(c, hipSetDevice(c->device));
  SPM_HIP(c, hipMemcpyAsync(uk_dc, c->uk_dc, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  return SPM_OK;
}

int spm_put_time(spm_ctx *c, int ts, double time) {
  if (!c || ts < 0) return SPM_ERR_INVALID;
  c->ts = ts;
  c->time = time;
  return SPM_OK;
}

int spm_get_time(spm_ctx *c, int *ts, double *time) {
  if (!c || !ts || !time) return SPM_ERR_INVALID;
  *ts = c->ts;
  *time = c->time;
  return SPM_OK;
}

int spm_put_rng(spm_ctx *c, const uint32_t mt[SPM_MT_N], int mti) {
  if (!c || !mt || mti < 0 || mti > SPM_MT_N + 1) return SPM_ERR_INVALID;
  std::memcpy(c->mt, mt, sizeof(uint32_t) * SPM_MT_N);
  c->mti = mti;
  return SPM_OK;
}

int spm_get_rng(spm_ctx *c, uint32_t mt[SPM_MT_N], int *mti) {
  if (!c || !mt || !mti) return SPM_ERR_INVALID;
  // the host stream is current as soon as the draw has been made
  std::memcpy(mt, c->mt, sizeof(uint32_t) * SPM_MT_N);
  *mti = c->mti;
  return SPM_OK;
}

// every entry point but spm_step first rebuilds what a fused step left
// stale (u, phi, up), so callers see the reference's end-of-step fields
#define ENTRY_NO_SETTLE(body)                      \
  do {                                             \
    if (!c) return SPM_ERR_INVALID;                \
    c->err.clear();                                \
    SPM_HIP(c, hipSetDevice(c->device));           \
    int rc_ = (body);                              \
    return rc_;                                    \
  } while (0)
#define ENTRY(body)                                \
  do {                                             \
    if (!c) return SPM_ERR_INVALID;                \
    c->err.clear();                                \
    SPM_HIP(c, hipSetDevice(c->device));           \
    int rc_ = spm::settle(c);                      \
    if (!rc_) rc_ = (body);                        \
    return rc_;                                    \
  } while (0)

static int init_fluid(spm_ctx *c) {
    // Init_zeta_k (init_fluid.cxx:10-23) from u = 0
    for (int d = 0; d < 3; ++d) RC(k_zero(c, c->u[d], c->nfield));
    double *const u3[3] = {c->u[0], c->u[1], c->u[2]};
    RC(fft_r2c_many(c, u3, 3));
    RC(k_u_k2zeta_k(c, c->u[0], c->u[1], c->u[2], c->zeta[0], c->zeta[1]));
    // sp_3d_ns.cxx:508-524
    RC(p_make_phi(c, 1));
    RC(k_zeta_k2u_k(c, c->zeta[0], c->zeta[1], c->u[0], c->u[1], c->u[2], 0));
    RC(fft_c2r_many(c, u3, 3));
    RC(make_f_sole_add(c));
    RC(u2zeta_k(c));
    SPM_HIP(c, hipMemsetAsync(c->uk_dc, 0, sizeof(double) * 3, c->stream));
    SPM_HIP(c, hipStreamSynchronize(c->stream));
    return 0;
}

static int calc_f_hydro_api(spm_ctx *c) {
  RC(reset_phi_u(c));
  return p_calc_f_hydro(c);
}

static int step_n(spm_ctx *c, int nsteps) {
  const bool fused = c->schedule == SPM_SCHEDULE_FUSED;
  if (!(fused && !c->electro && !c->le)) RC(spm::settle(c));
  // The fused passes carry only the modes the 2/3 rule keeps and leave zeta
  // untouched outside them; the reference's first act in a step is the same
  // truncation (Zeta_k2advection_k, operate_omega.cxx:102-108).
  const bool fused_ns = fused && !c->electro && !c->le;
  if (fused_ns && nsteps > 0) {
    RC(k_truncate(c, c->zeta[0]));
    RC(k_truncate(c, c->zeta[1]));
  }
  for (int s = 0; s < nsteps; ++s) {
    if (c->electro)
      RC(electrolyte_step(c));
    else if (c->le)
      RC(le_step(c, fix_mask(c), s == nsteps - 1));
    else
      RC(fused ? fused_step(c) : time_evolution_hydro(c));
    c->time += c->prm.dt;
    c->ts++;
    // keep the event pool bounded when timing long runs
    if (c->timing && c->launches.size() > 20000) {
      SPM_HIP(c, hipStreamSynchronize(c->stream));
      RC(fold_timing(c));
    }
  }
  // u, phi and up are rebuilt when something reads them (settle)
  if (fused_ns && nsteps > 0) c->stale = true;
  return 0;
}


static int check_finite(spm_ctx *c) {
  SPM_HIP(c, hipMemsetAsync(c->flag, 0, sizeof(int), c->stream));
  for (int d = 0; d < 3; ++d) RC(k_check_finite(c, c->u[d], false));
  for (int d = 0; d < 2; ++d) RC(k_check_finite(c, c->zeta[d], true));
  int h = 0;
  SPM_HIP(c, hipMemcpyAsync(&h, c->flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  if (h) {
    c->err = "NaN or Inf in u or zeta";
    return SPM_ERR_NONFINITE;
  }
  return 0;
}

int spm_init_fluid(spm_ctx *c) { ENTRY(init_fluid(c)); }

int spm_ns_solver_slaved_euler(spm_ctx *c) { ENTRY(ns_solver(c)); }
int spm_zeta_k2u(spm_ctx *c) { ENTRY(zeta_k2u(c, fix_mask(c))); }
int spm_md_solver_position(spm_ctx *c) { ENTRY(p_position(c, c->ts == 0)); }
int spm_calc_f_hydro(spm_ctx *c) { ENTRY(calc_f_hydro_api(c)); }
int spm_force(spm_ctx *c) { ENTRY(p_force(c)); }
int spm_md_solver_velocity(spm_ctx *c) { ENTRY(p_velocity(c, c->ts == 0)); }
int spm_add_random_force_thermostat(spm_ctx *c) { ENTRY(thermostat(c)); }
int spm_make_phi_u_particle(spm_ctx *c) { ENTRY(p_make_phi(c, 1)); }
int spm_make_f_particle_dt_sole(spm_ctx *c) { ENTRY(make_f_sole_add(c)); }
int spm_u2zeta_k(spm_ctx *c) { ENTRY(u2zeta_k(c)); }

int spm_step(spm_ctx *c, int nsteps) {
  if (nsteps < 0) return SPM_ERR_INVALID;
  ENTRY_NO_SETTLE(step_n(c, nsteps));
}

int spm_synchronize(spm_ctx *c) {
  if (!c) return SPM_ERR_INVALID;
  SPM_HIP(c, hipSetDevice(c->device));
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  return SPM_OK;
}

int spm_check_finite(spm_ctx *c) { ENTRY(check_finite(c)); }

int spm_fused_stage(spm_ctx *c, int stage) { ENTRY(fused_stage(c, stage)); }

int spm_set_schedule(spm_ctx *c, int schedule) {
  if (!c || (schedule != SPM_SCHEDULE_FUSED && schedule != SPM_SCHEDULE_LITERAL)) return SPM_ERR_INVALID;
  if (schedule == SPM_SCHEDULE_FUSED && (!pow2_grid(c) || !fused_fits(c))) {
    c->err = "the fused schedule needs power-of-two grids and fields under 4 GiB";
    return SPM_ERR_UNSUPPORTED;
  }
  c->schedule = schedule;
  return SPM_OK;
}

int spm_timing_enable(spm_ctx *c, int on) {
  if (!c) return SPM_ERR_INVALID;
  SPM_HIP(c, hipSetDevice(c->device));
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  RC(fold_timing(c));
  c->timing = on != 0;
  return SPM_OK;
}

int spm_timing_read(spm_ctx *c, double ms[SPM_NUM_KERNEL_FAMILIES], double bytes[SPM_NUM_KERNEL_FAMILIES],
                    long launches[SPM_NUM_KERNEL_FAMILIES]) {
  if (!c || !ms || !bytes || !launches) return SPM_ERR_INVALID;
  SPM_HIP(c, hipSetDevice(c->device));
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  RC(fold_timing(c));
  for (int f = 0; f < SPM_NUM_KERNEL_FAMILIES; ++f) {
    ms[f] = c->fam_ms[f];
    bytes[f] = c->fam_bytes[f];
    launches[f] = c->fam_count[f];
    c->fam_ms[f] = c->fam_bytes[f] = 0.0;
    c->fam_count[f] = 0;
  }
  return SPM_OK;
}


// ---- electrolyte ----
#define ELECTRO_ENTRY(body)                                  \
  do {                                                       \
    if (!c) return SPM_ERR_INVALID;                          \
    if (!c->electro) {                                       \
      c->err = "spm_set_electrolyte has not been called";    \
      return SPM_ERR_INVALID;                                \
    }                                                        \
    ENTRY(body);                                             \
  } while (0)

int spm_set_electrolyte(spm_ctx *c, const spm_electrolyte *el) { ENTRY(set_electrolyte(c, el)); }
int spm_init_rho_ion(spm_ctx *c) { ELECTRO_ENTRY(init_rho_ion(c)); }
int spm_init_rho_ion_pb(spm_ctx *c, int *iterations) { ELECTRO_ENTRY(init_rho_ion_pb(c, iterations)); }
int spm_nssolute_solver_euler(spm_ctx *c) { ELECTRO_ENTRY(nssolute_solver(c)); }
int spm_rescale_solute(spm_ctx *c) { ELECTRO_ENTRY(rescale_solute(c)); }
int spm_make_coulomb_force(spm_ctx *c) { ELECTRO_ENTRY(coulomb_force(c)); }

// ion spectra in the reference layout, FP64 on the host (rounded to FP32
// into an FP32 context's spectra)
int spm_put_concentration(spm_ctx *c, int n, const double *host) {
  if (!c || !host || !c->electro || n < 0 || n >= c->el.n_spec) return SPM_ERR_INVALID;
  SPM_HIP(c, hipSetDevice(c->device));
  if (c->prm.precision == SPM_FP32) {
    const size_t m = spm::ref_count(c->grid);
    std::vector<float> tmp(m);
    for (size_t i = 0; i < m; ++i) tmp[i] = (float)host[i];
    RC(spm::rows_to_device(c, c->conc[n], tmp.data(), sizeof(float)));
  } else {
    RC(spm::rows_to_device(c, c->conc[n], host, sizeof(double)));
  }
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  return SPM_OK;
}

int spm_get_concentration(spm_ctx *c, int n, double *host) {
  if (!c || !host || !c->electro || n < 0 || n >= c->el.n_spec) return SPM_ERR_INVALID;
  SPM_HIP(c, hipSetDevice(c->device));
  if (c->prm.precision == SPM_FP32) {
    const size_t m = spm::ref_count(c->grid);
    std::vector<float> tmp(m);
    RC(spm::rows_to_host(c, tmp.data(), c->conc[n], sizeof(float)));
    SPM_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < m; ++i) host[i] = tmp[i];
    return SPM_OK;
  }
  RC(spm::rows_to_host(c, host, c->conc[n], sizeof(double)));
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  return SPM_OK;
}

int spm_put_total_solute(spm_ctx *c, const double *total) {
  if (!c || !total || !c->electro) return SPM_ERR_INVALID;
  for (int n = 0; n < c->el.n_spec; ++n) c->total_solute[n] = total[n];
  return SPM_OK;
}

int spm_get_total_solute(spm_ctx *c, double *total) {
  if (!c || !total || !c->electro) return SPM_ERR_INVALID;
  for (int n = 0; n < c->el.n_spec; ++n) total[n] = c->total_solute[n];
  return SPM_OK;
}

// ---- Lees-Edwards shear ----
#define LE_ENTRY(body)                                       \
  do {                                                       \
    if (!c) return SPM_ERR_INVALID;                          \
    if (!c->le) {                                            \
      c->err = "spm_set_lees_edwards has not been called";   \
      return SPM_ERR_INVALID;                                \
    }                                                        \
    ENTRY(body);                                             \
  } while (0)

int spm_set_lees_edwards(spm_ctx *c, double shear_rate) { ENTRY(set_lees_edwards(c, shear_rate)); }
int spm_put_oblique(spm_ctx *c, double degree_oblique, double shear_rate_eff) {
  LE_ENTRY(put_oblique(c, degree_oblique, shear_rate_eff));
}
int spm_get_oblique(spm_ctx *c, double *degree_oblique, double *shear_rate_eff) {
  if (!degree_oblique || !shear_rate_eff) return SPM_ERR_INVALID;
  LE_ENTRY(get_oblique(c, degree_oblique, shear_rate_eff));
}
int spm_ns_solver_slaved_euler_shear_obl(spm_ctx *c) { LE_ENTRY(le_ns_solver(c)); }
int spm_le_shear_stress(spm_ctx *c, spm_le_stress *out) {
  if (!out) return SPM_ERR_INVALID;
  LE_ENTRY(le_shear_stress(c, out));
}
int spm_zeta_k2u_k_obl(spm_ctx *c) { LE_ENTRY(le_zeta_k2u_k_api(c)); }
int spm_u_k2zeta_k_obl(spm_ctx *c) { LE_ENTRY(le_u_k2zeta_k(c)); }
int spm_make_phi_u_particle_obl(spm_ctx *c) { LE_ENTRY(le_make_phi_api(c)); }
int spm_calc_f_hydro_obl(spm_ctx *c) { LE_ENTRY(le_f_hydro_api(c)); }
int spm_md_solver_position_obl(spm_ctx *c) { LE_ENTRY(le_position_api(c)); }
int spm_md_solver_velocity_obl(spm_ctx *c) { LE_ENTRY(le_velocity_api(c)); }

}  // extern "C"
