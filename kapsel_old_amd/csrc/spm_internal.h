// spm_internal.h -- shared definitions of the MI355X SPM library.
//
// Data layout in HBM (one context = one box on one GPU):
//   fields    NX*NY*NZ_ elements of T (double on the FP64 path), row-major
//             [i][j][k], NZ_ = 2*(NZ/2+1); real space uses k < NZ, a
//             spectrum holds NZ/2+1 interleaved complex values per row.
//   particles SoA copy of the reference Particle struct: slot s of
//             particle n at pbuf[s*Np + n] (s indexes the 124 doubles of
//             the struct after `spec`), spec in a separate int array.
//   lists     per-tile particle lists (make_phi) and per-cell particle
//             lists (LJ), both as CSR (start, count, sorted members).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/spm.h"

namespace spm {

// ---------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------
template <typename T>
struct cplx_t;
template <>
struct cplx_t<double> {
  using type = double2;
};
template <>
struct cplx_t<float> {
  using type = float2;
};
template <typename T>
using cplx = typename cplx_t<T>::type;

// Workgroups are dispatched to the 8 XCDs round robin (id mod 8), and each
// XCD has its own L2.  Give XCD k the contiguous run of tiles
// [k*n/8, (k+1)*n/8) so that tiles sharing a cache line at their edges meet
// in one L2.  The tail beyond a multiple of 8 keeps its own numbering.
// edge of the tiles that carry make_phi's particle lists (grid points)
#ifndef SPM_PHI_TILE
#define SPM_PHI_TILE 8
#endif
constexpr int kPhiTile = SPM_PHI_TILE;
constexpr int kPhiTilePoints = kPhiTile * kPhiTile * kPhiTile;

#ifndef SPM_XCD_REMAP
#define SPM_XCD_REMAP 1
#endif
__device__ __forceinline__ int xcd_tile(int w, int n) {
  if (!SPM_XCD_REMAP) return w;
  const int per = n >> 3;
  return w < (per << 3) ? (w & 7) * per + (w >> 3) : w;
}

template <typename C2>
__device__ __forceinline__ C2 cadd(C2 a, C2 b) {
  return C2{a.x + b.x, a.y + b.y};
}
template <typename C2>
__device__ __forceinline__ C2 csub(C2 a, C2 b) {
  return C2{a.x - b.x, a.y - b.y};
}
template <typename C2>
__device__ __forceinline__ C2 cmul(C2 a, C2 b) {
  return C2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// ---------------------------------------------------------------------
// particle SoA slots: offsets (in doubles) of the fields of spm_particle
// after `spec`, so slot = (offsetof(field) - 8) / 8.
// ---------------------------------------------------------------------
enum PSlot : int {
  P_X = 0, P_X_PREV = 3, P_X_NOPBC = 6, P_V = 9, P_V_OLD = 12, P_V_SLIP = 15,
  P_F_HYDRO = 18, P_F_HYDRO_PREV = 21, P_F_HYDRO1 = 24, P_F_SLIP = 27,
  P_F_SLIP_PREV = 30, P_FR = 33, P_FR_PREV = 36, P_OMEGA = 39,
  P_OMEGA_OLD = 42, P_OMEGA_SLIP = 45, P_TORQUE_HYDRO = 48,
  P_TORQUE_HYDRO_PREV = 51, P_TORQUE_HYDRO1 = 54, P_TORQUE_SLIP = 57,
  P_TORQUE_SLIP_PREV = 60, P_MOMENTUM_DEPEND_FR = 63, P_QR = 66,
  P_QR_OLD = 75, P_Q = 84, P_Q_OLD = 88, P_NSLOTS = 124
};

// Grid-wide constants handed to kernels by value.
struct Grid {
  int nx, ny, nz, nz_, hnz;  // hnz = NZ/2+1; nz_ = row pitch in elements (>= 2 hnz)
  int hp;                    // row pitch in complex elements (nz_ / 2)
  int trn_x, trn_y, trn_z;   // (N+2)/3, input.cxx:297
  double wave_x, wave_y, wave_z;
  double lx, ly, lz;
  double dx;
};

// Particle-model constants (Set_global_parameters, input.cxx:274-433).
struct PModel {
  int np;
  int rotation;
  double dx, radius, hxi, max_radius;  // max_radius: stencil bound
  int nh;                              // half-width of the stencil cube
  double rho, dx3;
  double imass[SPM_MAX_SPECIES], imoi[SPM_MAX_SPECIES];
  double mass_ratio[SPM_MAX_SPECIES];
  double lj_dia, lj_cutoff, epsilon;
  int lj_powers, lj_on;
  double g_on_fluid;
  int g_direction;
};

// moments per particle the sheared f_hydro writes for the stress output
constexpr int kStressSlots = 10;

// surface charge (times the elementary charge) of each particle species
struct QTable {
  double q[SPM_MAX_SPECIES];
};

}  // namespace spm

// ---------------------------------------------------------------------
// the context (opaque to callers of include/spm.h)
// ---------------------------------------------------------------------
struct spm_ctx {
  spm_params prm;
  spm::Grid grid;
  spm::PModel pm;
  int device = 0;
  int n_cu = 256;  // compute units of the device (persistent launches)
  hipStream_t stream = nullptr;
  // second stream of the fused step: the particle work that needs only the
  // previous step's velocities (positions, tile lists, LJ cell lists and
  // forces) runs on it beside the Navier-Stokes passes; fork and join events
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::string err;
  size_t nfield = 0;  // elements per field
  int ts = 0;
  double time = 0.0;
  int schedule = SPM_SCHEDULE_FUSED;
  // u, phi and up still to be rebuilt from zeta and the particles: a fused
  // spm_step defers that to the first entry point that reads or replaces
  // them (settle), since the next fused step does not read them
  bool stale = false;

  // fields (double path)
  double *zeta[2] = {nullptr, nullptr};
  double *u[3] = {nullptr, nullptr, nullptr};
  double *up[3] = {nullptr, nullptr, nullptr};
  double *phi = nullptr;
  double *f[3] = {nullptr, nullptr, nullptr};
  double *work[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  double *uk_dc = nullptr;  // device, 3 doubles

  // FFT tables: roots exp(+2 pi i t/N), t < N, for N = NX, NY, NZ/2, and
  // the real-to-complex post twiddle exp(+2 pi i k/NZ), k <= NZ/2.
  double2 *tw_x = nullptr, *tw_y = nullptr, *tw_z = nullptr, *tw_rz = nullptr;

  // particles
  int np = 0;
  double *pbuf = nullptr;  // P_NSLOTS * np
  int *spec = nullptr;
  int *stencil = nullptr;  // 3 * n_stencil ints (Sekibun_cell)
  int n_stencil = 0;
  int *pxint = nullptr;     // 3 * np: anchor cell of each particle (Particle_cell)
  double *pres = nullptr;   // 3 * np: residue x - anchor*DX

  // per-tile particle lists for make_phi (tiles of 8x8x8 points)
  int tiles[3] = {0, 0, 0};
  int ntiles = 0;
  int *tile_count = nullptr, *tile_start = nullptr, *tile_cursor = nullptr;
  int *tile_list = nullptr;
  int tile_list_cap = 0;
  // the same under Lees-Edwards shear, each particle also entered in the
  // tiles of its images across the y faces (allocated by spm_set_lees_edwards)
  int *tile_count_obl = nullptr, *tile_start_obl = nullptr, *tile_cursor_obl = nullptr;
  int *tile_list_obl = nullptr;
  long tile_list_obl_cap = 0;

  // LJ cell list
  int cells[3] = {0, 0, 0};
  int ncells = 0;
  int *cell_count = nullptr, *cell_start = nullptr, *cell_cursor = nullptr;
  int *cell_list = nullptr;
  int *pcell = nullptr;     // np: LJ cell of each particle
  int *flag = nullptr;      // device int: finiteness check result

  // thermostat noise (host MT19937, device apply)
  uint32_t mt[SPM_MT_N];
  int mti = SPM_MT_N + 1;
  double *noise_host[2] = {nullptr, nullptr};  // pinned
  double *noise_dev = nullptr;
  hipEvent_t noise_done[2] = {nullptr, nullptr};
  int noise_slot = 0;
  double noise_v = 0.0, noise_o = 0.0;
  // noise drawn ahead for pool_steps whole steps (ensembles replaying a
  // captured step): with pool_mode the thermostat takes slot
  // *pool_counter mod pool_steps of pool_dev on the device and advances the
  // counter, so a replayed graph needs no host work per step
  bool pool_mode = false;
  int pool_steps = 0;
  double *pool_dev = nullptr;   // pool_steps * 6 * np
  double *pool_host = nullptr;  // pinned staging of the same size
  int *pool_counter = nullptr;  // device

  // host staging
  std::vector<double> pstage;

  // particle species counts and sum |surface charge| over particles, in
  // particle order (kept on the host at spm_put_particles)
  int spec_count[SPM_MAX_SPECIES] = {};
  std::vector<int> spec_host;

  // electrolyte (SW_EQ == Electrolyte): ion spectra, their right-hand sides,
  // surface normals, the solvent right-hand side, reduction scratch
  // Lees-Edwards shear (SW_EQ == Shear_Navier_Stokes_Lees_Edwards), fluid
  // only: Shear_rate, degree_oblique (host), Shear_rate_eff (device, after
  // the reduction partials of le_red)
  bool le = false;
  double le_rate = 0.0, le_deg = 0.0;
  double a_r_cutoff = 0.0;  // A_R_cutoff (input.cxx:1461-1480), lj_cutoff = a_r_cutoff * lj_dia
  double *le_red = nullptr;
  // stress output (Mean_shear_stress): per particle kStressSlots moments
  // from the last f_hydro_obl, then one LJ pair virial per particle; the
  // realized strain (Shear_strain_realized) on the device
  double *le_stress = nullptr;
  double *le_strain = nullptr;
  bool electro = false;
  spm_electrolyte el{};
  double *conc[2] = {nullptr, nullptr};
  double *conc_rhs[2] = {nullptr, nullptr};
  double *surface_normal[3] = {nullptr, nullptr, nullptr};
  double *f_ns1[2] = {nullptr, nullptr};
  double *red = nullptr;  // partial sums, then result scalars
  // FP64 Poisson side of an FP32 context (an FP64 context uses its fields):
  // charge density, potential, its gradient, the particles' charge profile
  double *el_q = nullptr, *el_p = nullptr, *el_g[3] = {nullptr, nullptr, nullptr}, *el_pq = nullptr;
  double total_solute[2] = {0.0, 0.0};
  spm::QTable qtab{};
  double valency_e[2] = {0.0, 0.0};

  // learned closure added to f_hydro / torque_hydro (include/spm_closure.h)
  // and the particle motion of the step (free, or prescribed: held fixed)
  struct spm_closure *closure = nullptr;
  int motion = 0;

  // optional per-kernel-family timing (HIP events on `stream`)
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct TimedLaunch {
    int family;
    size_t e0, e1;
    double bytes;
  };
  std::vector<TimedLaunch> launches;
  double fam_ms[SPM_NUM_KERNEL_FAMILIES] = {};
  double fam_bytes[SPM_NUM_KERNEL_FAMILIES] = {};
  long fam_count[SPM_NUM_KERNEL_FAMILIES] = {};
};

// ---------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------
#define SPM_HIP(ctx, call)                                                      \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess) {                                                     \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);           \
      return SPM_ERR_HIP;                                                       \
    }                                                                           \
  } while (0)

#define SPM_LAUNCH_CHECK(ctx)                                                   \
  do {                                                                          \
    hipError_t e_ = hipGetLastError();                                          \
    if (e_ != hipSuccess) {                                                     \
      (ctx)->err = std::string("kernel launch: ") + hipGetErrorString(e_);      \
      return SPM_ERR_HIP;                                                       \
    }                                                                           \
  } while (0)

// ---------------------------------------------------------------------
// launchers implemented in the .hip files
// ---------------------------------------------------------------------
namespace spm {
// Grid points per axis the transforms take: powers of two in [8, 1024],
// or one of the mixed-radix lengths 48, 80, 96, 160, 192 (factors 2, 3, 5).
inline bool grid_pow2(int n) { return n >= 8 && n <= 1024 && (n & (n - 1)) == 0; }
inline bool grid_size_ok(int n) { return grid_pow2(n) || n == 48 || n == 80 || n == 96 || n == 160 || n == 192; }
// last error of a call without a context (this thread)
std::string &global_error();
// timing brackets around a launch; no-ops unless timing is enabled
void tbeg(spm_ctx *c);
void tend(spm_ctx *c, int family, double bytes);

// fft.hip: 3-D real transforms in the reference layout (in place).
int fft_r2c(spm_ctx *c, double *a);
int fft_c2r(spm_ctx *c, double *a);
// FP64 transforms whatever the context's field type (the Poisson side of an
// FP32 electrolyte context; its row pitch is also a valid FP64 pitch)
int fft_r2c_f64(spm_ctx *c, double *a);
int fft_c2r_many_f64(spm_ctx *c, double *const *a, int n);
// batched versions over several fields (same grid)
int fft_r2c_many(spm_ctx *c, double *const *a, int n);
int fft_c2r_many(spm_ctx *c, double *const *a, int n);
// with the band-limited y and z passes (inverse of spectra zero outside the
// 2/3 box; forward whose modes outside it are left unspecified)
int fft_r2c_many_band(spm_ctx *c, double *const *a, int n);
int fft_c2r_many_band(spm_ctx *c, double *const *a, int n);
// the same without the z pass (x and y of an inverse; y and x of a forward)
int fft_c2r_band_xy(spm_ctx *c, double *const *a, int n);
int fft_r2c_band_yx(spm_ctx *c, double *const *a, int n);
// z pass of the sheared solver: C2R of u and w, u x w, R2C into u (FP64)
int fz_cross_obl(spm_ctx *c, double *const u[3], double *const w[3]);
int fft_init_tables(spm_ctx *c);
// context-free transform of one FP64 field (spm_fft3d): roots in `ws`, at
// least fft3d_workspace_bytes(); 0 bytes means the sizes are unsupported
size_t fft3d_workspace_bytes(int nx, int ny, int nz);
int fft3d_raw(int nx, int ny, int nz, int direction, double *a, void *ws, hipStream_t st);
int fft_pass_y(spm_ctx *c, double *const *a, int n, int sign);
int fft_pass_z_c2r(spm_ctx *c, double *const *a, int n);
int fft_pass_z_r2c(spm_ctx *c, double *const *a, int n);

// fused.hip: FFT passes carrying the pointwise work of the step
int fx_zeta2u(spm_ctx *c, int fix_mask);
int fz_advect(spm_ctx *c);
int fx_update(spm_ctx *c, double nu, double dt);
int fx_update_zeta2u(spm_ctx *c, double nu, double dt, int fix_mask);
int fx_project(spm_ctx *c);

// spectral.hip
int k_truncate(spm_ctx *c, double *a);
int k_zeta_k2u_k(spm_ctx *c, const double *z0, const double *z1, double *u0, double *u1, double *u2, int fix_cell_mask);
int k_u_k2zeta_k(spm_ctx *c, const double *u0, const double *u1, const double *u2, double *z0, double *z1);
int k_products(spm_ctx *c, double *u0, double *u1, double *u2, double *a0, double *a1);
int k_advection_contract(spm_ctx *c, double *u0, double *u1, double *u2, const double *a0, const double *a1, double *o0, double *o1);
int k_slaved_euler(spm_ctx *c, double *z0, double *z1, const double *f0, const double *f1, double nu, double dt);
int k_solenoidal_trunc(spm_ctx *c, double *u0, double *u1, double *u2);
int k_make_f(spm_ctx *c);
int k_add_f(spm_ctx *c);
int k_zero(spm_ctx *c, double *a, size_t n);
int k_zero_f64(spm_ctx *c, double *a, size_t n);  // n doubles, whatever the field type
int k_zero_int(spm_ctx *c, int *a, size_t n);
int k_check_finite(spm_ctx *c, const double *a, bool spectrum);
// field copies between the reference layout on the host and the device
// pitch (spm_api.cpp)
inline size_t ref_count(const Grid &g) { return (size_t)g.nx * g.ny * 2 * g.hnz; }
int rows_to_device(spm_ctx *c, void *dev, const void *host, size_t elem);
int rows_to_host(spm_ctx *c, void *host, const void *dev, size_t elem);

// particles.hip
int p_position(spm_ctx *c, int euler);
int p_calc_f_hydro(spm_ctx *c);
int p_force(spm_ctx *c);
int p_velocity(spm_ctx *c, int euler);
// the same without their first parts (tile lists; LJ forces), for a caller
// that has run those already
int p_calc_f_hydro_listed(spm_ctx *c);
int p_velocity_forced(spm_ctx *c, int euler);
int p_thermostat_apply(spm_ctx *c, const double *noise_dev);
// noise_dev = slot *pool_counter of pool_dev, then the counter advances
int p_noise_from_pool(spm_ctx *c);
// ensembles (ensemble.cpp): spm_step in phases (0 before the steps, 1 one
// step, 2 after them), the thermostat's noise pool, and the timing
// autotuner of the mixed-radix launch shapes for the context's grid
int box_step_phase(spm_ctx *c, int phase);
// rebuilds u, phi and up if a fused step left them stale
int settle(spm_ctx *c);
bool box_uses_noise(const spm_ctx *c);
int pool_setup(spm_ctx *c, int steps);
int pool_fill(spm_ctx *c, int k, int slot0);
int pool_reset(spm_ctx *c);
int tune_mixed_transforms(spm_ctx *c);
// the autotuner's table as text, one line per tuned pass
std::string tune_report();
// spherical particles under Lees-Edwards shear; sre is Shear_rate_eff on the device
int p_position_obl(spm_ctx *c, int euler, const double *sre);
int p_calc_f_hydro_obl(spm_ctx *c, double *const ucp[3], const double *sre);
int p_velocity_obl(spm_ctx *c, int euler);
int p_make_phi_obl(spm_ctx *c, const double *sre);
// tiles per axis that one region of a sheared particle's list can span
int obl_tile_span(const spm_ctx *c, int axis);
int p_f_add_obl(spm_ctx *c, double *const ucp[3]);
// Add_random_force_thermostat: host noise draw + device update
int host_thermostat(spm_ctx *c);
// U_oblique2u of the real-space oblique velocity u into ucp (Cartesian)
int le_oblique2u(spm_ctx *c, double *const ucp[3]);
int p_build_tiles(spm_ctx *c);
int p_make_phi(spm_ctx *c, int sw_up);
int p_make_f(spm_ctx *c);
int p_init_lists(spm_ctx *c);
int p_calc_f_hydro_field(spm_ctx *c, double *const fld[3]);
// closure.hip: the attached network's force/torque correction (after
// p_calc_f_hydro, whose tile lists and anchors it reads)
int closure_apply(spm_ctx *c);
int electrolyte_step(spm_ctx *c);
// Lees-Edwards (shear.hip)
int le_red_slots();
int le_zeta2u_k(spm_ctx *c, double *const u[3], int fix_mask);
// store_u: leave u holding the covariant spectrum, as the reference does
// (the next step does not read it, so only the last step of a call stores it)
int le_u_k2zeta_k(spm_ctx *c, bool store_u = true);
int le_ns_solver(spm_ctx *c);
int le_step(spm_ctx *c, int fix_mask, bool store_u = true);

// electrolyte kernels (electrolyte.hip)
int e_red_slots();
int e_poisson_grad(spm_ctx *c, const double *pot, double *const grad[3], int trunc, const double cst[3]);
int e_grad_scaled(spm_ctx *c, const double *a, double scale, double *const out[3]);
int e_divergence(spm_ctx *c, double *const flux[3], double *div);
// fluid false: only the ion spectra (Ion_diffusion_solver_Euler)
int e_nssolute_update(spm_ctx *c, double nu, double dt, bool fluid = true);
int e_coulomb_potential(spm_ctx *c, double *pot);
int e_half_potential(spm_ctx *c, double *pot, const double ext[3], int external);
int e_charge_add(spm_ctx *c, double *charge, const double *conc, const double *phi, double ze);
int e_np_flux(spm_ctx *c, double *const flux[3], const double *cx, double *const grad[3], double *const normal[3],
              double *const u[3], double onsager, double ze);
// f = charge (-grad + E_ext); grad FP64, f of the field type (may be grad in FP64)
int e_coulomb(spm_ctx *c, const double *charge, double *const grad[3], double *const f[3], const double ext[3],
              int external);
int e_to_t(spm_ctx *c, double *const a[3], double *const b[3]);  // FP64 real fields -> field type
int e_add_scaled(spm_ctx *c, double *const u[3], double *const f[3], double dmy);
int e_fill_real(spm_ctx *c, double *a, double v);
int e_reduce(spm_ctx *c, int mode, const double *a, const double *phi, int slot, int ratio, double numer);
int e_scale_by(spm_ctx *c, double *a, long n, int slot, bool f64 = false);
int e_read_slot(spm_ctx *c, int slot, double *out);
int p_copy_hydro1(spm_ctx *c);
int p_make_phi_qq(spm_ctx *c, double *phi, double *surface);
int p_surface_normal(spm_ctx *c);
}  // namespace spm
