// shear.hip -- Lees-Edwards shear in oblique coordinates (SW_EQ ==
// Shear_Navier_Stokes_Lees_Edwards): Time_evolution_hydro_OBL
// (sp_3d_ns.cxx:214-390); the particle kernels it calls are in
// particles.hip.
//
// The grid shears with the flow: degree_oblique (the strain since the last
// remap) tilts the y axis, wavevectors keep their indices, and the metric
// enters as K2 = (WX kx)^2 + (WY ky - WX deg kx)^2 + (WZ kz)^2 and as the
// contravariant / covariant conversions of fft_wrapper.h:327-424.  Each
// kernel is pointwise (one thread per complex mode or grid point); the FFT
// passes are the library's.  FP64 only.
#include "spm_internal.h"

namespace spm {

namespace {

constexpr int kBlock = 256;
constexpr int kRedBlocks = 1024;

__device__ __forceinline__ double sq(double x) { return x * x; }

struct ModeO {
  long im;  // index of the real part in the [NX][NY][NZ_] array
  int kx, ky, kz;
  bool keep;  // inside the 2/3-rule box
};

__device__ __forceinline__ bool mode_of(const Grid &g, long t, ModeO &m) {
  const long nmodes = (long)g.nx * g.ny * g.hnz;
  if (t >= nmodes) return false;
  const int kk = (int)(t % g.hnz);
  const long ij = t / g.hnz;
  const int j = (int)(ij % g.ny);
  const int i = (int)(ij / g.ny);
  m.im = ij * g.nz_ + 2 * kk;
  m.kx = i > g.nx / 2 ? i - g.nx : i;
  m.ky = j > g.ny / 2 ? j - g.ny : j;
  m.kz = kk;
  m.keep = abs(m.kx) < g.trn_x && abs(m.ky) < g.trn_y && m.kz < g.trn_z;
  return true;
}

// K2 and IK2 of the oblique metric (sp_3d_ns.cxx:319-334)
__device__ __forceinline__ double k2_obl(const Grid &g, const ModeO &m, double deg) {
  return sq(g.wave_x * m.kx) + sq(g.wave_y * m.ky - g.wave_x * deg * m.kx) + sq(g.wave_z * m.kz);
}

// contra2co_single / co2contra_single (fft_wrapper.h:398-424)
__device__ __forceinline__ void contra2co1(double deg, double v[3]) {
  const double d0 = v[0], d1 = v[1];
  v[0] = d0 + deg * d1;
  v[1] = deg * d0 + (1. + deg * deg) * d1;
}
__device__ __forceinline__ void co2contra1(double deg, double v[3]) {
  const double d0 = v[0], d1 = v[1];
  v[0] = (1. + deg * deg) * d0 - deg * d1;
  v[1] = -deg * d0 + d1;
}

// Zeta_k2u_k_OBL (fft_wrapper.cxx:513-584) at one mode: zeta -> the
// contravariant velocity (real parts ur, imaginary parts ui); mode 0 is
// left to the caller (it takes uk_dc).
__device__ __forceinline__ void zeta2u_mode(const Grid &g, const ModeO &m, double deg, const double *__restrict__ z0,
                                            const double *__restrict__ z1, double ur[3], double ui[3]) {
  double kx = g.wave_x * m.kx, ky = g.wave_y * m.ky, kz = g.wave_z * m.kz;
  const double k2 = k2_obl(g, m, deg);
  const double ik2 = k2 > 0.0 ? 1. / k2 : 0.0;
  const double d1r = z0[m.im], d1i = z0[m.im + 1], d2r = z1[m.im], d2i = z1[m.im + 1];
  double wr[3], wi[3];
  if (m.kx != 0) {
    wr[0] = -(1. / kx) * (ky * d1r + kz * d2r);
    wi[0] = -(1. / kx) * (ky * d1i + kz * d2i);
    wr[1] = d1r; wi[1] = d1i;
    wr[2] = d2r; wi[2] = d2i;
  } else if (m.kz != 0) {
    const double dmy = -(ky / kz);
    wr[0] = d2r; wi[0] = d2i;
    wr[1] = d1r; wi[1] = d1i;
    wr[2] = dmy * d1r; wi[2] = dmy * d1i;
  } else {
    wr[0] = d1r; wi[0] = d1i;
    wr[1] = 0.0; wi[1] = 0.0;
    wr[2] = d2r; wi[2] = d2i;
  }
  contra2co1(deg, wr);
  contra2co1(deg, wi);
  kx *= ik2;
  ky *= ik2;
  kz *= ik2;
  ur[0] = (ky * wi[2] - kz * wi[1]);
  ur[1] = (kz * wi[0] - kx * wi[2]);
  ur[2] = (kx * wi[1] - ky * wi[0]);
  ui[0] = -(ky * wr[2] - kz * wr[1]);
  ui[1] = -(kz * wr[0] - kx * wr[2]);
  ui[2] = -(kx * wr[1] - ky * wr[0]);
}

// Zeta_k2u_k_OBL over the spectrum; mode 0 takes u = uk_dc (real part),
// whose FIX_CELL components are zeroed first when fix_mask has their bits.
// 32 B read, 48 B written per mode.
__global__ void zeta2u_obl_kernel(Grid g, const double *__restrict__ z0, const double *__restrict__ z1, double *u0,
                                  double *u1, double *u2, double *uk_dc, int fix_mask, double deg) {
  ModeO m;
  if (!mode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m)) return;
  double ur[3], ui[3];
  zeta2u_mode(g, m, deg, z0, z1, ur, ui);
  if (m.im == 0)
    for (int d = 0; d < 3; ++d) {
      if (fix_mask & (1 << d)) uk_dc[d] = 0.0;
      ur[d] = uk_dc[d];
    }
  u0[m.im] = ur[0];
  u0[m.im + 1] = ui[0];
  u1[m.im] = ur[1];
  u1[m.im + 1] = ui[1];
  u2[m.im] = ur[2];
  u2[m.im + 1] = ui[2];
}

// Zeta_k2u_k_OBL (mode 0: uk_dc, real part) and Zeta_k2omega_k_OBL
// (fft_wrapper.cxx:513-632) from one read of zeta: 32 B read, 96 B written
// per mode; omega(0) = 0.
__global__ void zeta2uw_obl_kernel(Grid g, const double *__restrict__ z0, const double *__restrict__ z1, double *u0,
                                   double *u1, double *u2, double *w0, double *w1, double *w2,
                                   const double *__restrict__ uk_dc, double deg) {
  ModeO m;
  if (!mode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m)) return;
  double ur[3], ui[3];
  zeta2u_mode(g, m, deg, z0, z1, ur, ui);
  if (m.im == 0)
    for (int d = 0; d < 3; ++d) ur[d] = uk_dc[d];
  u0[m.im] = ur[0];
  u0[m.im + 1] = ui[0];
  u1[m.im] = ur[1];
  u1[m.im + 1] = ui[1];
  u2[m.im] = ur[2];
  u2[m.im + 1] = ui[2];
  const double kx = g.wave_x * m.kx, ky = g.wave_y * m.ky, kz = g.wave_z * m.kz;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const long im = m.im + s;
    const double d1 = z0[im], d2 = z1[im];
    double o0, o1, o2;
    if (m.kx != 0) {
      o0 = -(1. / kx) * (ky * d1 + kz * d2);
      o1 = d1;
      o2 = d2;
    } else if (m.kz != 0) {
      o0 = d2;
      o1 = d1;
      o2 = -(ky / kz) * d1;
    } else {
      o0 = d1;
      o1 = 0.0;
      o2 = d2;
    }
    if (im == 0) o0 = o1 = o2 = 0.0;
    w0[im] = o0;
    w1[im] = o1;
    w2[im] = o2;
  }
}

// The tail of Zeta_k2advection_k_OBL (operate_omega.cxx:158-172) and the
// update of NS_solver_slavedEuler_Shear_OBL (fluid_solver.cxx:407-428) in
// one pass over the kept modes (the update reads the advection only there):
// the shear terms u0 -= 2 rate uk_cp, u1 -= 2 rate deg uk_cp,
// U_k2rotation_k (fft_wrapper.cxx:679-718), Omega_k2zeta_k_OBL
// (fft_wrapper.cxx:175-198), then zeta += (exp(-nu dt K2) - 1)(zeta -
// adv / (nu K2)) with the oblique K2.  `rate` is Shear_rate_eff on the
// device.  uk_cp, the y velocity spectrum of the step's zeta
// (operate_omega.cxx:124-128), is formed again here from zeta rather than
// kept in a field of its own.  80 B read, 32 B written per kept mode.
__global__ void advect_update_obl_kernel(Grid g, const double *__restrict__ u0, const double *__restrict__ u1,
                                         const double *__restrict__ u2, double *z0, double *z1,
                                         const double *__restrict__ uk_dc, const double *__restrict__ rate_p,
                                         double deg, double nu, double dt) {
  ModeO m;
  if (!mode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m) || !m.keep) return;
  const double rate = *rate_p;
  double cr[3], ci[3];
  zeta2u_mode(g, m, deg, z0, z1, cr, ci);
  if (m.im == 0) cr[1] = uk_dc[1];
  const double cp[2] = {cr[1], ci[1]};
  const double ks[3] = {m.kx * g.wave_x, m.ky * g.wave_y, m.kz * g.wave_z};
  double re[3], iv[3];
  re[0] = u0[m.im] - 2. * rate * cp[0];
  iv[0] = u0[m.im + 1] - 2. * rate * cp[1];
  re[1] = u1[m.im] - 2. * rate * deg * cp[0];
  iv[1] = u1[m.im + 1] - 2. * rate * deg * cp[1];
  re[2] = u2[m.im];
  iv[2] = u2[m.im + 1];
  double rr[3], ri[3];
  rr[0] = (ks[1] * iv[2] - ks[2] * iv[1]);
  ri[0] = -(ks[1] * re[2] - ks[2] * re[1]);
  rr[1] = (ks[2] * iv[0] - ks[0] * iv[2]);
  ri[1] = -(ks[2] * re[0] - ks[0] * re[2]);
  rr[2] = (ks[0] * iv[1] - ks[1] * iv[0]);
  ri[2] = -(ks[0] * re[1] - ks[1] * re[0]);
  const int a = m.kx != 0 ? 1 : (m.kz != 0 ? 1 : 0);
  const int b = m.kx != 0 ? 2 : (m.kz != 0 ? 0 : 2);
  const double adv0[2] = {rr[a], ri[a]}, adv1[2] = {rr[b], ri[b]};
  const double k2 = k2_obl(g, m, deg);
  const double ik2 = k2 > 0.0 ? 1. / k2 : 0.0;
  const double dmy0 = -nu * dt, dmy1 = 1. / nu;
  const double dmy = exp(dmy0 * k2) - 1.;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const long im = m.im + s;
    z0[im] += dmy * (z0[im] - dmy1 * ik2 * adv0[s]);
    z1[im] += dmy * (z1[im] - dmy1 * ik2 * adv1[s]);
  }
}

// The integer strain: once degree_oblique reaches 1 the grid is remapped,
// u0 += u1 (sp_3d_ns.cxx:281-305).  Gather form of the reference's
// scatter: row (i, j) goes to (i + j - NY/2) mod NX.  48 B read, 24 B
// written per point.
__global__ void remap_kernel(Grid g, const double *__restrict__ c0, const double *__restrict__ c1,
                             const double *__restrict__ c2, double *u0, double *u1, double *u2) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  const int k = (int)(t % g.nz), j = (int)((t / g.nz) % g.ny), id = (int)(t / ((long)g.nz * g.ny));
  const int io = j - g.ny / 2;
  const int i = ((id - io) % g.nx + g.nx) % g.nx;
  const long src = ((long)i * g.ny + j) * g.nz_ + k, dst = ((long)id * g.ny + j) * g.nz_ + k;
  u0[dst] = c0[src] + c1[src];
  u1[dst] = c1[src];
  u2[dst] = c2[src];
}

// U_oblique2u (fft_wrapper.h:265-325) in gather form: destination row
// (id, j) takes source rows i = id - io(j) and i + sign(j - NY/2) of the
// oblique field w, with the reference's weights, plus the mean shear flow
// rate (j - NY/2).  Each source row serves two destination rows: 24 B read,
// 24 B written per point.
__global__ void oblique2u_kernel(Grid g, const double *__restrict__ w0, const double *__restrict__ w1,
                                 const double *__restrict__ w2, double *u0, double *u1, double *u2, double deg,
                                 double rate) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  const int k = (int)(t % g.nz), j = (int)((t / g.nz) % g.ny), id = (int)(t / ((long)g.nz * g.ny));
  const int NX = g.nx, NY = g.ny;
  double sign = j - NY / 2;
  if (!(sign == 0)) sign = sign / fabs(sign);
  const int io = (int)((int)(sign * deg * (j - NY / 2.)) * sign + sign);
  const double alpha = (io - deg * (j - NY / 2.)) * sign;
  const double beta = 1. - alpha;
  const int i = ((id - io) % NX + NX) % NX;
  const int ip = (int)fmod(i + sign + 2 * NX, NX);
  const long s = ((long)i * NY + j) * g.nz_ + k, sp = ((long)ip * NY + j) * g.nz_ + k;
  const long dst = ((long)id * NY + j) * g.nz_ + k;
  u0[dst] = (beta * w0[s] + alpha * w0[sp]) + deg * (beta * w1[s] + alpha * w1[sp]) + rate * (j - NY / 2.);
  u1[dst] = beta * w1[s] + alpha * w1[sp];
  u2[dst] = beta * w2[s] + alpha * w2[sp];
}

// The x velocity that U_oblique2u (fft_wrapper.h:265-325) forms at grid
// point (id, j, k) from the oblique field w0, w1: destination row (id, j)
// interpolates source rows i = id - io(j) and i + sign(j - NY/2) with the
// reference's weights, and adds the mean shear flow rate (j - NY/2).
__device__ __forceinline__ double cart_u0(const Grid &g, const double *__restrict__ w0, const double *__restrict__ w1,
                                          int id, int j, int k, double deg, double rate) {
  const int NX = g.nx, NY = g.ny;
  double sign = j - NY / 2;
  if (!(sign == 0)) sign = sign / fabs(sign);
  const int io = (int)((int)(sign * deg * (j - NY / 2.)) * sign + sign);
  const double alpha = (io - deg * (j - NY / 2.)) * sign;
  const double beta = 1. - alpha;
  const int i = ((id - io) % NX + NX) % NX;
  const int ip = (int)fmod(i + sign + 2 * NX, NX);
  const long s = ((long)i * NY + j) * g.nz_ + k, sp = ((long)ip * NY + j) * g.nz_ + k;
  return (beta * w0[s] + alpha * w0[sp]) + deg * (beta * w1[s] + alpha * w1[sp]) + rate * (j - NY / 2.);
}

// Calc_shear_rate_eff (sp_3d_ns.h:174-196) of the Cartesian velocity
// U_oblique2u makes of the oblique field (sp_3d_ns.cxx:319-341).  For each
// (i, k) the sum over j < NY-1 of (ux(j+1) - ux(j)) / DX telescopes to
// (ux(NY-1) - ux(0)) / DX, so only rows 0 and NY-1 of ux are formed; without
// particles nothing else reads the Cartesian field (with particles
// Calc_f_hydro_correct_precision_OBL and Make_f_particle_dt_nonsole do).
// 32 B read per (i, k) column; partial sums in a fixed order.
__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
  return s;
}

__global__ __launch_bounds__(kBlock) void shear_rate_rows_kernel(Grid g, const double *__restrict__ w0,
                                                                 const double *__restrict__ w1, double deg,
                                                                 double rate, double *part) {
  __shared__ double sh[kBlock / 64];
  const long n = (long)g.nx * g.nz;
  double v = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const int k = (int)(t % g.nz), i = (int)(t / g.nz);
    v += (cart_u0(g, w0, w1, i, g.ny - 1, k, deg, rate) - cart_u0(g, w0, w1, i, 0, k, deg, rate)) / g.dx;
  }
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void shear_rate_final_kernel(const double *__restrict__ part, int nparts,
                                                                  double factor, double *out) {
  __shared__ double sh[kBlock / 64];
  double v = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) v += part[i];
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) *out = s * factor;
}

// The end of the step (sp_3d_ns.cxx:383-388): Solenoidal_uk_OBL
// (operate_omega.cxx:222-262), contra2co (fft_wrapper.h:327-361) and
// U_k2zeta_k_OBL (fft_wrapper.cxx:341-406), with uk_dc = co2contra of the
// k = 0 velocity.  48 B read, 80 B written per mode.
__global__ void project_zeta_obl_kernel(Grid g, double *u0, double *u1, double *u2, double *z0, double *z1,
                                        double *uk_dc, double deg, int store_u) {
  ModeO m;
  if (!mode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m)) return;
  const double kx = m.kx * g.wave_x, ky = m.ky * g.wave_y, kz = m.kz * g.wave_z;
  const double kxc = (1. + deg * deg) * m.kx * g.wave_x - deg * m.ky * g.wave_y;
  const double kyc = -deg * m.kx * g.wave_x + m.ky * g.wave_y;
  const double k2 = k2_obl(g, m, deg);
  const double ik2 = k2 > 0.0 ? 1. / k2 : 0.0;
  double re[3], iv[3];
  double *const u[3] = {u0, u1, u2};
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const long im = m.im + s;
    double v[3] = {u0[im], u1[im], u2[im]};
    const double dmy = ik2 * (v[0] * kx + v[1] * ky + v[2] * kz);
    v[0] -= dmy * kxc;
    v[1] -= dmy * kyc;
    v[2] -= dmy * kz;
    contra2co1(deg, v);
    for (int d = 0; d < 3; ++d) {
      if (store_u) u[d][im] = v[d];
      (s == 0 ? re : iv)[d] = v[d];
    }
  }
  if (m.im == 0) {
    double dc[3] = {re[0], re[1], re[2]};
    co2contra1(deg, dc);
    for (int d = 0; d < 3; ++d) uk_dc[d] = dc[d];
  }
  const double ks[3] = {kx, ky, kz};
  double rr[3], ri[3];
  rr[0] = (ks[1] * iv[2] - ks[2] * iv[1]);
  ri[0] = -(ks[1] * re[2] - ks[2] * re[1]);
  rr[1] = (ks[2] * iv[0] - ks[0] * iv[2]);
  ri[1] = -(ks[2] * re[0] - ks[0] * re[2]);
  rr[2] = (ks[0] * iv[1] - ks[1] * iv[0]);
  ri[2] = -(ks[0] * re[1] - ks[1] * re[0]);
  const int a = m.kx != 0 ? 1 : (m.kz != 0 ? 1 : 0);
  const int b = m.kx != 0 ? 2 : (m.kz != 0 ? 0 : 2);
  z0[m.im] = rr[a];
  z0[m.im + 1] = ri[a];
  z1[m.im] = rr[b];
  z1[m.im + 1] = ri[b];
}

inline unsigned nblk(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }
inline long nmodes(const spm_ctx *c) { return (long)c->grid.nx * c->grid.ny * c->grid.hnz; }
inline long npoints(const spm_ctx *c) { return (long)c->grid.nx * c->grid.ny * c->grid.nz; }

#define RCS(x)             \
  do {                     \
    const int rc_ = (x);   \
    if (rc_) return rc_;   \
  } while (0)

}  // namespace

int le_red_slots() { return kRedBlocks + 2; }

// Zeta_k2u_k_OBL of c->zeta into the spectra `u` (uk_dc's FIX_CELL
// components zeroed first where fix_mask says so).
int le_zeta2u_k(spm_ctx *c, double *const u[3], int fix_mask) {
  tbeg(c);
  hipLaunchKernelGGL(zeta2u_obl_kernel, dim3(nblk(nmodes(c))), dim3(kBlock), 0, c->stream, c->grid, c->zeta[0],
                     c->zeta[1], u[0], u[1], u[2], c->uk_dc, fix_mask, c->le_deg);
  tend(c, SPM_K_SPECTRAL, 80.0 * nmodes(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

// Solenoidal_uk_OBL + contra2co + U_k2zeta_k_OBL of the spectra in c->u.
int le_u_k2zeta_k(spm_ctx *c, bool store_u) {
  tbeg(c);
  hipLaunchKernelGGL(project_zeta_obl_kernel, dim3(nblk(nmodes(c))), dim3(kBlock), 0, c->stream, c->grid, c->u[0],
                     c->u[1], c->u[2], c->zeta[0], c->zeta[1], c->uk_dc, c->le_deg, store_u ? 1 : 0);
  tend(c, SPM_K_SPECTRAL, (store_u ? 128.0 : 80.0) * nmodes(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

// NS_solver_slavedEuler_Shear_OBL (fluid_solver.cxx:398-429) with
// Zeta_k2advection_k_OBL (operate_omega.cxx:115-173).  Scratch: f[0..2]
// (omega).
int le_ns_solver(spm_ctx *c) {
  const Grid &g = c->grid;
  const double deg = c->le_deg;
  RCS(k_truncate(c, c->zeta[0]));
  RCS(k_truncate(c, c->zeta[1]));
  double *const u3[3] = {c->u[0], c->u[1], c->u[2]};
  double *const w3[3] = {c->f[0], c->f[1], c->f[2]};
  tbeg(c);
  hipLaunchKernelGGL(zeta2uw_obl_kernel, dim3(nblk(nmodes(c))), dim3(kBlock), 0, c->stream, g, c->zeta[0],
                     c->zeta[1], u3[0], u3[1], u3[2], w3[0], w3[1], w3[2], c->uk_dc, deg);
  tend(c, SPM_K_SPECTRAL, 128.0 * nmodes(c));
  SPM_LAUNCH_CHECK(c);
  // zeta is truncated, so u and omega vanish outside the 2/3 box: x and y
  // passes, then one z pass that also forms u x omega in real space, then
  // the y and x passes of the product (only its kept modes are read on)
  double *const uw[6] = {u3[0], u3[1], u3[2], w3[0], w3[1], w3[2]};
  RCS(fft_c2r_band_xy(c, uw, 6));
  RCS(fz_cross_obl(c, u3, w3));
  RCS(fft_r2c_band_yx(c, u3, 3));
  tbeg(c);
  hipLaunchKernelGGL(advect_update_obl_kernel, dim3(nblk(nmodes(c))), dim3(kBlock), 0, c->stream, g, u3[0], u3[1],
                     u3[2], c->zeta[0], c->zeta[1], c->uk_dc, c->le_red + kRedBlocks, deg,
                     c->prm.eta * (1. / c->prm.rho), c->prm.dt);
  tend(c, SPM_K_SPECTRAL, 112.0 * nmodes(c) * ((double)(2 * g.trn_x - 1) / g.nx) *
                              ((double)(2 * g.trn_y - 1) / g.ny) * ((double)g.trn_z / g.hnz));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int le_oblique2u(spm_ctx *c, double *const ucp[3]) {
  tbeg(c);
  hipLaunchKernelGGL(oblique2u_kernel, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid, c->u[0], c->u[1],
                     c->u[2], ucp[0], ucp[1], ucp[2], c->le_deg, c->le_rate);
  tend(c, SPM_K_REALSPACE, 48.0 * npoints(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

// Time_evolution_hydro_OBL (sp_3d_ns.cxx:214-390).  The remap reads a copy
// of u held in up[0..2]; with particles the Cartesian velocity ucp lives in
// f[0..2].  Without particles f is zero, so U2u_oblique and Add_f_particle
// leave u unchanged and the particle block is skipped.
__global__ void strain_kernel(const double *__restrict__ sre, double *strain, double dt) { *strain += *sre * dt; }

int le_step(spm_ctx *c, int fix_mask, bool store_u) {
  const Grid &g = c->grid;
  RCS(le_ns_solver(c));
  double *const u3[3] = {c->u[0], c->u[1], c->u[2]};
  RCS(le_zeta2u_k(c, u3, fix_mask));
  RCS(fft_c2r_many_band(c, u3, 3));  // the updated zeta is still zero outside the box
  // Shear_rate_eff = Shear_rate; degree_oblique += Shear_rate_eff * dt
  const double rate = c->le_rate;
  c->le_deg += rate * c->prm.dt;
  if (c->le_deg >= 1.) {
    double *const copy[3] = {c->up[0], c->up[1], c->up[2]};
    const size_t fb = sizeof(double) * c->nfield;
    for (int d = 0; d < 3; ++d) SPM_HIP(c, hipMemcpyAsync(copy[d], u3[d], fb, hipMemcpyDeviceToDevice, c->stream));
    tbeg(c);
    hipLaunchKernelGGL(remap_kernel, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, g, copy[0], copy[1], copy[2],
                       u3[0], u3[1], u3[2]);
    tend(c, SPM_K_REALSPACE, 48.0 * npoints(c));
    SPM_LAUNCH_CHECK(c);
    c->le_deg -= 1.;
  }
  // Calc_shear_rate_eff: ivolume = Ivolume DX^3, Ivolume = 1 / (LX LY LZ)
  const double ivolume = (1. / (g.lx * g.ly * g.lz)) * (g.dx * g.dx * g.dx);
  tbeg(c);
  hipLaunchKernelGGL(shear_rate_rows_kernel, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, g, u3[0], u3[1],
                     c->le_deg, rate, c->le_red);
  hipLaunchKernelGGL(shear_rate_final_kernel, dim3(1), dim3(kBlock), 0, c->stream, c->le_red, kRedBlocks,
                     ivolume * g.ny / (g.ny - 1), c->le_red + kRedBlocks);
  tend(c, SPM_K_REALSPACE, 32.0 * g.nx * g.nz);
  // Shear_strain_realized += Shear_rate_eff dt (sp_3d_ns.cxx:600), on the device
  hipLaunchKernelGGL(strain_kernel, dim3(1), dim3(1), 0, c->stream, c->le_red + kRedBlocks, c->le_strain,
                     c->prm.dt);
  SPM_LAUNCH_CHECK(c);
  if (c->np > 0) {  // sp_3d_ns.cxx:343-373
    const double *sre = c->le_red + kRedBlocks;
    double *const ucp[3] = {c->f[0], c->f[1], c->f[2]};
    const int euler = c->ts == 0;
    RCS(le_oblique2u(c, ucp));
    RCS(p_position_obl(c, euler, sre));
    RCS(p_calc_f_hydro_obl(c, ucp, sre));
    RCS(p_velocity_obl(c, euler));
    if (c->prm.kbt > 0.) RCS(host_thermostat(c));
    RCS(p_make_phi_obl(c, sre));
    RCS(p_f_add_obl(c, ucp));
  }
  RCS(fft_r2c_many(c, u3, 3));
  return le_u_k2zeta_k(c, store_u);
}

}  // namespace spm
