// electrolyte.hip -- kernels of the charged-colloid (Electrolyte) step.
//
// The reference's SW_EQ == Electrolyte branch (sp_3d_ns.cxx:46-57,89-121)
// adds Nernst-Planck ion transport and a Poisson solve to the SPM step:
//   Rhs_NS_Nernst_Planck  fluid_solver.cxx:243-280
//   Rhs_NS_solute         fluid_solver.cxx:181-227, solute_rhs.cxx:42-235
//   NSsolute_solver_Euler fluid_solver.cxx:306-326
//   Rescale_solute        solute_rhs.cxx:97-126
//   Make_Coulomb_force_x_on_fluid  operate_electrolyte.cxx:90-177
// The host sequence is in spm_api.cpp; the particle-shaped kernels
// (charge profile, surface normals) are in particles.hip.  All FP64.
//
// k-space kernels take one complex mode (i, j, kz) per thread, real-space
// kernels one grid point (k < NZ) per thread.  Grid sums (|surface charge|,
// fluid volume, ion counts) are two-pass reductions in a fixed tree order,
// so results do not depend on scheduling.
#include "spm_internal.h"

namespace spm {

namespace {

constexpr int kBlock = 256;

struct EMode {
  long im;
  double kx, ky, kz;  // physical wavenumbers
  bool keep;          // inside the 2/3-rule box (the ijk_range blocks)
};

__device__ __forceinline__ bool emode_of(const Grid &g, long t, EMode &m, double &k2) {
  const long nmodes = (long)g.nx * g.ny * g.hnz;
  if (t >= nmodes) return false;
  const int kk = (int)(t % g.hnz);
  const long ij = t / g.hnz;
  const int j = (int)(ij % g.ny);
  const int i = (int)(ij / g.ny);
  const int ix = i > g.nx / 2 ? i - g.nx : i;
  const int iy = j > g.ny / 2 ? j - g.ny : j;
  m.im = ij * g.nz_ + 2 * kk;
  m.kx = (double)ix * g.wave_x;
  m.ky = (double)iy * g.wave_y;
  m.kz = (double)kk * g.wave_z;
  m.keep = abs(ix) < g.trn_x && abs(iy) < g.trn_y && kk < g.trn_z;
  // K2 as the reference's table (fft_wrapper.cxx:51-58)
  k2 = (g.wave_x * ix) * (g.wave_x * ix) + (g.wave_y * iy) * (g.wave_y * iy) + (g.wave_z * kk) * (g.wave_z * kk);
  return true;
}

__device__ __forceinline__ bool point_of(const Grid &g, long t, long &im) {
  const long npts = (long)g.nx * g.ny * g.nz;
  if (t >= npts) return false;
  const int k = (int)(t % g.nz);
  im = (t / g.nz) * g.nz_ + k;
  return true;
}

// Charge_field_k2Coulomb_potential_k_PBC (operate_electrolyte.cxx:76-88),
// optionally Truncate_two_third_rule_ooura, then A_k2da_k
// (fft_wrapper.cxx:103-135) and Add_constant_field_k (fluid_solver.cxx:
// 228-241) on mode 0.  16 B read, 48 B written per mode.
__global__ void poisson_grad_kernel(Grid g, const double *__restrict__ pot, double ieps, int trunc, double *g0,
                                    double *g1, double *g2, double c0, double c1, double c2) {
  EMode m;
  double k2;
  if (!emode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m, k2)) return;
  const double ik2 = k2 > 0.0 ? 1.0 / k2 : 0.0;
  double re = pot[m.im], im = pot[m.im + 1];
  re *= (ik2 * ieps);
  im *= (ik2 * ieps);
  if (trunc && !m.keep) re = im = 0.0;
  g0[m.im] = m.kx * im;
  g0[m.im + 1] = -m.kx * re;
  g1[m.im] = m.ky * im;
  g1[m.im + 1] = -m.ky * re;
  g2[m.im] = m.kz * im;
  g2[m.im + 1] = -m.kz * re;
  if (m.im == 0) {
    g0[0] -= c0;
    g1[0] -= c1;
    g2[0] -= c2;
  }
}

// Diffusion_flux_single (solute_rhs.cxx:216-235) in k space: the gradient
// of kBT Onsager C.  16 B read, 48 B written per mode.
__global__ void grad_scaled_kernel(Grid g, const double *__restrict__ a, double scale, double *g0, double *g1,
                                   double *g2) {
  EMode m;
  double k2;
  if (!emode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m, k2)) return;
  const double re = scale * a[m.im], im = scale * a[m.im + 1];
  g0[m.im] = m.kx * im;
  g0[m.im + 1] = -m.kx * re;
  g1[m.im] = m.ky * im;
  g1[m.im + 1] = -m.ky * re;
  g2[m.im] = m.kz * im;
  g2[m.im + 1] = -m.kz * re;
}

// U_k2divergence_k (fft_wrapper.cxx:649-677).  48 B read, 16 B written.
__global__ void divergence_kernel(Grid g, const double *__restrict__ u0, const double *__restrict__ u1,
                                  const double *__restrict__ u2, double *div) {
  EMode m;
  double k2;
  if (!emode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m, k2)) return;
  const long a = m.im, b = m.im + 1;
  const double d0 = m.kx * u0[b] + m.ky * u1[b] + m.kz * u2[b];
  const double d1 = -m.kx * u0[a] - m.ky * u1[a] - m.kz * u2[a];
  div[a] = d0;
  div[b] = d1;
}

// Add_zeta_viscous_term (operate_omega.cxx:172-184) and Field_solver_Euler
// (fluid_solver.cxx:122-135) over the ijk_range blocks, for zeta and the
// ion spectra.  Modes outside the blocks are left as they are.
__global__ void nssolute_update_kernel(Grid g, double *z0, double *z1, const double *__restrict__ r0,
                                       const double *__restrict__ r1, double *c0, double *c1,
                                       const double *__restrict__ q0, const double *__restrict__ q1, int nspec,
                                       double nu, double dt, int fluid) {
  EMode m;
  double k2;
  if (!emode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m, k2) || !m.keep) return;
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const long t = m.im + e;
    if (fluid) {
      z0[t] += (dt * (r0[t] + -(nu * k2 * z0[t])));
      z1[t] += (dt * (r1[t] + -(nu * k2 * z1[t])));
    }
    c0[t] += (dt * q0[t]);
    if (nspec > 1) c1[t] += (dt * q1[t]);
  }
}

// Conc_k2charge_field's accumulation (operate_electrolyte.cxx:66-72):
// charge += Z e C (1 - phi).  24 B read, 8 B written per point.
__global__ void charge_add_kernel(Grid g, double *charge, const double *__restrict__ conc,
                                  const double *__restrict__ phi, double ze) {
  long im;
  if (!point_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, im)) return;
  charge[im] += ze * conc[im] * (1. - phi[im]);
}

// Solute_solver_rhs_nonlinear_x_single (solute_rhs.cxx:149-187),
// Solute_impermeability (:42-95) and Add_advection_flux (:188-207) on one
// point.  The reference projects the flux once per particle whose stencil
// holds the point; a projection is idempotent, so once is the same field
// up to rounding.  88 B read, 24 B written per point.
__global__ void np_flux_kernel(Grid g, double *f0, double *f1, double *f2, const double *__restrict__ cx,
                               const double *__restrict__ gp0, const double *__restrict__ gp1,
                               const double *__restrict__ gp2, const double *__restrict__ n0,
                               const double *__restrict__ n1, const double *__restrict__ n2,
                               const double *__restrict__ u0, const double *__restrict__ u1,
                               const double *__restrict__ u2, double onsager, double ze) {
  long im;
  if (!point_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, im)) return;
  const double c = cx[im];
  double fl[3] = {f0[im], f1[im], f2[im]};
  const double gp[3] = {gp0[im], gp1[im], gp2[im]};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double dmy_interaction = gp[d] * ze;
    fl[d] += c * (onsager * dmy_interaction);
  }
  const double nv[3] = {n0[im], n1[im], n2[im]};
  const double norm2 = nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2];
  if (norm2 > 0.) {
    double dmy = 0.;
#pragma unroll
    for (int d = 0; d < 3; ++d) dmy += fl[d] * nv[d];
    dmy /= norm2;
#pragma unroll
    for (int d = 0; d < 3; ++d) fl[d] -= nv[d] * dmy;
  }
  const double uu[3] = {u0[im], u1[im], u2[im]};
#pragma unroll
  for (int d = 0; d < 3; ++d) fl[d] += c * -uu[d];
  f0[im] = fl[0];
  f1[im] = fl[1];
  f2[im] = fl[2];
}

// Make_Coulomb_force_x_on_fluid's last loops (operate_electrolyte.cxx:
// 125-176): f = charge (-grad Phi + E_ext).  32 B read, 24 B written.
__global__ void coulomb_kernel(Grid g, const double *__restrict__ charge, double *f0, double *f1, double *f2,
                               double e0, double e1, double e2, int external) {
  long im;
  if (!point_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, im)) return;
  const double q = charge[im];
  double ef0 = -f0[im], ef1 = -f1[im], ef2 = -f2[im];
  if (external) {
    ef0 += e0;
    ef1 += e1;
    ef2 += e2;
  }
  f0[im] = q * ef0;
  f1[im] = q * ef1;
  f2[im] = q * ef2;
}

// u += f dt/rho (sp_3d_ns.cxx:108-119).
__global__ void add_scaled_kernel(Grid g, double *u0, double *u1, double *u2, const double *__restrict__ f0,
                                  const double *__restrict__ f1, const double *__restrict__ f2, double dmy) {
  long im;
  if (!point_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, im)) return;
  u0[im] += (f0[im] * dmy);
  u1[im] += (f1[im] * dmy);
  u2[im] += (f2[im] * dmy);
}

// a = v on every real-space point (k < NZ; the padding is left alone).
// Charge_field_k2Coulomb_potential_k_PBC (operate_electrolyte.cxx:76-88):
// the potential spectrum rho_k / (eps k^2), zero at k = 0.  16 B read and
// written per mode.
__global__ void coulomb_potential_kernel(Grid g, double *pot, double ieps) {
  EMode m;
  double k2;
  if (!emode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m, k2)) return;
  const double f = (k2 > 0.0 ? 1.0 / k2 : 0.0) * ieps;
  pot[m.im] *= f;
  pot[m.im + 1] *= f;
}

// Calc_free_energy_PB (operate_electrolyte.cxx:340-355): half the potential,
// less the applied field's E.x (Add_external_electric_field_x, :16-43)
__global__ void half_potential_kernel(Grid g, double *pot, double e0, double e1, double e2, int external) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  const int k = (int)(t % g.nz), j = (int)((t / g.nz) % g.ny), i = (int)(t / ((long)g.nz * g.ny));
  const long im = ((long)i * g.ny + j) * g.nz_ + k;
  double v = pot[im] * 0.5;
  if (external) v -= ((e0 * (double)i + e1 * (double)j + e2 * (double)k) * g.dx);
  pot[im] = v;
}

__global__ void fill_real_kernel(Grid g, double *a, double v) {
  long im;
  if (!point_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, im)) return;
  a[im] = v;
}

// a *= *factor over n elements.
__global__ void scale_by_kernel(double *a, long n, const double *__restrict__ factor) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] *= *factor;
}

// ---- reductions over k < NZ: pass 1 (kRedBlocks partial sums) ---------
constexpr int kRedBlocks = 1024;

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

// mode 0: sum |a|; 1: sum (1 - phi) a; 2: sum (1 - phi);
// 3: sum (1 - phi) a (log a - 1) (Calc_free_energy_PB's ideal-gas term,
// operate_electrolyte.cxx:318-330); 4: sum a phi (its electrostatic term,
// phi then holds the potential, :356-363)
__global__ __launch_bounds__(kBlock) void reduce_partial_kernel(Grid g, int mode, const double *__restrict__ a,
                                                                const double *__restrict__ phi, double *part) {
  __shared__ double sh[kBlock / 64];
  const long npts = (long)g.nx * g.ny * g.nz;
  double v = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < npts; t += (long)gridDim.x * blockDim.x) {
    const long im = (t / g.nz) * g.nz_ + t % g.nz;
    switch (mode) {
      case 0: v += fabs(a[im]); break;
      case 1: v += (1. - phi[im]) * a[im]; break;
      case 2: v += (1. - phi[im]); break;
      case 3: v += (1. - phi[im]) * a[im] * (log(a[im]) - 1.); break;
      default: v += a[im] * phi[im]; break;
    }
  }
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// pass 2: out = numer / (sum dx^3) if ratio, else sum dx^3
__global__ __launch_bounds__(kBlock) void reduce_final_kernel(const double *__restrict__ part, int nparts,
                                                              double dx3, int ratio, double numer, double *out) {
  __shared__ double sh[kBlock / 64];
  double v = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) v += part[i];
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) *out = ratio ? numer / (s * dx3) : s * dx3;
}

inline unsigned nblk(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }
inline long nmodes(const spm_ctx *c) { return (long)c->grid.nx * c->grid.ny * c->grid.hnz; }
inline long npoints(const spm_ctx *c) { return (long)c->grid.nx * c->grid.ny * c->grid.nz; }

}  // namespace

// ---------------------------------------------------------------------
// launchers (the reduction scratch c->red holds kRedBlocks partials and
// then result slots)
// ---------------------------------------------------------------------
int e_red_slots() { return kRedBlocks + 16; }

int e_poisson_grad(spm_ctx *c, const double *pot, double *const grad[3], int trunc, const double cst[3]) {
  tbeg(c);
  hipLaunchKernelGGL(poisson_grad_kernel, dim3(nblk(nmodes(c))), dim3(kBlock), 0, c->stream, c->grid, pot,
                     1. / c->el.dielectric, trunc, grad[0], grad[1], grad[2], cst[0], cst[1], cst[2]);
  tend(c, SPM_K_SPECTRAL, 64.0 * nmodes(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_grad_scaled(spm_ctx *c, const double *a, double scale, double *const out[3]) {
  tbeg(c);
  hipLaunchKernelGGL(grad_scaled_kernel, dim3(nblk(nmodes(c))), dim3(kBlock), 0, c->stream, c->grid, a, scale,
                     out[0], out[1], out[2]);
  tend(c, SPM_K_SPECTRAL, 64.0 * nmodes(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_divergence(spm_ctx *c, double *const flux[3], double *div) {
  tbeg(c);
  hipLaunchKernelGGL(divergence_kernel, dim3(nblk(nmodes(c))), dim3(kBlock), 0, c->stream, c->grid, flux[0],
                     flux[1], flux[2], div);
  tend(c, SPM_K_SPECTRAL, 64.0 * nmodes(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_coulomb_potential(spm_ctx *c, double *pot) {
  hipLaunchKernelGGL(coulomb_potential_kernel, dim3(nblk(nmodes(c))), dim3(kBlock), 0, c->stream, c->grid, pot,
                     1.0 / c->el.dielectric);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_half_potential(spm_ctx *c, double *pot, const double ext[3], int external) {
  hipLaunchKernelGGL(half_potential_kernel, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid, pot, ext[0],
                     ext[1], ext[2], external);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_nssolute_update(spm_ctx *c, double nu, double dt, bool fluid) {
  tbeg(c);
  hipLaunchKernelGGL(nssolute_update_kernel, dim3(nblk(nmodes(c))), dim3(kBlock), 0, c->stream, c->grid,
                     c->zeta[0], c->zeta[1], c->f_ns1[0], c->f_ns1[1], c->conc[0],
                     c->el.n_spec > 1 ? c->conc[1] : c->conc[0], c->conc_rhs[0],
                     c->el.n_spec > 1 ? c->conc_rhs[1] : c->conc_rhs[0], c->el.n_spec, nu, dt, fluid ? 1 : 0);
  tend(c, SPM_K_SPECTRAL, ((fluid ? 64.0 : 0.0) + 48.0 * c->el.n_spec) * nmodes(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_charge_add(spm_ctx *c, double *charge, const double *conc, const double *phi, double ze) {
  tbeg(c);
  hipLaunchKernelGGL(charge_add_kernel, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid, charge, conc,
                     phi, ze);
  tend(c, SPM_K_REALSPACE, 32.0 * npoints(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_np_flux(spm_ctx *c, double *const flux[3], const double *cx, double *const grad[3], double *const normal[3],
              double *const u[3], double onsager, double ze) {
  tbeg(c);
  hipLaunchKernelGGL(np_flux_kernel, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid, flux[0], flux[1],
                     flux[2], cx, grad[0], grad[1], grad[2], normal[0], normal[1], normal[2], u[0], u[1], u[2],
                     onsager, ze);
  tend(c, SPM_K_REALSPACE, 112.0 * npoints(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_coulomb(spm_ctx *c, const double *charge, double *const f[3], const double ext[3], int external) {
  tbeg(c);
  hipLaunchKernelGGL(coulomb_kernel, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid, charge, f[0], f[1],
                     f[2], ext[0], ext[1], ext[2], external);
  tend(c, SPM_K_REALSPACE, 56.0 * npoints(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_add_scaled(spm_ctx *c, double *const u[3], double *const f[3], double dmy) {
  tbeg(c);
  hipLaunchKernelGGL(add_scaled_kernel, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid, u[0], u[1],
                     u[2], f[0], f[1], f[2], dmy);
  tend(c, SPM_K_REALSPACE, 72.0 * npoints(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_fill_real(spm_ctx *c, double *a, double v) {
  tbeg(c);
  hipLaunchKernelGGL(fill_real_kernel, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid, a, v);
  tend(c, SPM_K_REALSPACE, 8.0 * npoints(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

// result slot `slot` of c->red = sum dx^3, or numer / (sum dx^3) if ratio
int e_reduce(spm_ctx *c, int mode, const double *a, const double *phi, int slot, int ratio, double numer) {
  tbeg(c);
  hipLaunchKernelGGL(reduce_partial_kernel, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, c->grid, mode, a, phi,
                     c->red);
  hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(kBlock), 0, c->stream, c->red, kRedBlocks, c->pm.dx3, ratio,
                     numer, c->red + kRedBlocks + slot);
  tend(c, SPM_K_REALSPACE, (mode == 1 ? 16.0 : 8.0) * npoints(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_scale_by(spm_ctx *c, double *a, long n, int slot) {
  tbeg(c);
  hipLaunchKernelGGL(scale_by_kernel, dim3(nblk(n)), dim3(kBlock), 0, c->stream, a, n, c->red + kRedBlocks + slot);
  tend(c, SPM_K_REALSPACE, 16.0 * n);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_read_slot(spm_ctx *c, int slot, double *out) {
  SPM_HIP(c, hipMemcpyAsync(out, c->red + kRedBlocks + slot, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

}  // namespace spm
