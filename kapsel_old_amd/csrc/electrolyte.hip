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
// (charge profile, surface normals) are in particles.hip.
//
// Precision: the transport fields (ion spectra and their right-hand sides,
// fluxes, velocity, vorticity, phi) are of the context's field type T (FP64
// or FP32); the Poisson side (charge density, potential, its gradient, the
// particle charge profile, surface normals) is FP64 in both.  Kernels that
// mix the two compute in FP64 and store T.
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
template <typename T>
__global__ void grad_scaled_kernel(Grid g, const T *__restrict__ a, double scale, T *g0, T *g1, T *g2) {
  EMode m;
  double k2;
  if (!emode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m, k2)) return;
  const double re = scale * (double)a[m.im], im = scale * (double)a[m.im + 1];
  g0[m.im] = (T)(m.kx * im);
  g0[m.im + 1] = (T)(-m.kx * re);
  g1[m.im] = (T)(m.ky * im);
  g1[m.im + 1] = (T)(-m.ky * re);
  g2[m.im] = (T)(m.kz * im);
  g2[m.im + 1] = (T)(-m.kz * re);
}

// U_k2divergence_k (fft_wrapper.cxx:649-677).  48 B read, 16 B written.
template <typename T>
__global__ void divergence_kernel(Grid g, const T *__restrict__ u0, const T *__restrict__ u1,
                                  const T *__restrict__ u2, T *div) {
  EMode m;
  double k2;
  if (!emode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m, k2)) return;
  const long a = m.im, b = m.im + 1;
  const double d0 = m.kx * (double)u0[b] + m.ky * (double)u1[b] + m.kz * (double)u2[b];
  const double d1 = -m.kx * (double)u0[a] - m.ky * (double)u1[a] - m.kz * (double)u2[a];
  div[a] = (T)d0;
  div[b] = (T)d1;
}

// Add_zeta_viscous_term (operate_omega.cxx:172-184) and Field_solver_Euler
// (fluid_solver.cxx:122-135) over the ijk_range blocks, for zeta and the
// ion spectra.  Modes outside the blocks are left as they are.
template <typename T>
__global__ void nssolute_update_kernel(Grid g, T *z0, T *z1, const T *__restrict__ r0, const T *__restrict__ r1,
                                       T *c0, T *c1, const T *__restrict__ q0, const T *__restrict__ q1, int nspec,
                                       double nu, double dt, int fluid) {
  EMode m;
  double k2;
  if (!emode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m, k2) || !m.keep) return;
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const long t = m.im + e;
    if (fluid) {
      const double a0 = z0[t], a1 = z1[t];
      z0[t] = (T)(a0 + (dt * ((double)r0[t] + -(nu * k2 * a0))));
      z1[t] = (T)(a1 + (dt * ((double)r1[t] + -(nu * k2 * a1))));
    }
    c0[t] = (T)((double)c0[t] + (dt * (double)q0[t]));
    if (nspec > 1) c1[t] = (T)((double)c1[t] + (dt * (double)q1[t]));
  }
}

// Conc_k2charge_field's accumulation (operate_electrolyte.cxx:66-72):
// charge += Z e C (1 - phi).  24 B read, 8 B written per point.
template <typename T>
__global__ void charge_add_kernel(Grid g, double *charge, const T *__restrict__ conc,
                                  const double *__restrict__ phi, double ze) {
  long im;
  if (!point_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, im)) return;
  charge[im] += ze * (double)conc[im] * (1. - phi[im]);
}

// Solute_solver_rhs_nonlinear_x_single (solute_rhs.cxx:149-187),
// Solute_impermeability (:42-95) and Add_advection_flux (:188-207) on one
// point.  The reference projects the flux once per particle whose stencil
// holds the point; a projection is idempotent, so once is the same field
// up to rounding.  88 B read, 24 B written per point.
template <typename T>
__global__ void np_flux_kernel(Grid g, T *f0, T *f1, T *f2, const T *__restrict__ cx,
                               const double *__restrict__ gp0, const double *__restrict__ gp1,
                               const double *__restrict__ gp2, const double *__restrict__ n0,
                               const double *__restrict__ n1, const double *__restrict__ n2,
                               const T *__restrict__ u0, const T *__restrict__ u1, const T *__restrict__ u2,
                               double onsager, double ze) {
  long im;
  if (!point_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, im)) return;
  const double c = cx[im];
  double fl[3] = {(double)f0[im], (double)f1[im], (double)f2[im]};
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
  const double uu[3] = {(double)u0[im], (double)u1[im], (double)u2[im]};
#pragma unroll
  for (int d = 0; d < 3; ++d) fl[d] += c * -uu[d];
  f0[im] = (T)fl[0];
  f1[im] = (T)fl[1];
  f2[im] = (T)fl[2];
}

// Make_Coulomb_force_x_on_fluid's last loops (operate_electrolyte.cxx:
// 125-176): f = charge (-grad Phi + E_ext).  32 B read, 24 B written.
// The potential gradient (FP64, gp) and the force (T, f) may be the same
// FP64 field.
template <typename T>
__global__ void coulomb_kernel(Grid g, const double *__restrict__ charge, const double *gp0, const double *gp1,
                               const double *gp2, T *f0, T *f1, T *f2, double e0, double e1, double e2,
                               int external) {
  long im;
  if (!point_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, im)) return;
  const double q = charge[im];
  double ef0 = -gp0[im], ef1 = -gp1[im], ef2 = -gp2[im];
  if (external) {
    ef0 += e0;
    ef1 += e1;
    ef2 += e2;
  }
  f0[im] = (T)(q * ef0);
  f1[im] = (T)(q * ef1);
  f2[im] = (T)(q * ef2);
}

// FP64 real field -> field of type T (the potential gradient the step's
// Calc_f_hydro reads from f)
template <typename T>
__global__ void to_t_kernel(Grid g, const double *__restrict__ a, T *b) {
  long im;
  if (!point_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, im)) return;
  b[im] = (T)a[im];
}

// u += f dt/rho (sp_3d_ns.cxx:108-119).
template <typename T>
__global__ void add_scaled_kernel(Grid g, T *u0, T *u1, T *u2, const T *__restrict__ f0, const T *__restrict__ f1,
                                  const T *__restrict__ f2, double dmy) {
  long im;
  if (!point_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, im)) return;
  u0[im] = (T)((double)u0[im] + ((double)f0[im] * dmy));
  u1[im] = (T)((double)u1[im] + ((double)f1[im] * dmy));
  u2[im] = (T)((double)u2[im] + ((double)f2[im] * dmy));
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

template <typename T>
__global__ void fill_real_kernel(Grid g, T *a, double v) {
  long im;
  if (!point_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, im)) return;
  a[im] = (T)v;
}

// a *= *factor over n elements.
template <typename T>
__global__ void scale_by_kernel(T *a, long n, const double *__restrict__ factor) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = (T)((double)a[t] * *factor);
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
// phi then holds the potential, :356-363).  Modes 0 and 4 read FP64 fields
// (surface charge; charge and potential), modes 1-3 fields of type T
// (ion densities, phi).
template <typename T>
__global__ __launch_bounds__(kBlock) void reduce_partial_kernel(Grid g, int mode, const void *__restrict__ av,
                                                                const void *__restrict__ phiv, double *part) {
  __shared__ double sh[kBlock / 64];
  const long npts = (long)g.nx * g.ny * g.nz;
  const double *a64 = static_cast<const double *>(av), *p64 = static_cast<const double *>(phiv);
  const T *a = static_cast<const T *>(av), *phi = static_cast<const T *>(phiv);
  double v = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < npts; t += (long)gridDim.x * blockDim.x) {
    const long im = (t / g.nz) * g.nz_ + t % g.nz;
    switch (mode) {
      case 0: v += fabs(a64[im]); break;
      case 1: v += (1. - (double)phi[im]) * (double)a[im]; break;
      case 2: v += (1. - (double)phi[im]); break;
      case 3: v += (1. - (double)phi[im]) * (double)a[im] * (log((double)a[im]) - 1.); break;
      default: v += a64[im] * p64[im]; break;
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

template <typename T>
inline T *TP(double *p) {
  return reinterpret_cast<T *>(p);
}
inline double tsz(const spm_ctx *c) { return c->prm.precision == SPM_FP32 ? 4.0 : 8.0; }

int e_grad_scaled(spm_ctx *c, const double *a, double scale, double *const out[3]) {
  tbeg(c);
  if (c->prm.precision == SPM_FP32)
    hipLaunchKernelGGL(grad_scaled_kernel<float>, dim3(nblk(nmodes(c))), dim3(kBlock), 0, c->stream, c->grid,
                       reinterpret_cast<const float *>(a), scale, TP<float>(out[0]), TP<float>(out[1]),
                       TP<float>(out[2]));
  else
    hipLaunchKernelGGL(grad_scaled_kernel<double>, dim3(nblk(nmodes(c))), dim3(kBlock), 0, c->stream, c->grid, a,
                       scale, out[0], out[1], out[2]);
  tend(c, SPM_K_SPECTRAL, 8.0 * tsz(c) * nmodes(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_divergence(spm_ctx *c, double *const flux[3], double *div) {
  tbeg(c);
  if (c->prm.precision == SPM_FP32)
    hipLaunchKernelGGL(divergence_kernel<float>, dim3(nblk(nmodes(c))), dim3(kBlock), 0, c->stream, c->grid,
                       TP<float>(flux[0]), TP<float>(flux[1]), TP<float>(flux[2]), TP<float>(div));
  else
    hipLaunchKernelGGL(divergence_kernel<double>, dim3(nblk(nmodes(c))), dim3(kBlock), 0, c->stream, c->grid,
                       flux[0], flux[1], flux[2], div);
  tend(c, SPM_K_SPECTRAL, 8.0 * tsz(c) * nmodes(c));
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
  double *c1 = c->el.n_spec > 1 ? c->conc[1] : c->conc[0];
  double *q1 = c->el.n_spec > 1 ? c->conc_rhs[1] : c->conc_rhs[0];
  if (c->prm.precision == SPM_FP32)
    hipLaunchKernelGGL(nssolute_update_kernel<float>, dim3(nblk(nmodes(c))), dim3(kBlock), 0, c->stream, c->grid,
                       TP<float>(c->zeta[0]), TP<float>(c->zeta[1]), TP<float>(c->f_ns1[0]), TP<float>(c->f_ns1[1]),
                       TP<float>(c->conc[0]), TP<float>(c1), TP<float>(c->conc_rhs[0]), TP<float>(q1), c->el.n_spec,
                       nu, dt, fluid ? 1 : 0);
  else
    hipLaunchKernelGGL(nssolute_update_kernel<double>, dim3(nblk(nmodes(c))), dim3(kBlock), 0, c->stream, c->grid,
                       c->zeta[0], c->zeta[1], c->f_ns1[0], c->f_ns1[1], c->conc[0], c1, c->conc_rhs[0], q1,
                       c->el.n_spec, nu, dt, fluid ? 1 : 0);
  tend(c, SPM_K_SPECTRAL, ((fluid ? 8.0 : 0.0) + 6.0 * c->el.n_spec) * tsz(c) * nmodes(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_charge_add(spm_ctx *c, double *charge, const double *conc, const double *phi, double ze) {
  tbeg(c);
  if (c->prm.precision == SPM_FP32)
    hipLaunchKernelGGL(charge_add_kernel<float>, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid, charge,
                       reinterpret_cast<const float *>(conc), phi, ze);
  else
    hipLaunchKernelGGL(charge_add_kernel<double>, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid, charge,
                       conc, phi, ze);
  tend(c, SPM_K_REALSPACE, (24.0 + tsz(c)) * npoints(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_np_flux(spm_ctx *c, double *const flux[3], const double *cx, double *const grad[3], double *const normal[3],
              double *const u[3], double onsager, double ze) {
  tbeg(c);
  if (c->prm.precision == SPM_FP32)
    hipLaunchKernelGGL(np_flux_kernel<float>, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid,
                       TP<float>(flux[0]), TP<float>(flux[1]), TP<float>(flux[2]),
                       reinterpret_cast<const float *>(cx), grad[0], grad[1], grad[2], normal[0], normal[1],
                       normal[2], TP<float>(u[0]), TP<float>(u[1]), TP<float>(u[2]), onsager, ze);
  else
    hipLaunchKernelGGL(np_flux_kernel<double>, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid, flux[0],
                       flux[1], flux[2], cx, grad[0], grad[1], grad[2], normal[0], normal[1], normal[2], u[0], u[1],
                       u[2], onsager, ze);
  tend(c, SPM_K_REALSPACE, (48.0 + 10.0 * tsz(c)) * npoints(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_coulomb(spm_ctx *c, const double *charge, double *const grad[3], double *const f[3], const double ext[3],
              int external) {
  tbeg(c);
  if (c->prm.precision == SPM_FP32)
    hipLaunchKernelGGL(coulomb_kernel<float>, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid, charge,
                       grad[0], grad[1], grad[2], TP<float>(f[0]), TP<float>(f[1]), TP<float>(f[2]), ext[0], ext[1],
                       ext[2], external);
  else
    hipLaunchKernelGGL(coulomb_kernel<double>, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid, charge,
                       grad[0], grad[1], grad[2], f[0], f[1], f[2], ext[0], ext[1], ext[2], external);
  tend(c, SPM_K_REALSPACE, (32.0 + 3.0 * tsz(c)) * npoints(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_to_t(spm_ctx *c, double *const a[3], double *const b[3]) {
  tbeg(c);
  for (int d = 0; d < 3; ++d) {
    if (c->prm.precision == SPM_FP32)
      hipLaunchKernelGGL(to_t_kernel<float>, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid, a[d],
                         TP<float>(b[d]));
    else
      hipLaunchKernelGGL(to_t_kernel<double>, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid, a[d],
                         b[d]);
  }
  tend(c, SPM_K_REALSPACE, 3.0 * (8.0 + tsz(c)) * npoints(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_add_scaled(spm_ctx *c, double *const u[3], double *const f[3], double dmy) {
  tbeg(c);
  if (c->prm.precision == SPM_FP32)
    hipLaunchKernelGGL(add_scaled_kernel<float>, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid,
                       TP<float>(u[0]), TP<float>(u[1]), TP<float>(u[2]), TP<float>(f[0]), TP<float>(f[1]),
                       TP<float>(f[2]), dmy);
  else
    hipLaunchKernelGGL(add_scaled_kernel<double>, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid, u[0],
                       u[1], u[2], f[0], f[1], f[2], dmy);
  tend(c, SPM_K_REALSPACE, 9.0 * tsz(c) * npoints(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_fill_real(spm_ctx *c, double *a, double v) {
  tbeg(c);
  if (c->prm.precision == SPM_FP32)
    hipLaunchKernelGGL(fill_real_kernel<float>, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid,
                       TP<float>(a), v);
  else
    hipLaunchKernelGGL(fill_real_kernel<double>, dim3(nblk(npoints(c))), dim3(kBlock), 0, c->stream, c->grid, a, v);
  tend(c, SPM_K_REALSPACE, tsz(c) * npoints(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

// result slot `slot` of c->red = sum dx^3, or numer / (sum dx^3) if ratio
int e_reduce(spm_ctx *c, int mode, const double *a, const double *phi, int slot, int ratio, double numer) {
  tbeg(c);
  if (c->prm.precision == SPM_FP32)
    hipLaunchKernelGGL(reduce_partial_kernel<float>, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, c->grid, mode,
                       (const void *)a, (const void *)phi, c->red);
  else
    hipLaunchKernelGGL(reduce_partial_kernel<double>, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, c->grid, mode,
                       (const void *)a, (const void *)phi, c->red);
  hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(kBlock), 0, c->stream, c->red, kRedBlocks, c->pm.dx3, ratio,
                     numer, c->red + kRedBlocks + slot);
  tend(c, SPM_K_REALSPACE, (mode == 1 || mode == 3 ? 2.0 : 1.0) * (mode == 0 || mode == 4 ? 8.0 : tsz(c)) * npoints(c));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

// a *= factor over the first n elements of a field of type T (f64: FP64)
int e_scale_by(spm_ctx *c, double *a, long n, int slot, bool f64) {
  tbeg(c);
  if (c->prm.precision == SPM_FP32 && !f64)
    hipLaunchKernelGGL(scale_by_kernel<float>, dim3(nblk(n)), dim3(kBlock), 0, c->stream, TP<float>(a), n,
                       c->red + kRedBlocks + slot);
  else
    hipLaunchKernelGGL(scale_by_kernel<double>, dim3(nblk(n)), dim3(kBlock), 0, c->stream, a, n,
                       c->red + kRedBlocks + slot);
  tend(c, SPM_K_REALSPACE, 2.0 * (f64 ? 8.0 : tsz(c)) * n);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int e_read_slot(spm_ctx *c, int slot, double *out) {
  SPM_HIP(c, hipMemcpyAsync(out, c->red + kRedBlocks + slot, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  SPM_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

}  // namespace spm
