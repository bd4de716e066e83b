// spectral.hip -- pointwise k-space and real-space kernels of the SPM step.
//
// One thread per complex mode (i, j, kz) for k-space kernels, one thread per
// grid point for real-space kernels.  Wavenumbers are rebuilt from the
// indices instead of read from the reference's KX_int/KY_int/KZ_int/K2/IK2
// tables (fft_wrapper.cxx:29-62), which saves five table reads per mode:
//   kx = i > NX/2 ? i - NX : i   (Calc_KX_Ooura, fft_wrapper.h:687-699)
//   K2 = (WAVE_X kx)^2 + (WAVE_Y ky)^2 + (WAVE_Z kz)^2, IK2 = 1/K2 (0 at k=0)
// Each kernel is HBM-bound: it reads and writes every element it touches
// once; bytes per mode are stated beside each kernel.
#include "spm_internal.h"

namespace spm {

namespace {

constexpr int kBlock = 256;

struct Mode {
  long im;  // index of the real part in the [NX][NY][NZ_] array
  int kx, ky, kz;
  bool keep;  // inside the 2/3-rule box
};

__device__ __forceinline__ bool mode_of(const Grid &g, long t, Mode &m) {
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

template <typename T>
__device__ __forceinline__ T sq(T x) { return x * x; }

// Truncate_two_third_rule_ooura (fft_wrapper.h:723-766).  8 B written per
// discarded mode component.
template <typename T>
__global__ void truncate_kernel(Grid g, T *a) {
  using C2 = cplx<T>;
  Mode m;
  if (!mode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m)) return;
  if (!m.keep) *reinterpret_cast<C2 *>(a + m.im) = C2{0, 0};
}

// Zeta_k2u_k (fft_wrapper.cxx:408-511): 32 B read, 48 B written per mode.
// Mode 0 takes u = uk_dc (real part only); bits of fix_mask zero uk_dc
// first (FIX_CELL, sp_3d_ns.cxx:67-73).
template <typename T>
__global__ void zeta_k2u_k_kernel(Grid g, const T *__restrict__ z0, const T *__restrict__ z1, T *u0,
                                  T *u1, T *u2, double *uk_dc, int fix_mask) {
  using C2 = cplx<T>;
  Mode m;
  if (!mode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m)) return;
  T kx = (T)(g.wave_x * m.kx), ky = (T)(g.wave_y * m.ky), kz = (T)(g.wave_z * m.kz);
  const T k2 = (T)(sq(g.wave_x * (double)m.kx) + sq(g.wave_y * (double)m.ky) + sq(g.wave_z * (double)m.kz));
  const T ik2 = k2 > 0.0 ? (T)1 / k2 : 0.0;
  const C2 d1 = *reinterpret_cast<const C2 *>(z0 + m.im);
  const C2 d2 = *reinterpret_cast<const C2 *>(z1 + m.im);
  T wr0, wi0, wr1, wi1, wr2, wi2;
  if (m.kx != 0) {
    wr0 = -((T)1 / kx) * (ky * d1.x + kz * d2.x);
    wi0 = -((T)1 / kx) * (ky * d1.y + kz * d2.y);
    wr1 = d1.x; wi1 = d1.y;
    wr2 = d2.x; wi2 = d2.y;
  } else if (m.ky != 0) {
    const T dmy = -(kz / ky);
    wr0 = d2.x; wi0 = d2.y;
    wr1 = dmy * d1.x; wi1 = dmy * d1.y;
    wr2 = d1.x; wi2 = d1.y;
  } else {
    wr0 = d1.x; wi0 = d1.y;
    wr1 = d2.x; wi1 = d2.y;
    wr2 = 0.0; wi2 = 0.0;
  }
  kx *= ik2;
  ky *= ik2;
  kz *= ik2;
  C2 o0{(ky * wi2 - kz * wi1), -(ky * wr2 - kz * wr1)};
  C2 o1{(kz * wi0 - kx * wi2), -(kz * wr0 - kx * wr2)};
  C2 o2{(kx * wi1 - ky * wi0), -(kx * wr1 - ky * wr0)};
  if (m.im == 0) {
    for (int d = 0; d < 3; ++d)
      if (fix_mask & (1 << d)) uk_dc[d] = 0.0;
    o0.x = (T)uk_dc[0];
    o1.x = (T)uk_dc[1];
    o2.x = (T)uk_dc[2];
  }
  *reinterpret_cast<C2 *>(u0 + m.im) = o0;
  *reinterpret_cast<C2 *>(u1 + m.im) = o1;
  *reinterpret_cast<C2 *>(u2 + m.im) = o2;
}

// U_k2zeta_k (fft_wrapper.cxx:258-339): 48 B read, 32 B written per mode;
// uk_dc = Re u(k=0).
template <typename T>
__global__ void u_k2zeta_k_kernel(Grid g, const T *__restrict__ u0, const T *__restrict__ u1,
                                  const T *__restrict__ u2, T *z0, T *z1, double *uk_dc) {
  using C2 = cplx<T>;
  Mode m;
  if (!mode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m)) return;
  const T ks0 = (T)(m.kx * g.wave_x), ks1 = (T)(m.ky * g.wave_y), ks2 = (T)(m.kz * g.wave_z);
  const C2 a = *reinterpret_cast<const C2 *>(u0 + m.im);
  const C2 b = *reinterpret_cast<const C2 *>(u1 + m.im);
  const C2 cc = *reinterpret_cast<const C2 *>(u2 + m.im);
  T wr[3], wi[3];
  wr[0] = (ks1 * cc.y - ks2 * b.y);
  wi[0] = -(ks1 * cc.x - ks2 * b.x);
  wr[1] = (ks2 * a.y - ks0 * cc.y);
  wi[1] = -(ks2 * a.x - ks0 * cc.x);
  wr[2] = (ks0 * b.y - ks1 * a.y);
  wi[2] = -(ks0 * b.x - ks1 * a.x);
  int a0, a1;
  if (m.kx != 0) {
    a0 = 1; a1 = 2;
  } else if (m.ky != 0) {
    a0 = 2; a1 = 0;
  } else {
    a0 = 0; a1 = 1;
  }
  *reinterpret_cast<C2 *>(z0 + m.im) = C2{wr[a0], wi[a0]};
  *reinterpret_cast<C2 *>(z1 + m.im) = C2{wr[a1], wi[a1]};
  if (m.im == 0) {
    uk_dc[0] = a.x;
    uk_dc[1] = b.x;
    uk_dc[2] = cc.x;
  }
}

// Real-space products of U2advection_k (operate_omega.cxx:33-49):
// 24 B read, 40 B written per point.
template <typename T>
__global__ void products_kernel(Grid g, T *u0, T *u1, T *u2, T *a0, T *a1) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long npts = (long)g.nx * g.ny * g.nz;
  if (t >= npts) return;
  const long im = (t / g.nz) * g.nz_ + t % g.nz;
  const T v1 = u0[im], v2 = u1[im], v3 = u2[im];
  u0[im] = v2 * v3;
  u1[im] = v3 * v1;
  u2[im] = v1 * v2;
  a0[im] = (v2 - v3) * (v2 + v3);
  a1[im] = (v3 - v1) * (v3 + v1);
}

// k-space contraction of U2advection_k (operate_omega.cxx:60-97) followed by
// Omega_k2zeta_k (fft_wrapper.cxx:137-173), written over a0/a1.  Real and
// imaginary slots use the same wavenumbers, as in the reference.
// 80 B read, 32 B written per mode.
template <typename T>
__global__ void advection_contract_kernel(Grid g, const T *__restrict__ u0, const T *__restrict__ u1,
                                          const T *__restrict__ u2, const T *a0, const T *a1, T *o0, T *o1) {
  using C2 = cplx<T>;
  Mode m;
  if (!mode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m)) return;
  const T k1 = (T)(m.kx * g.wave_x), k2 = (T)(m.ky * g.wave_y), k3 = (T)(m.kz * g.wave_z);
  const T k1k2 = k1 * k2, k2k3 = k2 * k3, k3k1 = k3 * k1;
  const C2 U23 = *reinterpret_cast<const C2 *>(u0 + m.im);
  const C2 U31 = *reinterpret_cast<const C2 *>(u1 + m.im);
  const C2 U12 = *reinterpret_cast<const C2 *>(u2 + m.im);
  const C2 A = *reinterpret_cast<const C2 *>(a0 + m.im);
  const C2 B = *reinterpret_cast<const C2 *>(a1 + m.im);
  T om[3][2];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const T u2u3 = p ? U23.y : U23.x, u3u1 = p ? U31.y : U31.x, u1u2 = p ? U12.y : U12.x;
    const T u22_u32 = p ? A.y : A.x, u32_u12 = p ? B.y : B.x;
    const T u12_u22 = -u32_u12 - u22_u32;
    om[0][p] = -(k2k3 * u22_u32 + (k3 - k2) * (k3 + k2) * u2u3 + k3k1 * u1u2 - k1k2 * u3u1);
    om[1][p] = -(k3k1 * u32_u12 + (k1 - k3) * (k1 + k3) * u3u1 + k1k2 * u2u3 - k2k3 * u1u2);
    om[2][p] = -(k1k2 * u12_u22 + (k2 - k1) * (k2 + k1) * u1u2 + k2k3 * u3u1 - k3k1 * u2u3);
  }
  int s0, s1;
  if (m.kx != 0) {
    s0 = 1; s1 = 2;
  } else if (m.ky != 0) {
    s0 = 2; s1 = 0;
  } else {
    s0 = 0; s1 = 1;
  }
  // o0, o1 may alias a0, a1: each thread reads its mode before writing it
  *reinterpret_cast<C2 *>(o0 + m.im) = C2{om[s0][0], om[s0][1]};
  *reinterpret_cast<C2 *>(o1 + m.im) = C2{om[s1][0], om[s1][1]};
}

// Slaved-Euler update of NS_solver_slavedEuler (fluid_solver.cxx:337-353)
// on the four 2/3-rule blocks; modes outside them are the zeros left by
// the in-place truncation at the start of Zeta_k2advection_k.
// 64 B read, 32 B written per mode.
template <typename T>
__global__ void slaved_euler_kernel(Grid g, T *z0, T *z1, const T *__restrict__ f0,
                                    const T *__restrict__ f1, double nu, double dt) {
  using C2 = cplx<T>;
  Mode m;
  if (!mode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m)) return;
  C2 a{0.0, 0.0}, b{0.0, 0.0};
  if (m.keep) {
    const T k2 = (T)(sq(g.wave_x * (double)m.kx) + sq(g.wave_y * (double)m.ky) + sq(g.wave_z * (double)m.kz));
    const T ik2 = k2 > 0.0 ? (T)1 / k2 : 0.0;
    const T dmy0 = (T)(-nu * dt), dmy1 = (T)(1. / nu);
    const T dmy = exp(dmy0 * k2) - (T)1;
    a = *reinterpret_cast<const C2 *>(z0 + m.im);
    b = *reinterpret_cast<const C2 *>(z1 + m.im);
    const C2 fa = *reinterpret_cast<const C2 *>(f0 + m.im);
    const C2 fb = *reinterpret_cast<const C2 *>(f1 + m.im);
    a.x += dmy * (a.x - dmy1 * ik2 * fa.x);
    a.y += dmy * (a.y - dmy1 * ik2 * fa.y);
    b.x += dmy * (b.x - dmy1 * ik2 * fb.x);
    b.y += dmy * (b.y - dmy1 * ik2 * fb.y);
  }
  *reinterpret_cast<C2 *>(z0 + m.im) = a;
  *reinterpret_cast<C2 *>(z1 + m.im) = b;
}

// 2/3 truncation of three spectra + Solenoidal_uk (operate_omega.cxx:186-219),
// the k-space core of Solenoidal_u (operate_omega.h:96-103).
// 48 B read, 48 B written per mode.
template <typename T>
__global__ void solenoidal_trunc_kernel(Grid g, T *u0, T *u1, T *u2) {
  using C2 = cplx<T>;
  Mode m;
  if (!mode_of(g, (long)blockIdx.x * blockDim.x + threadIdx.x, m)) return;
  C2 a{0.0, 0.0}, b{0.0, 0.0}, cc{0.0, 0.0};
  if (m.keep) {
    const T kx = (T)(m.kx * g.wave_x), ky = (T)(m.ky * g.wave_y), kz = (T)(m.kz * g.wave_z);
    const T k2 = (T)(sq(g.wave_x * (double)m.kx) + sq(g.wave_y * (double)m.ky) + sq(g.wave_z * (double)m.kz));
    const T ik2 = k2 > 0.0 ? (T)1 / k2 : 0.0;
    a = *reinterpret_cast<const C2 *>(u0 + m.im);
    b = *reinterpret_cast<const C2 *>(u1 + m.im);
    cc = *reinterpret_cast<const C2 *>(u2 + m.im);
    const T dr = ik2 * (a.x * kx + b.x * ky + cc.x * kz);
    const T di = ik2 * (a.y * kx + b.y * ky + cc.y * kz);
    a.x -= dr * kx; a.y -= di * kx;
    b.x -= dr * ky; b.y -= di * ky;
    cc.x -= dr * kz; cc.y -= di * kz;
  }
  *reinterpret_cast<C2 *>(u0 + m.im) = a;
  *reinterpret_cast<C2 *>(u1 + m.im) = b;
  *reinterpret_cast<C2 *>(u2 + m.im) = cc;
}

// Make_f_particle_dt_nonsole (f_particle.h:41-64): f = up - phi u.
// 56 B read, 24 B written per point.
template <typename T>
__global__ void make_f_kernel(Grid g, const T *__restrict__ phi, const T *__restrict__ up0,
                              const T *__restrict__ up1, const T *__restrict__ up2,
                              const T *__restrict__ u0, const T *__restrict__ u1,
                              const T *__restrict__ u2, T *f0, T *f1, T *f2) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long npts = (long)g.nx * g.ny * g.nz;
  if (t >= npts) return;
  const long im = (t / g.nz) * g.nz_ + t % g.nz;
  const T p = phi[im];
  f0[im] = up0[im] - p * u0[im];
  f1[im] = up1[im] - p * u1[im];
  f2[im] = up2[im] - p * u2[im];
}

// Add_f_particle (f_particle.h:83-99): u += f.  48 B read, 24 B written.
template <typename T>
__global__ void add_f_kernel(Grid g, T *u0, T *u1, T *u2, const T *__restrict__ f0,
                             const T *__restrict__ f1, const T *__restrict__ f2) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long npts = (long)g.nx * g.ny * g.nz;
  if (t >= npts) return;
  const long im = (t / g.nz) * g.nz_ + t % g.nz;
  u0[im] += f0[im];
  u1[im] += f1[im];
  u2[im] += f2[im];
}

__global__ void zero_kernel(double *a, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = 0.0;
}

template <typename T>
__global__ void finite_kernel(Grid g, const T *__restrict__ a, int spectrum, int *flag) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nrow = (long)g.nx * g.ny;
  const int w = spectrum ? 2 * g.hnz : g.nz;
  if (t >= nrow * w) return;
  const T v = a[(t / w) * g.nz_ + t % w];
  if (!isfinite(v)) atomicOr(flag, 1);
}

inline unsigned nblk(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }
inline long nmodes(const Grid &g) { return (long)g.nx * g.ny * g.hnz; }
inline long npoints(const Grid &g) { return (long)g.nx * g.ny * g.nz; }

template <typename T>
inline T *F(double *p) {
  return reinterpret_cast<T *>(p);
}
template <typename T>
inline const T *F(const double *p) {
  return reinterpret_cast<const T *>(p);
}

// Runs BODY with T = float on an FP32 context, T = double otherwise.
#define SPM_BY_PRECISION(c, ...)                  \
  do {                                            \
    if ((c)->prm.precision == SPM_FP32) {         \
      using T = float;                            \
      __VA_ARGS__;                                \
    } else {                                      \
      using T = double;                           \
      __VA_ARGS__;                                \
    }                                             \
  } while (0)

inline double esz(const spm_ctx *c) { return c->prm.precision == SPM_FP32 ? 4.0 : 8.0; }

}  // namespace

int k_truncate(spm_ctx *c, double *a) {
  tbeg(c);
  SPM_BY_PRECISION(c, hipLaunchKernelGGL(truncate_kernel<T>, dim3(nblk(nmodes(c->grid))), dim3(kBlock), 0, c->stream,
                                         c->grid, F<T>(a)));
  tend(c, SPM_K_SPECTRAL, 2.0 * esz(c) * (double)nmodes(c->grid) *
                              (1.0 - (double)(2 * c->grid.trn_x - 1) * (2 * c->grid.trn_y - 1) * c->grid.trn_z /
                                         (double)nmodes(c->grid)));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int k_zeta_k2u_k(spm_ctx *c, const double *z0, const double *z1, double *u0, double *u1, double *u2, int fix_mask) {
  tbeg(c);
  SPM_BY_PRECISION(c, hipLaunchKernelGGL(zeta_k2u_k_kernel<T>, dim3(nblk(nmodes(c->grid))), dim3(kBlock), 0,
                                         c->stream, c->grid, F<T>(z0), F<T>(z1), F<T>(u0), F<T>(u1), F<T>(u2),
                                         c->uk_dc, fix_mask));
  tend(c, SPM_K_SPECTRAL, 10.0 * esz(c) * (double)nmodes(c->grid));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int k_u_k2zeta_k(spm_ctx *c, const double *u0, const double *u1, const double *u2, double *z0, double *z1) {
  tbeg(c);
  SPM_BY_PRECISION(c, hipLaunchKernelGGL(u_k2zeta_k_kernel<T>, dim3(nblk(nmodes(c->grid))), dim3(kBlock), 0,
                                         c->stream, c->grid, F<T>(u0), F<T>(u1), F<T>(u2), F<T>(z0), F<T>(z1),
                                         c->uk_dc));
  tend(c, SPM_K_SPECTRAL, 10.0 * esz(c) * (double)nmodes(c->grid));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int k_products(spm_ctx *c, double *u0, double *u1, double *u2, double *a0, double *a1) {
  tbeg(c);
  SPM_BY_PRECISION(c, hipLaunchKernelGGL(products_kernel<T>, dim3(nblk(npoints(c->grid))), dim3(kBlock), 0, c->stream,
                                         c->grid, F<T>(u0), F<T>(u1), F<T>(u2), F<T>(a0), F<T>(a1)));
  tend(c, SPM_K_REALSPACE, 8.0 * esz(c) * (double)npoints(c->grid));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int k_advection_contract(spm_ctx *c, double *u0, double *u1, double *u2, const double *a0, const double *a1,
                         double *o0, double *o1) {
  // the zeta-form result goes to o0, o1 (which may be a0, a1)
  tbeg(c);
  SPM_BY_PRECISION(c, hipLaunchKernelGGL(advection_contract_kernel<T>, dim3(nblk(nmodes(c->grid))), dim3(kBlock), 0,
                                         c->stream, c->grid, F<T>(u0), F<T>(u1), F<T>(u2),
                                         reinterpret_cast<const T *>(a0), reinterpret_cast<const T *>(a1), F<T>(o0),
                                         F<T>(o1)));
  tend(c, SPM_K_SPECTRAL, 14.0 * esz(c) * (double)nmodes(c->grid));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int k_slaved_euler(spm_ctx *c, double *z0, double *z1, const double *f0, const double *f1, double nu, double dt) {
  tbeg(c);
  SPM_BY_PRECISION(c, hipLaunchKernelGGL(slaved_euler_kernel<T>, dim3(nblk(nmodes(c->grid))), dim3(kBlock), 0,
                                         c->stream, c->grid, F<T>(z0), F<T>(z1), F<T>(f0), F<T>(f1), nu, dt));
  tend(c, SPM_K_SPECTRAL, 12.0 * esz(c) * (double)nmodes(c->grid));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int k_solenoidal_trunc(spm_ctx *c, double *u0, double *u1, double *u2) {
  tbeg(c);
  SPM_BY_PRECISION(c, hipLaunchKernelGGL(solenoidal_trunc_kernel<T>, dim3(nblk(nmodes(c->grid))), dim3(kBlock), 0,
                                         c->stream, c->grid, F<T>(u0), F<T>(u1), F<T>(u2)));
  tend(c, SPM_K_SPECTRAL, 12.0 * esz(c) * (double)nmodes(c->grid));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int k_make_f(spm_ctx *c) {
  tbeg(c);
  SPM_BY_PRECISION(c, hipLaunchKernelGGL(make_f_kernel<T>, dim3(nblk(npoints(c->grid))), dim3(kBlock), 0, c->stream,
                                         c->grid, F<T>(c->phi), F<T>(c->up[0]), F<T>(c->up[1]), F<T>(c->up[2]),
                                         F<T>(c->u[0]), F<T>(c->u[1]), F<T>(c->u[2]), F<T>(c->f[0]), F<T>(c->f[1]),
                                         F<T>(c->f[2])));
  tend(c, SPM_K_REALSPACE, 10.0 * esz(c) * (double)npoints(c->grid));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int k_add_f(spm_ctx *c) {
  tbeg(c);
  SPM_BY_PRECISION(c, hipLaunchKernelGGL(add_f_kernel<T>, dim3(nblk(npoints(c->grid))), dim3(kBlock), 0, c->stream,
                                         c->grid, F<T>(c->u[0]), F<T>(c->u[1]), F<T>(c->u[2]), F<T>(c->f[0]),
                                         F<T>(c->f[1]), F<T>(c->f[2])));
  tend(c, SPM_K_REALSPACE, 9.0 * esz(c) * (double)npoints(c->grid));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

// Zeroes a whole buffer of n field elements (of the context's precision).
int k_zero(spm_ctx *c, double *a, size_t n) {
  const long nd = c->prm.precision == SPM_FP32 ? (long)((n + 1) / 2) : (long)n;
  tbeg(c);
  hipLaunchKernelGGL(zero_kernel, dim3(nblk(nd)), dim3(kBlock), 0, c->stream, a, nd);
  tend(c, SPM_K_REALSPACE, 8.0 * (double)nd);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int k_zero_f64(spm_ctx *c, double *a, size_t n) {
  tbeg(c);
  hipLaunchKernelGGL(zero_kernel, dim3(nblk((long)n)), dim3(kBlock), 0, c->stream, a, (long)n);
  tend(c, SPM_K_REALSPACE, 8.0 * (double)n);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

__global__ void zero_int_kernel(int *a, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// Zeroes n ints with a kernel (no memset node when captured in a graph).
int k_zero_int(spm_ctx *c, int *a, size_t n) {
  hipLaunchKernelGGL(zero_int_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a, (long)n);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

// Sets *c->flag to 1 if any used element of `a` is NaN or Inf.
int k_check_finite(spm_ctx *c, const double *a, bool spectrum) {
  const long n = (long)c->grid.nx * c->grid.ny * (spectrum ? 2 * c->grid.hnz : c->grid.nz);
  SPM_BY_PRECISION(c, hipLaunchKernelGGL(finite_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                                         c->grid, F<T>(a), spectrum ? 1 : 0, c->flag));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

}  // namespace spm
