/*
 * spm_oracle.c -- CPU FP64 restatement of KAPSEL's SPM time step.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline); see
 * spm_oracle.h for what may use it and how each function is pinned.
 * Single-threaded, like the reference's shipped GCC build (Makefile:43-48).
 *
 * Deviations from the reference, all without effect on the arithmetic of a
 * run with fixed parameters:
 *   - state is held in a context instead of globals and function-local
 *     statics (input.cxx:13-271; fft_wrapper.h:552; md_force.cxx:264;
 *     fluct.cxx:11-18), so several configurations can live in one process;
 *   - the 3-D FFT is this file's own radix-2 code (not Ooura's); it
 *     computes the same transform (pinned against oracle/_ref);
 *   - the padding slots k = NZ, NZ+1 of a real-space field are set to 0
 *     where the reference leaves leftovers (fft_wrapper.h:534-631).
 */
#include "spm_oracle.h"

#include <float.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

#define DIM 3
#define MT_N 624
#define MT_M 397

typedef struct {
  int n;
  double *c, *s;  /* cos, sin of 2*pi*t/n, t < n */
  int *rev;       /* bit reversal of t < n */
} fft_plan;

struct oracle_ctx {
  spm_params P;
  int NX, NY, NZ, NZ_, HNZ_, TRN_X, TRN_Y, TRN_Z;
  size_t nfield;                /* NX*NY*NZ_ */
  double DX, RHO, ETA, NU, IRHO, DX3, kBT, RADIUS, XI, HXI, SIGMA;
  double L[DIM], WAVE_X, WAVE_Y, WAVE_Z;
  double MASS[SPM_MAX_SPECIES], IMASS[SPM_MAX_SPECIES];
  double MOI[SPM_MAX_SPECIES], IMOI[SPM_MAX_SPECIES];
  double LJ_dia, A_R_cutoff;
  int *KX_int, *KY_int, *KZ_int;
  double *K2, *IK2;
  double *zeta[2], *u[3], *up[3], *phi, *f[3], *f_ns0[2], *adv_tmp[2];
  double uk_dc[3];
  int np_domain;
  int *sekibun;                 /* np_domain x 3 */
  spm_particle *p;
  int Np;
  int ts;
  double time, dt, hdt;
  uint32_t mt[MT_N];
  int mti;
  fft_plan px, py, pz;          /* pz has n = NZ/2 */
  double *ztw;                  /* cos, sin of 2*pi*k/NZ, k <= NZ/2 */
  /* electrolyte (SW_EQ == Electrolyte) */
  int electro;
  spm_electrolyte E;
  double *conc[2], *conc_rhs[2], *surface_normal[3], *f_ns1[2];
  double total_solute[2];
  double surface_charge_e[SPM_MAX_SPECIES];
  double valency_e[2];
  /* Lees-Edwards shear (SW_EQ == Shear_Navier_Stokes_Lees_Edwards) */
  int le;
  double shear_rate, shear_rate_eff, degree_oblique;
  /* Mean_shear_stress inputs: the sums over the grid of Hydro_force and
   * Hydro_force_new, dev_shear_stress_lj before the Ivolume factor, and
   * Shear_strain_realized */
  double hydro_force_sum, hydro_force_new_sum, lj_shear_stress, strain;
  double *ucp[3], *wv3[3], *wv1;
};

/* ------------------------------------------------------------------ */
/* small helpers (macro.h:38-45, 129-150)                              */
/* ------------------------------------------------------------------ */
static inline double SQ(double x) { return x * x; }
static inline double POW3(double x) { return x * x * x; }
static inline int Nint(double a) { return a >= 0.0 ? (int)(a + .5) : (int)(a - .5); }
static inline double DMIN(double a, double b) { return a < b ? a : b; }

static size_t IDX(const oracle_ctx *c, int i, int j, int k) {
  return ((size_t)i * c->NY + j) * c->NZ_ + k;
}

/* ------------------------------------------------------------------ */
/* FFT: own radix-2 code computing Ooura's transform                  */
/* X(k) = sum a(x) exp(+2 pi i k.x/N), half spectrum in [NX][NY][NZ_]  */
/* (fftsg3d.c:109-120, fft_wrapper.h:432-631)                          */
/* ------------------------------------------------------------------ */
/* Threads of the parallel loops (the line transforms and the per-particle
 * force sums); 1 runs the restatement serially.  No effect without OpenMP. */
void oracle_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

static int plan_init(fft_plan *pl, int n) {
  pl->n = n;
  pl->c = (double *)malloc(sizeof(double) * n);
  pl->s = (double *)malloc(sizeof(double) * n);
  pl->rev = (int *)malloc(sizeof(int) * n);
  if (!pl->c || !pl->s || !pl->rev) return -1;
  int lg = 0;
  while ((1 << lg) < n) lg++;
  for (int t = 0; t < n; t++) {
    long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
    pl->c[t] = (double)cosl(ang);
    pl->s[t] = (double)sinl(ang);
    int r = 0;
    for (int b = 0; b < lg; b++)
      if (t & (1 << b)) r |= 1 << (lg - 1 - b);
    pl->rev[t] = r;
  }
  return 0;
}

static void plan_free(fft_plan *pl) {
  free(pl->c);
  free(pl->s);
  free(pl->rev);
}

/* In-place complex FFT of n points on `batch` interleaved columns:
 * element j of column b is the complex number at a[2*(j*stride + b)].
 * sign = +1 computes sum x_j exp(+2 pi i jk/n), -1 the conjugate kernel. */
static void fft_batch(const fft_plan *pl, double *a, long stride, int batch, int sign) {
  const int n = pl->n;
  if (n & (n - 1)) {
    /* lengths with factors 3 and 5 (the GPU path's 48, 80, 96, 160, 192,
     * which the reference cannot run): the DFT by its definition */
    double t[2 * 192];
    for (int b = 0; b < batch; b++) {
      for (int k = 0; k < n; k++) {
        double sr = 0.0, si = 0.0;
        for (int j = 0; j < n; j++) {
          const int m = (int)(((long)j * k) % n);
          const double xr = a[2 * ((long)j * stride + b)], xi = a[2 * ((long)j * stride + b) + 1];
          const double wr = pl->c[m], wi = sign * pl->s[m];
          sr += xr * wr - xi * wi;
          si += xr * wi + xi * wr;
        }
        t[2 * k] = sr;
        t[2 * k + 1] = si;
      }
      for (int k = 0; k < n; k++) {
        a[2 * ((long)k * stride + b)] = t[2 * k];
        a[2 * ((long)k * stride + b) + 1] = t[2 * k + 1];
      }
    }
    return;
  }
  for (int j = 0; j < n; j++) {
    int r = pl->rev[j];
    if (r > j) {
      double *x = a + 2 * (long)j * stride, *y = a + 2 * (long)r * stride;
      for (int b = 0; b < 2 * batch; b++) {
        double t = x[b];
        x[b] = y[b];
        y[b] = t;
      }
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    const int half = len >> 1, step = n / len;
    for (int k = 0; k < half; k++) {
      const double wr = pl->c[k * step], wi = sign * pl->s[k * step];
      for (int start = 0; start < n; start += len) {
        double *x = a + 2 * (long)(start + k) * stride;
        double *y = a + 2 * (long)(start + k + half) * stride;
        for (int b = 0; b < batch; b++) {
          double yr = y[2 * b] * wr - y[2 * b + 1] * wi;
          double yi = y[2 * b] * wi + y[2 * b + 1] * wr;
          double xr = x[2 * b], xi = x[2 * b + 1];
          x[2 * b] = xr + yr;
          x[2 * b + 1] = xi + yi;
          y[2 * b] = xr - yr;
          y[2 * b + 1] = xi - yi;
        }
      }
    }
  }
}

/* A2a_k, fft_wrapper.h:432-527: real field (k < NZ) -> half spectrum. */
void oracle_A2a_k(oracle_ctx *c, double *a) {
  const int NX = c->NX, NY = c->NY, NZ = c->NZ, M = NZ / 2, H = c->HNZ_;
  /* every line is transformed on its own, so the lines are shared out over
   * OpenMP threads (oracle_set_threads); the result does not depend on how */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < NX; i++) {
    double zbuf[2 * 1024 + 2];
    double *z = zbuf;
    for (int j = 0; j < NY; j++) {
      double *row = a + IDX(c, i, j, 0);
      memcpy(z, row, sizeof(double) * NZ);
      fft_batch(&c->pz, z, 1, 1, +1);
      /* X[k] = E[k] + w^k O[k], w = exp(+2 pi i/NZ) */
      for (int k = 0; k <= M; k++) {
        int k1 = k % M, k2 = (M - k) % M;
        double zr = z[2 * k1], zi = z[2 * k1 + 1];
        double cr = z[2 * k2], ci = -z[2 * k2 + 1];
        double er = 0.5 * (zr + cr), ei = 0.5 * (zi + ci);
        /* O = (Z - conj Z')/(2i) */
        double dr = zr - cr, di = zi - ci;
        double orr = 0.5 * di, oi = -0.5 * dr;
        double wr = c->ztw[2 * k], wi = c->ztw[2 * k + 1];
        row[2 * k] = er + (wr * orr - wi * oi);
        row[2 * k + 1] = ei + (wr * oi + wi * orr);
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < NX; i++) fft_batch(&c->py, a + IDX(c, i, 0, 0), H, H, +1);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < NY; j++) fft_batch(&c->px, a + IDX(c, 0, j, 0), (long)NY * H, H, +1);
}

/* A_k2a, fft_wrapper.h:534-631: half spectrum -> real field, scale 1/N. */
/* The kz = 0 and kz = NZ/2 planes of a half spectrum are Hermitian in
 * (kx, ky) for a real field.  Ooura's inverse reads only one member of each
 * pair (k, -k) of those planes (rdft3dsort(-1), fftsg3d.c:645-668) and takes
 * the other as its conjugate; the four self-conjugate points keep only their
 * real part.  Kept on kz = 0: 1 <= j < NY/2, or j in {0, NY/2} with
 * i <= NX/2; on kz = NZ/2: j > NY/2, or j in {0, NY/2} with i > NX/2.  This
 * makes a spectrum that is not Hermitian (a derivative's Nyquist modes)
 * transform as the reference transforms it. */
static void ooura_hermitian_planes(const oracle_ctx *c, double *a) {
  const int NX = c->NX, NY = c->NY, n1h = NX / 2, n2h = NY / 2, M = c->NZ / 2;
  for (int plane = 0; plane < 2; plane++) {
    const int kz = plane == 0 ? 0 : M;
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++) {
        const int edge = (j == 0 || j == n2h);
        const int selfc = edge && (i == 0 || i == n1h);
        int keep;
        if (plane == 0)
          keep = (j >= 1 && j < n2h) || (edge && i <= n1h);
        else
          keep = (j > n2h) || (edge && i > n1h) || selfc;
        double *x = a + IDX(c, i, j, 2 * kz);
        if (selfc) {
          x[1] = 0.0;
        } else if (!keep) {
          const double *m = a + IDX(c, (NX - i) % NX, (NY - j) % NY, 2 * kz);
          x[0] = m[0];
          x[1] = -m[1];
        }
      }
  }
}

void oracle_A_k2a(oracle_ctx *c, double *a) {
  const int NX = c->NX, NY = c->NY, NZ = c->NZ, M = NZ / 2, H = c->HNZ_;
  ooura_hermitian_planes(c, a);
  const double scale = 1.0 / ((double)NX * NY * NZ);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < NY; j++) fft_batch(&c->px, a + IDX(c, 0, j, 0), (long)NY * H, H, -1);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < NX; i++) fft_batch(&c->py, a + IDX(c, i, 0, 0), H, H, -1);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < NX; i++) {
    double zbuf[2 * 1024 + 2];
    double *z = zbuf;
    for (int j = 0; j < NY; j++) {
      double *row = a + IDX(c, i, j, 0);
      for (int k = 0; k < M; k++) {
        double xr = row[2 * k], xi = row[2 * k + 1];
        double cr = row[2 * (M - k)], ci = -row[2 * (M - k) + 1];
        double er = xr + cr, ei = xi + ci;
        double dr = xr - cr, di = xi - ci;
        double wr = c->ztw[2 * k], wi = -c->ztw[2 * k + 1];
        double orr = dr * wr - di * wi, oi = dr * wi + di * wr;
        z[2 * k] = er - oi;      /* Z = E + i O */
        z[2 * k + 1] = ei + orr;
      }
      fft_batch(&c->pz, z, 1, 1, -1);
      for (int k = 0; k < NZ; k++) row[k] = z[k] * scale;
      row[NZ] = 0.0;
      row[NZ + 1] = 0.0;
    }
  }
}

/* ------------------------------------------------------------------ */
/* k tables, Init_K fft_wrapper.cxx:29-62, Calc_K*_Ooura h:687-699      */
/* ------------------------------------------------------------------ */
static void init_k(oracle_ctx *c) {
  const int NX = c->NX, NY = c->NY, NZ_ = c->NZ_;
  for (int i = 0; i < NX; i++)
    for (int j = 0; j < NY; j++)
      for (int k = 0; k < NZ_; k++) {
        size_t im = IDX(c, i, j, k);
        int kx = (i > NX / 2) ? i - NX : i;
        int ky = (j > NY / 2) ? j - NY : j;
        int kz = k / 2;
        c->KX_int[im] = kx;
        c->KY_int[im] = ky;
        c->KZ_int[im] = kz;
        c->K2[im] = SQ(c->WAVE_X * (double)kx) + SQ(c->WAVE_Y * (double)ky) + SQ(c->WAVE_Z * (double)kz);
        c->IK2[im] = c->K2[im] > 0.0 ? 1. / c->K2[im] : 0.0;
      }
}

/* Truncate_two_third_rule_ooura, fft_wrapper.h:723-766: zero every mode
 * outside |kx| < TRN_X, |ky| < TRN_Y, kz < TRN_Z (4 regions). */
void oracle_truncate(oracle_ctx *c, double *a) {
  const int NX = c->NX, NY = c->NY, NZ_ = c->NZ_;
  const int TX = c->TRN_X, TY = c->TRN_Y, tz2 = 2 * c->TRN_Z;
  for (int i = 0; i < NX; i++)
    for (int j = 0; j < NY; j++)
      for (int k = tz2; k < NZ_; k++) a[IDX(c, i, j, k)] = 0.0;
  for (int i = 0; i < NX; i++)
    for (int j = TY; j <= NY - TY; j++)
      for (int k = 0; k < tz2; k++) a[IDX(c, i, j, k)] = 0.0;
  for (int i = TX; i <= NX - TX; i++)
    for (int j = 0; j <= TY - 1; j++)
      for (int k = 0; k < tz2; k++) a[IDX(c, i, j, k)] = 0.0;
  for (int i = TX; i <= NX - TX; i++)
    for (int j = NY - TY + 1; j <= NY - 1; j++)
      for (int k = 0; k < tz2; k++) a[IDX(c, i, j, k)] = 0.0;
}

/* ------------------------------------------------------------------ */
/* spectral operators                                                  */
/* ------------------------------------------------------------------ */
/* Zeta_k2u_k, fft_wrapper.cxx:408-511. */
void oracle_zeta_k2u_k(oracle_ctx *c, double *const zeta[2], const double uk_dc[3], double *const u[3]) {
  const int NX = c->NX, NY = c->NY, H = c->HNZ_;
  for (int i = 0; i < NX; i++)
    for (int j = 0; j < NY; j++)
      for (int k = 0; k < H; k++) {
        size_t im = IDX(c, i, j, 2 * k);
        double kx = c->WAVE_X * c->KX_int[im];
        double ky = c->WAVE_Y * c->KY_int[im];
        double kz = c->WAVE_Z * c->KZ_int[im];
        double ik2 = c->IK2[im];
        double d1r = zeta[0][im], d1i = zeta[0][im + 1];
        double d2r = zeta[1][im], d2i = zeta[1][im + 1];
        double wr[3], wi[3];
        if (c->KX_int[im] != 0) {
          wr[0] = -(1. / kx) * (ky * d1r + kz * d2r);
          wi[0] = -(1. / kx) * (ky * d1i + kz * d2i);
          wr[1] = d1r; wi[1] = d1i;
          wr[2] = d2r; wi[2] = d2i;
        } else if (c->KY_int[im] != 0) {
          double dmy = -(kz / ky);
          wr[0] = d2r; wi[0] = d2i;
          wr[1] = dmy * d1r; wi[1] = dmy * d1i;
          wr[2] = d1r; wi[2] = d1i;
        } else {
          wr[0] = d1r; wi[0] = d1i;
          wr[1] = d2r; wi[1] = d2i;
          wr[2] = 0.0; wi[2] = 0.0;
        }
        kx *= ik2;
        ky *= ik2;
        kz *= ik2;
        u[0][im] = (ky * wi[2] - kz * wi[1]);
        u[0][im + 1] = -(ky * wr[2] - kz * wr[1]);
        u[1][im] = (kz * wi[0] - kx * wi[2]);
        u[1][im + 1] = -(kz * wr[0] - kx * wr[2]);
        u[2][im] = (kx * wi[1] - ky * wi[0]);
        u[2][im + 1] = -(kx * wr[1] - ky * wr[0]);
      }
  u[0][0] = uk_dc[0];
  u[1][0] = uk_dc[1];
  u[2][0] = uk_dc[2];
}

/* U_k2zeta_k, fft_wrapper.cxx:258-339. */
void oracle_u_k2zeta_k(oracle_ctx *c, double *const u[3], double *const zeta[2], double uk_dc[3]) {
  const int NX = c->NX, NY = c->NY, H = c->HNZ_;
  uk_dc[0] = u[0][0];
  uk_dc[1] = u[1][0];
  uk_dc[2] = u[2][0];
  for (int i = 0; i < NX; i++)
    for (int j = 0; j < NY; j++)
      for (int k = 0; k < H; k++) {
        size_t im = IDX(c, i, j, 2 * k);
        double ks0 = c->KX_int[im] * c->WAVE_X;
        double ks1 = c->KY_int[im] * c->WAVE_Y;
        double ks2 = c->KZ_int[im] * c->WAVE_Z;
        double wr[3], wi[3];
        wr[0] = (ks1 * u[2][im + 1] - ks2 * u[1][im + 1]);
        wi[0] = -(ks1 * u[2][im] - ks2 * u[1][im]);
        wr[1] = (ks2 * u[0][im + 1] - ks0 * u[2][im + 1]);
        wi[1] = -(ks2 * u[0][im] - ks0 * u[2][im]);
        wr[2] = (ks0 * u[1][im + 1] - ks1 * u[0][im + 1]);
        wi[2] = -(ks0 * u[1][im] - ks1 * u[0][im]);
        int a0, a1;
        if (c->KX_int[im] != 0) {
          a0 = 1; a1 = 2;
        } else if (c->KY_int[im] != 0) {
          a0 = 2; a1 = 0;
        } else {
          a0 = 0; a1 = 1;
        }
        zeta[0][im] = wr[a0];
        zeta[0][im + 1] = wi[a0];
        zeta[1][im] = wr[a1];
        zeta[1][im + 1] = wi[a1];
      }
}

/* Omega_k2zeta_k, fft_wrapper.cxx:137-173. */
static void omega_k2zeta_k(oracle_ctx *c, double *const omega[3], double *const zetak[2]) {
  for (size_t im = 0; im < c->nfield; im++) {
    if (c->KX_int[im] != 0) {
      zetak[0][im] = omega[1][im];
      zetak[1][im] = omega[2][im];
    } else if (c->KY_int[im] != 0) {
      zetak[0][im] = omega[2][im];
      zetak[1][im] = omega[0][im];
    } else {
      zetak[0][im] = omega[0][im];
      zetak[1][im] = omega[1][im];
    }
  }
}

/* U2advection_k, operate_omega.cxx:14-100 (u is overwritten). */
void oracle_u2advection_k(oracle_ctx *c, double *const u[3], double *const adv[2]) {
  const int NX = c->NX, NY = c->NY, NZ = c->NZ;
  for (int i = 0; i < NX; i++)
    for (int j = 0; j < NY; j++)
      for (int k = 0; k < NZ; k++) {
        size_t im = IDX(c, i, j, k);
        double u1 = u[0][im], u2 = u[1][im], u3 = u[2][im];
        u[0][im] = u2 * u3;
        u[1][im] = u3 * u1;
        u[2][im] = u1 * u2;
        adv[0][im] = (u2 - u3) * (u2 + u3);
        adv[1][im] = (u3 - u1) * (u3 + u1);
      }
  for (int d = 0; d < DIM; d++) oracle_A2a_k(c, u[d]);
  oracle_A2a_k(c, adv[0]);
  oracle_A2a_k(c, adv[1]);
  for (size_t im = 0; im < c->nfield; im++) {
    double k1 = c->KX_int[im] * c->WAVE_X;
    double k2 = c->KY_int[im] * c->WAVE_Y;
    double k3 = c->KZ_int[im] * c->WAVE_Z;
    double u2u3 = u[0][im], u3u1 = u[1][im], u1u2 = u[2][im];
    double u22_u32 = adv[0][im], u32_u12 = adv[1][im];
    double u12_u22 = -u32_u12 - u22_u32;
    double k1k2 = k1 * k2, k2k3 = k2 * k3, k3k1 = k3 * k1;
    u[0][im] = -(k2k3 * u22_u32 + (k3 - k2) * (k3 + k2) * u2u3 + k3k1 * u1u2 - k1k2 * u3u1);
    u[1][im] = -(k3k1 * u32_u12 + (k1 - k3) * (k1 + k3) * u3u1 + k1k2 * u2u3 - k2k3 * u1u2);
    u[2][im] = -(k1k2 * u12_u22 + (k2 - k1) * (k2 + k1) * u1u2 + k2k3 * u3u1 - k3k1 * u2u3);
  }
  omega_k2zeta_k(c, u, adv);
}

/* Solenoidal_uk, operate_omega.cxx:186-219. */
void oracle_solenoidal_uk(oracle_ctx *c, double *const u[3]) {
  for (size_t im = 0; im < c->nfield; im++) {
    double kx = c->KX_int[im] * c->WAVE_X;
    double ky = c->KY_int[im] * c->WAVE_Y;
    double kz = c->KZ_int[im] * c->WAVE_Z;
    double dmy = c->IK2[im] * (u[0][im] * kx + u[1][im] * ky + u[2][im] * kz);
    u[0][im] -= dmy * kx;
    u[1][im] -= dmy * ky;
    u[2][im] -= dmy * kz;
  }
}

static void u2u_k(oracle_ctx *c, double *const u[3]) {
  for (int d = 0; d < DIM; d++) oracle_A2a_k(c, u[d]);
}
static void u_k2u(oracle_ctx *c, double *const u[3]) {
  for (int d = 0; d < DIM; d++) oracle_A_k2a(c, u[d]);
}
/* Zeta_k2u, operate_omega.h:125-128. */
static void zeta_k2u(oracle_ctx *c) {
  oracle_zeta_k2u_k(c, c->zeta, c->uk_dc, c->u);
  u_k2u(c, c->u);
}
/* Solenoidal_u, operate_omega.h:96-103. */
static void solenoidal_u(oracle_ctx *c, double *const u[3]) {
  u2u_k(c, u);
  for (int d = 0; d < DIM; d++) oracle_truncate(c, u[d]);
  oracle_solenoidal_uk(c, u);
  u_k2u(c, u);
}

/* NS_solver_slavedEuler, fluid_solver.cxx:328-354, with Zeta_k2advection_k
 * (operate_omega.cxx:102-113) and the ijk_range blocks of
 * sp_3d_ns.cxx:15-20. */
int oracle_ns_solver_slaved_euler(oracle_ctx *c) {
  for (int d = 0; d < DIM - 1; d++) oracle_truncate(c, c->zeta[d]);
  oracle_zeta_k2u_k(c, c->zeta, c->uk_dc, c->u);
  u_k2u(c, c->u);
  oracle_u2advection_k(c, c->u, c->f_ns0);
  const double dmy0 = -c->NU * c->dt;
  const double dmy1 = 1. / c->NU;
  const int NX = c->NX, NY = c->NY, TX = c->TRN_X, TY = c->TRN_Y, TZ = c->TRN_Z;
  const int rng[4][6] = {{0, TX - 1, 0, TY - 1, 0, 2 * TZ - 1},
                         {0, TX - 1, NY - TY + 1, NY - 1, 0, 2 * TZ - 1},
                         {NX - TX + 1, NX - 1, 0, TY - 1, 0, 2 * TZ - 1},
                         {NX - TX + 1, NX - 1, NY - TY + 1, NY - 1, 0, 2 * TZ - 1}};
  for (int n = 0; n < 4; n++)
    for (int i = rng[n][0]; i <= rng[n][1]; i++)
      for (int j = rng[n][2]; j <= rng[n][3]; j++)
        for (int k = rng[n][4]; k <= rng[n][5]; k++) {
          size_t im = IDX(c, i, j, k);
          double dmy = exp(dmy0 * c->K2[im]) - 1.;
          c->zeta[0][im] += dmy * (c->zeta[0][im] - dmy1 * c->IK2[im] * c->f_ns0[0][im]);
          c->zeta[1][im] += dmy * (c->zeta[1][im] - dmy1 * c->IK2[im] * c->f_ns0[1][im]);
        }
  return 0;
}

/* ------------------------------------------------------------------ */
/* particles on the grid                                               */
/* ------------------------------------------------------------------ */
/* H and Phi, profile.h:29-49. */
static inline double Hfun(const oracle_ctx *c, double x) { return x > 0 ? exp(-SQ(c->DX / x)) : 0; }
static inline double Phi(const oracle_ctx *c, double x, double radius) {
  double dmy = Hfun(c, radius + c->HXI - x);
  return dmy / (dmy + Hfun(c, x - radius + c->HXI));
}
double oracle_phi(const oracle_ctx *c, double x) { return Phi(c, x, c->RADIUS); }

/* Particle_domain, profile.cxx:25-74. */
static int particle_domain(oracle_ctx *c) {
  const int nh = (int)c->P.a + (int)c->P.a_xi + 2;
  const double MAX_RADIUS = c->RADIUS + c->HXI + 1.7321 * c->DX;
  int cap = (2 * nh + 1) * (2 * nh + 1) * (2 * nh + 1);
  c->sekibun = (int *)malloc(sizeof(int) * 3 * cap);
  if (!c->sekibun) return -1;
  int mesh = 0;
  for (int i = -nh; i <= nh; i++)
    for (int j = -nh; j <= nh; j++)
      for (int k = -nh; k <= nh; k++) {
        double dmy = sqrt(SQ((double)i * c->DX) + SQ((double)j * c->DX) + SQ((double)k * c->DX));
        if (dmy < MAX_RADIUS) {
          c->sekibun[3 * mesh] = i;
          c->sekibun[3 * mesh + 1] = j;
          c->sekibun[3 * mesh + 2] = k;
          mesh++;
        }
      }
  c->np_domain = mesh;
  return 0;
}

/* Particle_cell, make_phi.h:102-130. */
static void particle_cell(const oracle_ctx *c, const double *xp, int *x_int, double *residue) {
  const double idx = 1. / c->DX;
  for (int d = 0; d < DIM; d++) {
    double dmy = xp[d] * idx;
    x_int[d] = (int)dmy;
    residue[d] = (dmy - (double)x_int[d]) * c->DX;
  }
}

/* Relative_coord, make_phi.h:143-160. */
static void relative_coord(const oracle_ctx *c, const int *cell, const int *x_int, const double *residue,
                           int *r_mesh, double *r) {
  const int Ns[3] = {c->NX, c->NY, c->NZ};
  for (int d = 0; d < DIM; d++) {
    r_mesh[d] = (x_int[d] + cell[d] + Ns[d]) % Ns[d];
    r[d] = (double)cell[d] * c->DX - residue[d];
  }
}

/* Distance0 / Distance, interaction.h:110-135. */
static double distance0(const oracle_ctx *c, const double *x1, const double *x2, double *x12) {
  double dmy = 0.0;
  for (int d = 0; d < DIM; d++) {
    x12[d] = x2[d] - x1[d];
    x12[d] -= (double)Nint(x12[d] / c->L[d]) * c->L[d];
    dmy += SQ(x12[d]);
  }
  return sqrt(dmy);
}

/* Angular2v_rot_on / _rot_off, make_phi.h:355-374. */
static void angular2v(const oracle_ctx *c, const double *w, const double *r, double *v) {
  if (c->P.rotation) {
    v[0] = w[1] * r[2] - w[2] * r[1];
    v[1] = w[2] * r[0] - w[0] * r[2];
    v[2] = w[0] * r[1] - w[1] * r[0];
  } else {
    v[0] = v[1] = v[2] = 0.0;
  }
}

/* Reset_phi_u, make_phi.h:426-468. */
static void reset_phi_u(oracle_ctx *c) {
  memset(c->phi, 0, sizeof(double) * c->nfield);
  for (int d = 0; d < DIM; d++) memset(c->up[d], 0, sizeof(double) * c->nfield);
}

/* Make_phi_u_primitive, make_phi.cxx:221-298. */
static void make_phi_u_primitive(oracle_ctx *c, int sw_up) {
  for (int n = 0; n < c->Np; n++) {
    const spm_particle *pn = &c->p[n];
    int x_int[3], r_mesh[3];
    double residue[3], r[3], x[3], x12[3], v_rot[3];
    particle_cell(c, pn->x, x_int, residue);
    for (int mesh = 0; mesh < c->np_domain; mesh++) {
      relative_coord(c, &c->sekibun[3 * mesh], x_int, residue, r_mesh, r);
      for (int d = 0; d < DIM; d++) x[d] = r_mesh[d] * c->DX;
      double dmy = distance0(c, x, pn->x, x12);
      double dmy_phi = Phi(c, dmy, c->RADIUS);
      size_t im = IDX(c, r_mesh[0], r_mesh[1], r_mesh[2]);
      c->phi[im] += dmy_phi;
      if (sw_up) {
        angular2v(c, pn->omega, r, v_rot);
        for (int d = 0; d < DIM; d++) c->up[d][im] += ((pn->v[d] + v_rot[d]) * dmy_phi);
      }
    }
  }
  if (sw_up) {
    for (int i = 0; i < c->NX; i++)
      for (int j = 0; j < c->NY; j++)
        for (int k = 0; k < c->NZ; k++) {
          size_t im = IDX(c, i, j, k);
          if (c->phi[im] > 1.) {
            double idmy_phi = 1. / c->phi[im];
            c->up[0][im] *= idmy_phi;
            c->up[1][im] *= idmy_phi;
            c->up[2][im] *= idmy_phi;
            c->phi[im] = 1.;
          }
        }
  }
}

int oracle_make_phi_u_particle(oracle_ctx *c) {
  reset_phi_u(c);
  make_phi_u_primitive(c, 1);
  return 0;
}
int oracle_make_phi_particle(oracle_ctx *c) {
  memset(c->phi, 0, sizeof(double) * c->nfield);
  make_phi_u_primitive(c, 0);
  return 0;
}

/* Make_f_particle_dt_sole + Add_f_particle, f_particle.h:41-99. */
int oracle_make_f_particle_dt_sole(oracle_ctx *c) {
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++)
      for (int k = 0; k < c->NZ; k++) {
        size_t im = IDX(c, i, j, k);
        for (int d = 0; d < DIM; d++) c->f[d][im] = c->up[d][im] - c->phi[im] * c->u[d][im];
      }
  solenoidal_u(c, c->f);
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++)
      for (int k = 0; k < c->NZ; k++) {
        size_t im = IDX(c, i, j, k);
        for (int d = 0; d < DIM; d++) c->u[d][im] += c->f[d][im];
      }
  return 0;
}

/* Calc_f_hydro_correct_precision, md_force.cxx:263-389 (DBG_MASS_GRID off,
 * no rigid bodies), preceded by Reset_phi_u as at sp_3d_ns.cxx:127. */
static void calc_f_hydro_field(oracle_ctx *c, double *const fld[3]);
int oracle_calc_f_hydro(oracle_ctx *c) {
  calc_f_hydro_field(c, c->u);
  return 0;
}
static void calc_f_hydro_field(oracle_ctx *c, double *const fld[3]) {
  reset_phi_u(c);
  const double dmy0 = -c->DX3 * c->RHO;
  const double dmy = dmy0 / c->dt;
#pragma omp parallel for schedule(dynamic, 16)
  for (int n = 0; n < c->Np; n++) {
    spm_particle *pn = &c->p[n];
    int x_int[3], r_mesh[3];
    double residue[3], r[3], x[3], x12[3], v_rot[3], dmy_fp[3];
    double force[3] = {0, 0, 0}, torque[3] = {0, 0, 0};
    particle_cell(c, pn->x, x_int, residue);
    for (int mesh = 0; mesh < c->np_domain; mesh++) {
      relative_coord(c, &c->sekibun[3 * mesh], x_int, residue, r_mesh, r);
      for (int d = 0; d < DIM; d++) x[d] = r_mesh[d] * c->DX;
      double dmyR = distance0(c, x, pn->x, x12);
      double dmy_phi = Phi(c, dmyR, c->RADIUS);
      angular2v(c, pn->omega, r, v_rot);
      size_t im = IDX(c, r_mesh[0], r_mesh[1], r_mesh[2]);
      for (int d = 0; d < DIM; d++) {
        dmy_fp[d] = ((pn->v[d] + v_rot[d]) - fld[d][im]) * dmy_phi;
        force[d] += dmy_fp[d];
      }
      torque[0] += (r[1] * dmy_fp[2] - r[2] * dmy_fp[1]);
      torque[1] += (r[2] * dmy_fp[0] - r[0] * dmy_fp[2]);
      torque[2] += (r[0] * dmy_fp[1] - r[1] * dmy_fp[0]);
    }
    for (int d = 0; d < DIM; d++) {
      pn->f_hydro[d] = (dmy * force[d]) * 1.0;
      pn->f_slip[d] = 0.0;
    }
    if (c->P.rotation)
      for (int d = 0; d < DIM; d++) {
        pn->torque_hydro[d] = (dmy * torque[d]) * 1.0;
        pn->torque_slip[d] = 0.0;
      }
  }
}

/* ------------------------------------------------------------------ */
/* forces                                                              */
/* ------------------------------------------------------------------ */
/* Lennard_Jones_f, interaction.h:156-181. */
static double lennard_jones_f(const oracle_ctx *c, double x, double sigma) {
  double answer = 0.0;
  double dmy = sigma / x;
  if (c->P.lj_powers == 0) {
    dmy = SQ(dmy) * SQ(dmy) * SQ(dmy);
    answer = (24. * c->P.epsilon) / SQ(x) * (2.0 * SQ(dmy) - dmy);
  } else if (c->P.lj_powers == 1) {
    dmy = SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy);
    answer = (48. * c->P.epsilon) / SQ(x) * (2.0 * SQ(dmy) - dmy);
  } else if (c->P.lj_powers == 2) {
    dmy = SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy);
    answer = (72. * c->P.epsilon) / SQ(x) * (2.0 * SQ(dmy) - dmy);
  }
  return answer;
}

static void lj_pair(oracle_ctx *c, const double *xi, const double *xj, double *fi, double *fj) {
  double r_ij_vec[3];
  double r_ij = distance0(c, xi, xj, r_ij_vec);
  if (r_ij < c->A_R_cutoff * c->LJ_dia) {
    double dmy = DMIN(DBL_MAX / r_ij, lennard_jones_f(c, r_ij, c->LJ_dia));
    for (int d = 0; d < DIM; d++) {
      double dmy1 = dmy * -r_ij_vec[d];
      fi[d] += dmy1;
      fj[d] -= dmy1;
    }
  }
}

/* Calc_f_Lennard_Jones_shear_cap_primitive_lnk, md_force.cxx:23-134:
 * cells of Cell_length = 16 grid spacings, 27 neighbours with wrap, pairs
 * taken with i > j.  Reproduces the reference's double visit of a
 * neighbour cell when fewer than 3 cells span an axis. */
void oracle_lj_lnk(oracle_ctx *c, const double *x, int n, double *fr) {
  const int Cell_length = 16;
  int lc[3] = {c->NX / Cell_length, c->NY / Cell_length, c->NZ / Cell_length};
  double lc_r = Cell_length * c->DX;
  int lcyz = lc[1] * lc[2], lcxyz = lc[0] * lcyz;
  if (lcxyz <= 0) return;
  int *head = (int *)malloc(sizeof(int) * lcxyz);
  int *lscl = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int d = 0; d < lcxyz; d++) head[d] = -1;
  for (int i = 0; i < n; i++) {
    int mc[3];
    for (int d = 0; d < DIM; d++) mc[d] = (int)(x[3 * i + d] / lc_r);
    int cn = mc[0] * lcyz + mc[1] * lc[2] + mc[2];
    lscl[i] = head[cn];
    head[cn] = i;
  }
  int ic[3], icl[3], icls[3];
  for (ic[0] = 0; ic[0] < lc[0]; ic[0]++)
    for (ic[1] = 0; ic[1] < lc[1]; ic[1]++)
      for (ic[2] = 0; ic[2] < lc[2]; ic[2]++) {
        int cn = ic[0] * lcyz + ic[1] * lc[2] + ic[2];
        for (icl[0] = ic[0] - 1; icl[0] <= ic[0] + 1; icl[0]++)
          for (icl[1] = ic[1] - 1; icl[1] <= ic[1] + 1; icl[1]++)
            for (icl[2] = ic[2] - 1; icl[2] <= ic[2] + 1; icl[2]++) {
              for (int d = 0; d < DIM; d++) {
                icls[d] = icl[d];
                if (icls[d] < 0) icls[d] = lc[d] - 1;
                if (icls[d] > lc[d] - 1) icls[d] = 0;
              }
              int cl = ((icls[0] + lc[0]) % lc[0]) * lcyz + ((icls[1] + lc[1]) % lc[1]) * lc[2] +
                       ((icls[2] + lc[2]) % lc[2]);
              for (int i = head[cn]; i != -1; i = lscl[i])
                for (int j = head[cl]; j != -1; j = lscl[j])
                  if (i > j) lj_pair(c, &x[3 * i], &x[3 * j], &fr[3 * i], &fr[3 * j]);
            }
      }
  free(head);
  free(lscl);
}

/* Calc_f_Lennard_Jones_shear_cap_primitive, md_force.cxx:136-179. */
void oracle_lj_allpairs(oracle_ctx *c, const double *x, int n, double *fr) {
  for (int i = 0; i < n; i++)
    for (int m = i + 1; m < n; m++) lj_pair(c, &x[3 * i], &x[3 * m], &fr[3 * i], &fr[3 * m]);
}

/* Force, particle_solver.h:97-109, with Add_f_gravity md_force.cxx:181-191. */
int oracle_force(oracle_ctx *c) {
  if (c->P.lj_truncate >= 0 && c->Np > 0) {
    double *x = (double *)malloc(sizeof(double) * 3 * c->Np);
    double *fr = (double *)malloc(sizeof(double) * 3 * c->Np);
    for (int n = 0; n < c->Np; n++)
      for (int d = 0; d < DIM; d++) {
        x[3 * n + d] = c->p[n].x[d];
        fr[3 * n + d] = c->p[n].fr[d];
      }
    oracle_lj_lnk(c, x, c->Np, fr);
    for (int n = 0; n < c->Np; n++)
      for (int d = 0; d < DIM; d++) c->p[n].fr[d] = fr[3 * n + d];
    free(x);
    free(fr);
  }
  if (c->P.g != 0.0) {
    const double Gravity_on_fluid = c->P.g * c->RHO * 4. / 3. * M_PI * SQ(c->RADIUS) * c->RADIUS;
    for (int n = 0; n < c->Np; n++)
      c->p[n].fr[c->P.g_direction] -= Gravity_on_fluid * (c->P.mass_ratio[c->p[n].spec] - 1.0);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* quaternions, quaternion.h:55-250 and rigid_body.cxx:10-64           */
/* ------------------------------------------------------------------ */
static void qprod(spm_quaternion *q, const spm_quaternion *qa, const spm_quaternion *qb, double alpha) {
  spm_quaternion r;
  r.s = alpha * (qa->s * qb->s - (qa->v[0] * qb->v[0] + qa->v[1] * qb->v[1] + qa->v[2] * qb->v[2]));
  r.v[0] = alpha * (qa->s * qb->v[0] + qb->s * qa->v[0] + qa->v[1] * qb->v[2] - qa->v[2] * qb->v[1]);
  r.v[1] = alpha * (qa->s * qb->v[1] + qb->s * qa->v[1] + qa->v[2] * qb->v[0] - qa->v[0] * qb->v[2]);
  r.v[2] = alpha * (qa->s * qb->v[2] + qb->s * qa->v[2] + qa->v[0] * qb->v[1] - qa->v[1] * qb->v[0]);
  *q = r;
}
static void qnormalize(spm_quaternion *q) {
  double qnorm = sqrt(q->s * q->s + q->v[0] * q->v[0] + q->v[1] * q->v[1] + q->v[2] * q->v[2]);
  qnorm = 1.0 / qnorm;
  q->s *= qnorm;
  q->v[0] *= qnorm;
  q->v[1] *= qnorm;
  q->v[2] *= qnorm;
}
static void qadd(spm_quaternion *q, const spm_quaternion *qb, double alpha) {
  q->s += qb->s * alpha;
  q->v[0] += qb->v[0] * alpha;
  q->v[1] += qb->v[1] * alpha;
  q->v[2] += qb->v[2] * alpha;
}
/* rigid_body_rotation SPACE2BODY: conj(q) (0,v) q. */
static void space2body(double *rot, const double *orig, const spm_quaternion *q) {
  spm_quaternion qx = {0.0, {orig[0], orig[1], orig[2]}};
  spm_quaternion qc = {q->s, {-q->v[0], -q->v[1], -q->v[2]}};
  qprod(&qx, &qx, q, 1.0);
  qprod(&qc, &qc, &qx, 1.0);
  rot[0] = qc.v[0];
  rot[1] = qc.v[1];
  rot[2] = qc.v[2];
}

/* MD_solver_orientation_Euler / _AB2, particle_rotation_solver.h:15-39. */
static void orientation_euler(spm_particle *p, double dt) {
  spm_quaternion dqdt, qw = {0.0, {p->omega[0], p->omega[1], p->omega[2]}};
  p->q_old = p->q;
  qprod(&dqdt, &qw, &p->q, 0.5);
  qadd(&p->q, &dqdt, dt);
  qnormalize(&p->q);
}
static void orientation_ab2(spm_particle *p, double dt) {
  double wb[3], wb_old[3];
  space2body(wb, p->omega, &p->q);
  space2body(wb_old, p->omega_old, &p->q_old);
  for (int d = 0; d < DIM; d++) wb[d] = 3.0 * wb[d] - wb_old[d];
  spm_quaternion dqdt, qw = {0.0, {wb[0], wb[1], wb[2]}};
  p->q_old = p->q;
  qprod(&dqdt, &p->q, &qw, 0.5);
  qadd(&p->q, &dqdt, dt);
  qnormalize(&p->q);
}

/* PBC, periodic_boundary.h:9-15. */
static void pbc(const oracle_ctx *c, double *x) {
  for (int d = 0; d < DIM; d++) x[d] = fmod(x[d] + c->L[d], c->L[d]);
}

/* MD_solver_position_Euler / _AB2, particle_solver.cxx:36-88. */
int oracle_md_solver_position(oracle_ctx *c) {
  for (int n = 0; n < c->Np; n++) {
    spm_particle *p = &c->p[n];
    for (int d = 0; d < DIM; d++) {
      p->x_previous[d] = p->x[d];
      double delta_x = (c->ts == 0) ? c->dt * p->v[d] : c->hdt * (3.0 * p->v[d] - p->v_old[d]);
      p->x_nopbc[d] += delta_x;
      p->x[d] += delta_x;
    }
    pbc(c, p->x);
    if (c->P.rotation) {
      if (c->ts == 0)
        orientation_euler(p, c->dt);
      else
        orientation_ab2(p, c->hdt);
    }
  }
  return 0;
}

/* reset_Forces, particle_solver.cxx:11-34. */
static void reset_forces(oracle_ctx *c) {
  for (int n = 0; n < c->Np; n++) {
    spm_particle *p = &c->p[n];
    for (int d = 0; d < DIM; d++) {
      p->fr_previous[d] = p->fr[d];
      p->fr[d] = 0.0;
      p->f_hydro_previous[d] = p->f_hydro[d];
      p->f_hydro[d] = 0.0;
      p->f_slip_previous[d] = p->f_slip[d];
      p->f_slip[d] = 0.0;
      p->torque_hydro_previous[d] = p->torque_hydro[d];
      p->torque_hydro[d] = 0.0;
      p->torque_slip_previous[d] = p->torque_slip[d];
      p->torque_slip[d] = 0.0;
    }
  }
}

/* MD_solver_velocity_Euler / _AB2_hydro, particle_solver.cxx:106-147,
 * 247-289 (no self propulsion: self force and torque are 0). */
static void velocity_update(oracle_ctx *c);
int oracle_md_solver_velocity(oracle_ctx *c) {
  oracle_force(c);
  velocity_update(c);
  return 0;
}

static void velocity_update(oracle_ctx *c) {
  for (int n = 0; n < c->Np; n++) {
    spm_particle *p = &c->p[n];
    int s = p->spec;
    if (c->ts == 0) {
      double dmy = c->dt * c->IMASS[s], dmy_rot = c->dt * c->IMOI[s];
      for (int d = 0; d < DIM; d++) {
        p->v_old[d] = p->v[d];
        p->omega_old[d] = p->omega[d];
        p->v[d] += (dmy * (p->f_hydro[d] + p->fr[d] + 0.0));
        p->omega[d] += (dmy_rot * (p->torque_hydro[d] + 0.0));
      }
    } else {
      double dmy = c->hdt * c->IMASS[s], dmy_rot = c->hdt * c->IMOI[s];
      for (int d = 0; d < DIM; d++) {
        p->v_old[d] = p->v[d];
        p->omega_old[d] = p->omega[d];
        p->v[d] += dmy * (2.0 * (p->f_hydro[d] + 0.0) + p->fr[d] + p->fr_previous[d]);
        p->omega[d] += dmy_rot * 2.0 * (p->torque_hydro[d] + 0.0);
      }
    }
  }
  reset_forces(c);
}

/* ------------------------------------------------------------------ */
/* MT19937 (mt19937ar.c:65-171) and thermostat (fluct.cxx:10-36)       */
/* ------------------------------------------------------------------ */
static void init_genrand(oracle_ctx *c, unsigned long s) {
  c->mt[0] = (uint32_t)(s & 0xffffffffUL);
  for (c->mti = 1; c->mti < MT_N; c->mti++)
    c->mt[c->mti] = (uint32_t)(1812433253UL * (c->mt[c->mti - 1] ^ (c->mt[c->mti - 1] >> 30)) + c->mti);
}

uint32_t oracle_genrand_int32(oracle_ctx *c) {
  static const uint32_t mag01[2] = {0x0U, 0x9908b0dfU};
  uint32_t y;
  if (c->mti >= MT_N) {
    int kk;
    for (kk = 0; kk < MT_N - MT_M; kk++) {
      y = (c->mt[kk] & 0x80000000U) | (c->mt[kk + 1] & 0x7fffffffU);
      c->mt[kk] = c->mt[kk + MT_M] ^ (y >> 1) ^ mag01[y & 1U];
    }
    for (; kk < MT_N - 1; kk++) {
      y = (c->mt[kk] & 0x80000000U) | (c->mt[kk + 1] & 0x7fffffffU);
      c->mt[kk] = c->mt[kk + (MT_M - MT_N)] ^ (y >> 1) ^ mag01[y & 1U];
    }
    y = (c->mt[MT_N - 1] & 0x80000000U) | (c->mt[0] & 0x7fffffffU);
    c->mt[MT_N - 1] = c->mt[MT_M - 1] ^ (y >> 1) ^ mag01[y & 1U];
    c->mti = 0;
  }
  y = c->mt[c->mti++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680U;
  y ^= (y << 15) & 0xefc60000U;
  y ^= (y >> 18);
  return y;
}

double oracle_genrand_real3(oracle_ctx *c) {
  return (((double)oracle_genrand_int32(c)) + 0.5) * (1.0 / 4294967296.0);
}

/* Gauss2, fluct.h:48-60 (Marsaglia polar method). */
void oracle_gauss2(oracle_ctx *c, double out[2]) {
  double x1, x2, rsq;
  do {
    x1 = 2. * oracle_genrand_real3(c) - 1.;
    x2 = 2. * oracle_genrand_real3(c) - 1.;
    rsq = x1 * x1 + x2 * x2;
  } while (rsq >= 1.0 || rsq == 0.0);
  double factor = sqrt(-2. * log(rsq) / rsq);
  out[0] = factor * x1;
  out[1] = factor * x2;
}

/* Add_random_force_thermostat, fluct.cxx:10-36 (non-rigid branch). */
int oracle_add_random_force_thermostat(oracle_ctx *c) {
  if (!(c->kBT > 0.)) return 0;
  const double Zeta_drag = 6. * M_PI * c->ETA * c->RADIUS;
  const double Zeta_drag_rot = 8. * M_PI * c->ETA * POW3(c->RADIUS);
  const double sdv_v = sqrt(Zeta_drag * c->dt * c->kBT * c->P.alpha_v);
  const double sdv_omega = sqrt(Zeta_drag_rot * c->dt * c->kBT * c->P.alpha_o);
  const double noise_intensity_v = 1.439 * sdv_v;
  const double noise_intensity_o = 1.22 * sdv_omega;
  for (int n = 0; n < c->Np; n++) {
    double dmy[6];
    oracle_gauss2(c, dmy);
    oracle_gauss2(c, dmy + 2);
    oracle_gauss2(c, dmy + 4);
    spm_particle *p = &c->p[n];
    double imass = c->IMASS[p->spec], imoi = c->IMOI[p->spec];
    for (int d = 0; d < DIM; d++) {
      p->v[d] += dmy[d] * noise_intensity_v * imass;
      if (c->P.rotation) p->omega[d] += dmy[d + 3] * noise_intensity_o * imoi;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* the step                                                            */
/* ------------------------------------------------------------------ */
/* sp_3d_ns.cxx:67-79: FIX_CELL, truncation, Zeta_k2u. */
int oracle_zeta_k2u(oracle_ctx *c) {
  for (int d = 0; d < DIM; d++)
    if (c->P.fix_cell[d]) c->uk_dc[d] = 0.0;
  for (int d = 0; d < DIM - 1; d++) oracle_truncate(c, c->zeta[d]);
  zeta_k2u(c);
  return 0;
}

/* U2zeta_k, operate_omega.h:186-189. */
int oracle_u2zeta_k(oracle_ctx *c) {
  u2u_k(c, c->u);
  oracle_u_k2zeta_k(c, c->u, c->zeta, c->uk_dc);
  return 0;
}

/* Init_zeta_k (init_fluid.cxx:10-23) and sp_3d_ns.cxx:508-524. */
int oracle_init_fluid(oracle_ctx *c) {
  for (int d = 0; d < DIM; d++) memset(c->u[d], 0, sizeof(double) * c->nfield);
  u2u_k(c, c->u);
  oracle_solenoidal_uk(c, c->u);
  oracle_u_k2zeta_k(c, c->u, c->zeta, c->uk_dc);
  oracle_make_phi_u_particle(c);
  zeta_k2u(c);
  oracle_make_f_particle_dt_sole(c);
  oracle_u2zeta_k(c);
  for (int d = 0; d < DIM; d++) c->uk_dc[d] = 0.0;
  return 0;
}

/* ------------------------------------------------------------------ */
/* electrolyte: Nernst-Planck ions + Poisson (SW_EQ == Electrolyte)    */
/* ------------------------------------------------------------------ */
static const int (*ijk_blocks(const oracle_ctx *c, int rng[4][6]))[6] {
  const int NX = c->NX, NY = c->NY, TX = c->TRN_X, TY = c->TRN_Y, TZ = c->TRN_Z;
  const int r[4][6] = {{0, TX - 1, 0, TY - 1, 0, 2 * TZ - 1},
                       {0, TX - 1, NY - TY + 1, NY - 1, 0, 2 * TZ - 1},
                       {NX - TX + 1, NX - 1, 0, TY - 1, 0, 2 * TZ - 1},
                       {NX - TX + 1, NX - 1, NY - TY + 1, NY - 1, 0, 2 * TZ - 1}};
  memcpy(rng, r, sizeof(r));
  return (const int (*)[6])rng;
}

/* DPhi_compact_sin, profile.h:168-181. */
static double dphi_compact_sin(const oracle_ctx *c, double x) {
  const double hxi = c->XI * .5, ixi = 1. / c->XI;
  const double dmy_x = c->RADIUS - x;
  return fabs(dmy_x) < hxi ? M_PI * .5 * ixi * cos(M_PI * dmy_x * ixi) : 0.;
}

/* Make_phi_qq_particle, operate_electrolyte.cxx:179-242: phi (no clamp)
 * and the surface charge profile, normalised so that its volume integral
 * of |.| is the total |surface charge|. */
void oracle_make_phi_qq(oracle_ctx *c, double *phi, double *surface) {
  double abs_total = 0.;
  for (int n = 0; n < c->Np; n++) {
    const spm_particle *pn = &c->p[n];
    const double q = c->surface_charge_e[pn->spec];
    int x_int[3], r_mesh[3];
    double residue[3], r[3], x[3], x12[3];
    particle_cell(c, pn->x, x_int, residue);
    for (int mesh = 0; mesh < c->np_domain; mesh++) {
      relative_coord(c, &c->sekibun[3 * mesh], x_int, residue, r_mesh, r);
      for (int d = 0; d < DIM; d++) x[d] = r_mesh[d] * c->DX;
      const double dmy = distance0(c, x, pn->x, x12);
      const size_t im = IDX(c, r_mesh[0], r_mesh[1], r_mesh[2]);
      phi[im] += Phi(c, dmy, c->RADIUS);
      surface[im] += q * dphi_compact_sin(c, dmy);
    }
    abs_total += fabs(q);
  }
  double dmy = 0.0;
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++)
      for (int k = 0; k < c->NZ; k++) dmy += fabs(surface[IDX(c, i, j, k)]);
  const double rescale = abs_total / (dmy * c->DX3);
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++)
      for (int k = 0; k < c->NZ; k++) surface[IDX(c, i, j, k)] *= rescale;
}

/* Make_surface_normal, make_phi.cxx:17-58. */
void oracle_make_surface_normal(oracle_ctx *c, double *const normal[3]) {
  for (int d = 0; d < DIM; d++) memset(normal[d], 0, sizeof(double) * c->nfield);
  for (int n = 0; n < c->Np; n++) {
    int x_int[3], r_mesh[3];
    double residue[3], r[3];
    particle_cell(c, c->p[n].x, x_int, residue);
    for (int mesh = 0; mesh < c->np_domain; mesh++) {
      relative_coord(c, &c->sekibun[3 * mesh], x_int, residue, r_mesh, r);
      const double dmy_r = sqrt(SQ(r[0]) + SQ(r[1]) + SQ(r[2]));
      if (fabs(dmy_r - c->RADIUS) < c->HXI) {
        const double ir = 1. / dmy_r;
        const size_t im = IDX(c, r_mesh[0], r_mesh[1], r_mesh[2]);
        for (int d = 0; d < DIM; d++) normal[d][im] += r[d] * ir;
      }
    }
  }
}

/* A_k2a_out, fft_wrapper.h:639-651. */
static void a_k2a_out(oracle_ctx *c, const double *a_k, double *a_x) {
  memcpy(a_x, a_k, sizeof(double) * c->nfield);
  oracle_A_k2a(c, a_x);
}

/* Conc_k2charge_field, operate_electrolyte.cxx:45-74. */
static void conc_k2charge_field(oracle_ctx *c, double *charge, double *phi_p, double *dmy_value) {
  memset(phi_p, 0, sizeof(double) * c->nfield);
  memset(charge, 0, sizeof(double) * c->nfield);
  oracle_make_phi_qq(c, phi_p, charge);
  for (int n = 0; n < c->E.n_spec; n++) {
    a_k2a_out(c, c->conc[n], dmy_value);
    for (int i = 0; i < c->NX; i++)
      for (int j = 0; j < c->NY; j++)
        for (int k = 0; k < c->NZ; k++) {
          const size_t im = IDX(c, i, j, k);
          charge[im] += c->valency_e[n] * dmy_value[im] * (1. - phi_p[im]);
        }
  }
}

/* Charge_field_k2Coulomb_potential_k_PBC, operate_electrolyte.cxx:76-88. */
static void coulomb_potential_k(oracle_ctx *c, double *pot) {
  const double ieps = 1. / c->E.dielectric;
  for (size_t im = 0; im < c->nfield; im++) pot[im] *= (c->IK2[im] * ieps);
}

/* A_k2da_k, fft_wrapper.h:681-685 over A_k2dja_k_primitive,
 * fft_wrapper.cxx:103-126: (re, im) -> (k im, -k re). */
static void a_k2da_k(oracle_ctx *c, const double *a, double *const da[3]) {
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++)
      for (int k = 0; k < c->HNZ_; k++) {
        const size_t im = IDX(c, i, j, 2 * k);
        const double ks[3] = {c->KX_int[im] * c->WAVE_X, c->KY_int[im] * c->WAVE_Y, c->KZ_int[im] * c->WAVE_Z};
        for (int d = 0; d < DIM; d++) {
          const double re = a[im], imv = a[im + 1];
          da[d][im] = ks[d] * imv;
          da[d][im + 1] = -ks[d] * re;
        }
      }
}

/* U_k2divergence_k, fft_wrapper.cxx:649-677. */
static void u_k2divergence_k(oracle_ctx *c, double *const u[3], double *div) {
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++)
      for (int k = 0; k < c->HNZ_; k++) {
        const size_t im0 = IDX(c, i, j, 2 * k), im1 = im0 + 1;
        const double ks0 = c->KX_int[im0] * c->WAVE_X, ks1 = c->KY_int[im0] * c->WAVE_Y,
                     ks2 = c->KZ_int[im0] * c->WAVE_Z;
        const double d0 = ks0 * u[0][im1] + ks1 * u[1][im1] + ks2 * u[2][im1];
        const double d1 = -ks0 * u[0][im0] - ks1 * u[1][im0] - ks2 * u[2][im0];
        div[im0] = d0;
        div[im1] = d1;
      }
}

/* the applied field's amplitude: E_ext, times sin(w t) for AC
 * (operate_electrolyte.cxx:116-123) */
static void external_field_now(const oracle_ctx *c, double ext[3]) {
  for (int d = 0; d < DIM; d++) {
    ext[d] = c->E.e_ext[d];
    if (c->E.ac) ext[d] *= sin(c->E.angular_frequency * c->time);
  }
}

/* Count_single_solute, solute_rhs.cxx:128-147. */
static double count_single_solute(const oracle_ctx *c, const double *conc_x, const double *phi_p) {
  double dmy = 0.;
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++)
      for (int k = 0; k < c->NZ; k++) {
        const size_t im = IDX(c, i, j, k);
        dmy += (1. - phi_p[im]) * conc_x[im];
      }
  return dmy * c->DX3;
}

/* Rhs_NS_Nernst_Planck + Rhs_NS_solute, fluid_solver.cxx:181-280 and
 * solute_rhs.cxx:42-235.  Working memory as NSsolute_solver_Euler passes
 * it (:306-326): u, solute_flux = up, grad_potential = f (f_particle),
 * rhs_ns = f_ns1, rhs_solute = Concentration_rhs0. */
static void rhs_ns_nernst_planck(oracle_ctx *c) {
  double *const *u = c->u, *const *flux = c->up, *const *grad = c->f;
  /* potential gradient in real space (fluid_solver.cxx:263-274) */
  conc_k2charge_field(c, u[0], u[1], u[2]);
  oracle_A2a_k(c, u[0]);
  coulomb_potential_k(c, u[0]);
  oracle_truncate(c, u[0]);
  a_k2da_k(c, u[0], grad);
  { /* Add_constant_field_k, fluid_solver.cxx:228-241 */
    double amp = (double)(c->NX * c->NY * c->NZ);
    if (c->E.ac) amp = sin(c->E.angular_frequency * c->time);
    for (int d = 0; d < DIM; d++) grad[d][0] -= c->E.e_ext[d] * amp;
  }
  u_k2u(c, grad);
  /* Rhs_NS_solute, fluid_solver.cxx:181-227 */
  for (int d = 0; d < DIM - 1; d++) oracle_truncate(c, c->zeta[d]);
  zeta_k2u(c);
  oracle_make_surface_normal(c, c->surface_normal);
  for (int n = 0; n < c->E.n_spec; n++) {
    oracle_truncate(c, c->conc[n]);
    { /* Diffusion_flux_single, solute_rhs.cxx:216-235 (rhs_ns[0] as working memory) */
      double *dmy = c->f_ns1[0];
      for (size_t im = 0; im < c->nfield; im++) dmy[im] = c->kBT * c->E.onsager[n] * c->conc[n][im];
      a_k2da_k(c, dmy, flux);
      u_k2u(c, flux);
    }
    double *cx = c->conc_rhs[n];
    a_k2a_out(c, c->conc[n], cx);
    /* Solute_solver_rhs_nonlinear_x_single, solute_rhs.cxx:149-187 */
    for (int i = 0; i < c->NX; i++)
      for (int j = 0; j < c->NY; j++)
        for (int k = 0; k < c->NZ; k++) {
          const size_t im = IDX(c, i, j, k);
          for (int d = 0; d < DIM; d++) {
            const double dmy_interaction = grad[d][im] * c->valency_e[n];
            flux[d][im] += cx[im] * (c->E.onsager[n] * dmy_interaction);
          }
        }
    /* Solute_impermeability, solute_rhs.cxx:42-95 */
    for (int p = 0; p < c->Np; p++) {
      int x_int[3], r_mesh[3];
      double residue[3], r[3];
      particle_cell(c, c->p[p].x, x_int, residue);
      for (int mesh = 0; mesh < c->np_domain; mesh++) {
        relative_coord(c, &c->sekibun[3 * mesh], x_int, residue, r_mesh, r);
        const size_t im = IDX(c, r_mesh[0], r_mesh[1], r_mesh[2]);
        double norm2 = 0.;
        for (int d = 0; d < DIM; d++) norm2 += SQ(c->surface_normal[d][im]);
        if (norm2 > 0.) {
          double dmy = 0.;
          for (int d = 0; d < DIM; d++) dmy += flux[d][im] * c->surface_normal[d][im];
          dmy /= norm2;
          for (int d = 0; d < DIM; d++) flux[d][im] -= c->surface_normal[d][im] * dmy;
        }
      }
    }
    /* Add_advection_flux, solute_rhs.cxx:188-207 */
    for (int i = 0; i < c->NX; i++)
      for (int j = 0; j < c->NY; j++)
        for (int k = 0; k < c->NZ; k++) {
          const size_t im = IDX(c, i, j, k);
          for (int d = 0; d < DIM; d++) flux[d][im] += cx[im] * -u[d][im];
        }
    u2u_k(c, flux);
    u_k2divergence_k(c, flux, cx);
  }
  /* Rhs_solvent (fluid_solver.cxx:168-179) and Add_zeta_viscous_term
   * (operate_omega.cxx:172-184) */
  oracle_u2advection_k(c, u, c->f_ns1);
  int rng_buf[4][6];
  const int (*rng)[6] = ijk_blocks(c, rng_buf);
  for (int b = 0; b < 4; b++)
    for (int i = rng[b][0]; i <= rng[b][1]; i++)
      for (int j = rng[b][2]; j <= rng[b][3]; j++)
        for (int k = rng[b][4]; k <= rng[b][5]; k++) {
          const size_t im = IDX(c, i, j, k);
          c->f_ns1[0][im] += -(c->NU * c->K2[im] * c->zeta[0][im]);
          c->f_ns1[1][im] += -(c->NU * c->K2[im] * c->zeta[1][im]);
        }
}

/* NSsolute_solver_Euler, fluid_solver.cxx:306-326, with Field_solver_Euler
 * (:122-135) over the ijk_range blocks. */
int oracle_nssolute_solver_euler(oracle_ctx *c) {
  if (!c->electro) return -1;
  rhs_ns_nernst_planck(c);
  int rng_buf[4][6];
  const int (*rng)[6] = ijk_blocks(c, rng_buf);
  for (int b = 0; b < 4; b++)
    for (int i = rng[b][0]; i <= rng[b][1]; i++)
      for (int j = rng[b][2]; j <= rng[b][3]; j++)
        for (int k = rng[b][4]; k <= rng[b][5]; k++) {
          const size_t im = IDX(c, i, j, k);
          for (int d = 0; d < DIM - 1; d++) c->zeta[d][im] += (c->dt * c->f_ns1[d][im]);
          for (int n = 0; n < c->E.n_spec; n++) c->conc[n][im] += (c->dt * c->conc_rhs[n][im]);
        }
  /* uk_dc += dt * rhs_uk_dc with rhs_uk_dc = 0 (Rhs_solvent) */
  return 0;
}

/* Rescale_solute, solute_rhs.cxx:97-126 (phi and up[0] as working memory,
 * as Time_evolution_hydro passes them, sp_3d_ns.cxx:90-93). */
int oracle_rescale_solute(oracle_ctx *c) {
  if (!c->electro) return -1;
  oracle_make_phi_particle(c);
  for (int n = 0; n < c->E.n_spec; n++) {
    a_k2a_out(c, c->conc[n], c->up[0]);
    const double dmy = count_single_solute(c, c->up[0], c->phi);
    const double factor = c->total_solute[n] / dmy;
    for (size_t im = 0; im < c->nfield; im++) c->conc[n][im] *= factor;
  }
  return 0;
}

/* Make_Coulomb_force_x_on_fluid, operate_electrolyte.cxx:90-177 (free
 * particles): force = rho_e (-grad Phi + E_ext), into f. */
int oracle_make_coulomb_force(oracle_ctx *c) {
  if (!c->electro) return -1;
  double *charge = c->up[0], *pot = c->up[1];
  conc_k2charge_field(c, charge, c->f[0], c->f[1]);
  memcpy(pot, charge, sizeof(double) * c->nfield); /* A2a_k_out */
  oracle_A2a_k(c, pot);
  coulomb_potential_k(c, pot);
  a_k2da_k(c, pot, c->f);
  u_k2u(c, c->f);
  double ext[3];
  external_field_now(c, ext);
  for (int d = 0; d < DIM; d++)
    for (int i = 0; i < c->NX; i++)
      for (int j = 0; j < c->NY; j++)
        for (int k = 0; k < c->NZ; k++) {
          const size_t im = IDX(c, i, j, k);
          double electric_field = -c->f[d][im];
          if (c->E.external_field) electric_field += ext[d];
          c->f[d][im] = charge[im] * electric_field;
        }
  return 0;
}

int oracle_set_electrolyte(oracle_ctx *c, const spm_electrolyte *el) {
  if (!el || c->electro || el->n_spec < 1 || el->n_spec > 2 || !(el->dielectric > 0)) return -1;
  c->E = *el;
  int fail = 0;
  for (int n = 0; n < 2; n++) {
    c->conc[n] = (double *)calloc(c->nfield, sizeof(double));
    c->conc_rhs[n] = (double *)calloc(c->nfield, sizeof(double));
    c->f_ns1[n] = (double *)calloc(c->nfield, sizeof(double));
    fail |= !c->conc[n] || !c->conc_rhs[n] || !c->f_ns1[n];
  }
  for (int d = 0; d < DIM; d++) {
    c->surface_normal[d] = (double *)calloc(c->nfield, sizeof(double));
    fail |= !c->surface_normal[d];
  }
  if (fail) return -2;
  for (int s = 0; s < SPM_MAX_SPECIES; s++) c->surface_charge_e[s] = el->surface_charge[s] * el->elementary_charge;
  for (int n = 0; n < el->n_spec; n++) c->valency_e[n] = el->valency[n] * el->elementary_charge;
  c->electro = 1;
  return 0;
}

/* Init_rho_ion, operate_electrolyte.cxx:695-815, uniform initial
 * distribution (Poisson_Boltzmann off): Set_uniform_ion_charge_density_
 * nosalt (:370-403) or _salt (:522-575), then Count_solute_each
 * (solute_rhs.h:143-156). */
int oracle_init_rho_ion(oracle_ctx *c) {
  if (!c->electro) return -1;
  const spm_electrolyte *E = &c->E;
  int count[SPM_MAX_SPECIES] = {0};
  for (int n = 0; n < c->Np; n++) count[c->p[n].spec]++;
  double surface_ion_number = 0.;
  for (int s = 0; s < c->P.n_species; s++) {
    if (E->surface_charge[s] == 0.) return -3;
    surface_ion_number -= E->surface_charge[s] * count[s];
  }
  oracle_make_phi_particle(c);
  double volume_phi = 0.;
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++)
      for (int k = 0; k < c->NZ; k++) volume_phi += (1. - c->phi[IDX(c, i, j, k)]);
  double dens[2];
  if (E->n_spec == 1) {
    if (c->P.n_species == 1 && E->valency[0] * E->surface_charge[0] >= 0.) return -3;
    const double counterion_number = surface_ion_number / E->valency[0];
    dens[0] = counterion_number / (volume_phi * c->DX3);
  } else {
    if (E->valency[0] <= 0. || E->valency[1] >= 0.) return -3;
    double pos = 0., neg = 0.;
    for (int s = 0; s < c->P.n_species; s++) {
      const double q = E->surface_charge[s], np = (double)count[s];
      if (q < 0.)
        pos += fabs(q * np / E->valency[0]);
      else if (q > 0.)
        neg += fabs(q * np / E->valency[1]);
    }
    const double rho_inf = c->kBT * E->dielectric / SQ(E->elementary_charge * E->debye_length);
    const double rho_pos = rho_inf / (E->valency[0] * (E->valency[0] - E->valency[1]));
    const double rho_neg = -rho_inf / (E->valency[1] * (E->valency[0] - E->valency[1]));
    dens[0] = pos / (volume_phi * c->DX3) + rho_pos;
    dens[1] = neg / (volume_phi * c->DX3) + rho_neg;
  }
  for (int n = 0; n < E->n_spec; n++) {
    for (int i = 0; i < c->NX; i++)
      for (int j = 0; j < c->NY; j++)
        for (int k = 0; k < c->NZ; k++) c->conc[n][IDX(c, i, j, k)] = dens[n];
    c->total_solute[n] = count_single_solute(c, c->conc[n], c->phi);
    oracle_A2a_k(c, c->conc[n]);
  }
  /* Count_solute_each (phi and up[0] as working memory) */
  oracle_make_phi_particle(c);
  for (int n = 0; n < E->n_spec; n++) {
    a_k2a_out(c, c->conc[n], c->up[0]);
    c->total_solute[n] = count_single_solute(c, c->up[0], c->phi);
  }
  return 0;
}

/* Calc_free_energy_PB, operate_electrolyte.cxx:302-371 (the particles'
 * phi and the scratch fields as the library uses them). */
static double free_energy_pb(oracle_ctx *c) {
  double *charge = c->up[1], *val = c->up[2];
  double ideal = 0.0, es = 0.0;
  oracle_make_phi_particle(c);
  for (int n = 0; n < c->E.n_spec; n++) {
    a_k2a_out(c, c->conc[n], val);
    for (int i = 0; i < c->NX; i++)
      for (int j = 0; j < c->NY; j++)
        for (int k = 0; k < c->NZ; k++) {
          const size_t im = IDX(c, i, j, k);
          ideal += (1. - c->phi[im]) * val[im] * (log(val[im]) - 1.);
        }
  }
  ideal *= (c->kBT * c->DX3);
  conc_k2charge_field(c, charge, c->up[0], val);
  memcpy(val, charge, sizeof(double) * c->nfield); /* A2a_k_out */
  oracle_A2a_k(c, val);
  coulomb_potential_k(c, val);
  oracle_A_k2a(c, val);
  double ext[3];
  external_field_now(c, ext);
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++)
      for (int k = 0; k < c->NZ; k++) {
        const size_t im = IDX(c, i, j, k);
        val[im] *= 0.5;
        if (c->E.external_field) val[im] -= ((ext[0] * (double)i + ext[1] * (double)j + ext[2] * (double)k) * c->DX);
        es += charge[im] * val[im];
      }
  es *= c->DX3;
  return ideal + es;
}

/* Init_rho_ion with Poisson_Boltzmann on, salt, more than one particle
 * (operate_electrolyte.cxx:786-805): the uniform start, then
 * Set_steadystate_ion_density (:374-412): Ion_diffusion_solver_Euler
 * (fluid_solver.cxx:281-304) with u = zeta = uk_dc = 0 until the free
 * energy changes by less than tol (the reference's TOL = 1e-4,
 * operate_electrolyte.h:21; tests also run it tighter), at most 9999
 * iterations. */
int oracle_init_rho_ion_pb(oracle_ctx *c, double tol, int *iterations) {
  if (!c->electro || c->E.n_spec != 2 || c->Np < 2) return -1;
  int rc = oracle_init_rho_ion(c);
  if (rc) return rc;
  double *z0 = (double *)calloc(c->nfield, sizeof(double));
  double *z1 = (double *)calloc(c->nfield, sizeof(double));
  if (!z0 || !z1) return -2;
  double *zsave[2] = {c->zeta[0], c->zeta[1]};
  double dcsave[3] = {c->uk_dc[0], c->uk_dc[1], c->uk_dc[2]};
  double f_prev = free_energy_pb(c);
  int m;
  for (m = 1; m < 10000; m++) {
    memset(z0, 0, sizeof(double) * c->nfield);
    memset(z1, 0, sizeof(double) * c->nfield);
    c->zeta[0] = z0;
    c->zeta[1] = z1;
    for (int d = 0; d < DIM; d++) c->uk_dc[d] = 0.0;
    rhs_ns_nernst_planck(c);
    int rng_buf[4][6];
    const int (*rng)[6] = ijk_blocks(c, rng_buf);
    for (int b = 0; b < 4; b++)
      for (int i = rng[b][0]; i <= rng[b][1]; i++)
        for (int j = rng[b][2]; j <= rng[b][3]; j++)
          for (int k = rng[b][4]; k <= rng[b][5]; k++) {
            const size_t im = IDX(c, i, j, k);
            for (int n = 0; n < c->E.n_spec; n++) c->conc[n][im] += (c->dt * c->conc_rhs[n][im]);
          }
    c->zeta[0] = zsave[0];
    c->zeta[1] = zsave[1];
    for (int d = 0; d < DIM; d++) c->uk_dc[d] = dcsave[d];
    oracle_rescale_solute(c);
    const double f = free_energy_pb(c);
    if (fabs(f - f_prev) < tol) break;
    f_prev = f;
  }
  free(z0);
  free(z1);
  if (iterations) *iterations = m < 10000 ? m : 9999;
  /* Count_solute_each */
  oracle_make_phi_particle(c);
  for (int n = 0; n < c->E.n_spec; n++) {
    a_k2a_out(c, c->conc[n], c->up[0]);
    c->total_solute[n] = count_single_solute(c, c->up[0], c->phi);
  }
  return 0;
}

int oracle_put_concentration(oracle_ctx *c, int n, const double *host) {
  if (!c->electro || n < 0 || n >= c->E.n_spec) return -1;
  memcpy(c->conc[n], host, sizeof(double) * c->nfield);
  return 0;
}
int oracle_get_concentration(oracle_ctx *c, int n, double *host) {
  if (!c->electro || n < 0 || n >= c->E.n_spec) return -1;
  memcpy(host, c->conc[n], sizeof(double) * c->nfield);
  return 0;
}
int oracle_put_total_solute(oracle_ctx *c, const double *t) {
  if (!c->electro) return -1;
  for (int n = 0; n < c->E.n_spec; n++) c->total_solute[n] = t[n];
  return 0;
}
int oracle_get_total_solute(oracle_ctx *c, double *t) {
  if (!c->electro) return -1;
  for (int n = 0; n < c->E.n_spec; n++) t[n] = c->total_solute[n];
  return 0;
}

/* Time_evolution_hydro with SW_EQ == Electrolyte (sp_3d_ns.cxx:46-57,
 * 61-212): no thermostat (:183); the first Calc_f_hydro reads the f field
 * (the potential gradient left by the solver) into f_hydro1 (:97-106). */
static void electrolyte_step(oracle_ctx *c) {
  oracle_nssolute_solver_euler(c);
  oracle_zeta_k2u(c);
  oracle_md_solver_position(c);
  oracle_rescale_solute(c);
  calc_f_hydro_field(c, c->f);
  for (int n = 0; n < c->Np; n++)
    for (int d = 0; d < DIM; d++) {
      c->p[n].f_hydro1[d] = c->p[n].f_hydro[d];
      c->p[n].torque_hydro1[d] = c->p[n].torque_hydro[d];
    }
  oracle_make_coulomb_force(c);
  const double dmy = c->dt * c->IRHO;
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++)
      for (int k = 0; k < c->NZ; k++) {
        const size_t im = IDX(c, i, j, k);
        for (int d = 0; d < DIM; d++) c->u[d][im] += (c->f[d][im] * dmy);
      }
  solenoidal_u(c, c->u);
  oracle_calc_f_hydro(c);
  oracle_md_solver_velocity(c);
  oracle_make_phi_u_particle(c);
  oracle_make_f_particle_dt_sole(c);
  oracle_u2zeta_k(c);
}

/* Time_evolution_hydro (sp_3d_ns.cxx:61-212) for Navier_Stokes with free
 * spherical particles, then jikan.time += dt and ts++ (:593-594). */
static void le_step(oracle_ctx *c);

int oracle_step(oracle_ctx *c, int nsteps) {
  for (int s = 0; s < nsteps; s++) {
    if (c->electro) {
      electrolyte_step(c);
      c->time += c->dt;
      c->ts++;
      continue;
    }
    if (c->le) {
      le_step(c);
      c->time += c->dt;
      c->ts++;
      continue;
    }
    oracle_ns_solver_slaved_euler(c);
    oracle_zeta_k2u(c);
    oracle_md_solver_position(c);
    oracle_calc_f_hydro(c);
    oracle_md_solver_velocity(c);
    if (c->kBT > 0.) oracle_add_random_force_thermostat(c);
    oracle_make_phi_u_particle(c);
    oracle_make_f_particle_dt_sole(c);
    oracle_u2zeta_k(c);
    c->time += c->dt;
    c->ts++;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Lees-Edwards shear in oblique coordinates, fluid only              */
/* (Time_evolution_hydro_OBL, sp_3d_ns.cxx:214-390)                    */
/* ------------------------------------------------------------------ */
/* contra2co_single / co2contra_single, fft_wrapper.h:398-424 */
static void contra2co_single(double deg, double v[3]) {
  const double d0 = v[0], d1 = v[1], d2 = v[2];
  v[0] = d0 + deg * d1;
  v[1] = deg * d0 + (1. + deg * deg) * d1;
  v[2] = d2;
}
static void co2contra_single(double deg, double v[3]) {
  const double d0 = v[0], d1 = v[1], d2 = v[2];
  v[0] = (1. + deg * deg) * d0 - deg * d1;
  v[1] = -deg * d0 + d1;
  v[2] = d2;
}

/* K2 / IK2 for the current degree_oblique, sp_3d_ns.cxx:319-334 */
static void obl_update_k2(oracle_ctx *c) {
  const double deg = c->degree_oblique;
  for (size_t im = 0; im < c->nfield; im++) {
    c->K2[im] = SQ(c->WAVE_X * c->KX_int[im]) + SQ(c->WAVE_Y * c->KY_int[im] - c->WAVE_X * deg * c->KX_int[im]) +
                SQ(c->WAVE_Z * c->KZ_int[im]);
    c->IK2[im] = c->K2[im] > 0.0 ? 1. / c->K2[im] : 0.0;
  }
}

/* Zeta_k2u_k_OBL, fft_wrapper.cxx:513-584 (u contravariant) */
static void zeta_k2u_k_obl(oracle_ctx *c, double *const zeta[2], const double uk_dc[3], double *const u[3]) {
  const double deg = c->degree_oblique;
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++)
      for (int k = 0; k < c->HNZ_; k++) {
        const size_t im = IDX(c, i, j, 2 * k);
        double kx = c->WAVE_X * c->KX_int[im], ky = c->WAVE_Y * c->KY_int[im], kz = c->WAVE_Z * c->KZ_int[im];
        const double ik2 = c->IK2[im];
        const double d1r = zeta[0][im], d1i = zeta[0][im + 1], d2r = zeta[1][im], d2i = zeta[1][im + 1];
        double wr[3], wi[3];
        if (c->KX_int[im] != 0) {
          wr[0] = -(1. / kx) * (ky * d1r + kz * d2r);
          wi[0] = -(1. / kx) * (ky * d1i + kz * d2i);
          wr[1] = d1r; wi[1] = d1i;
          wr[2] = d2r; wi[2] = d2i;
        } else if (c->KZ_int[im] != 0) {
          const double dmy = -(ky / kz);
          wr[0] = d2r; wi[0] = d2i;
          wr[1] = d1r; wi[1] = d1i;
          wr[2] = dmy * d1r; wi[2] = dmy * d1i;
        } else {
          wr[0] = d1r; wi[0] = d1i;
          wr[1] = 0.0; wi[1] = 0.0;
          wr[2] = d2r; wi[2] = d2i;
        }
        contra2co_single(deg, wr);
        contra2co_single(deg, wi);
        kx *= ik2;
        ky *= ik2;
        kz *= ik2;
        u[0][im] = (ky * wi[2] - kz * wi[1]);
        u[0][im + 1] = -(ky * wr[2] - kz * wr[1]);
        u[1][im] = (kz * wi[0] - kx * wi[2]);
        u[1][im + 1] = -(kz * wr[0] - kx * wr[2]);
        u[2][im] = (kx * wi[1] - ky * wi[0]);
        u[2][im + 1] = -(kx * wr[1] - ky * wr[0]);
      }
  u[0][0] = uk_dc[0];
  u[1][0] = uk_dc[1];
  u[2][0] = uk_dc[2];
}

/* Zeta_k2omega_k_OBL, fft_wrapper.cxx:586-632 (every interleaved slot) */
static void zeta_k2omega_k_obl(oracle_ctx *c, double *const zeta[2], double *const omega[3]) {
  for (size_t im = 0; im < c->nfield; im++) {
    const double kx = c->WAVE_X * c->KX_int[im], ky = c->WAVE_Y * c->KY_int[im], kz = c->WAVE_Z * c->KZ_int[im];
    const double d1 = zeta[0][im], d2 = zeta[1][im];
    if (c->KX_int[im] != 0) {
      omega[0][im] = -(1. / kx) * (ky * d1 + kz * d2);
      omega[1][im] = d1;
      omega[2][im] = d2;
    } else if (c->KZ_int[im] != 0) {
      const double dmy = -(ky / kz);
      omega[0][im] = d2;
      omega[1][im] = d1;
      omega[2][im] = dmy * d1;
    } else {
      omega[0][im] = d1;
      omega[1][im] = 0.0;
      omega[2][im] = d2;
    }
  }
  for (int d = 0; d < DIM; d++) omega[d][0] = 0.0;
}

/* U_k2rotation_k, fft_wrapper.cxx:679-718 (in place, Cartesian ks) */
static void u_k2rotation_k(oracle_ctx *c, double *const u[3]) {
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++)
      for (int k = 0; k < c->HNZ_; k++) {
        const size_t im = IDX(c, i, j, 2 * k);
        const double ks[3] = {c->KX_int[im] * c->WAVE_X, c->KY_int[im] * c->WAVE_Y, c->KZ_int[im] * c->WAVE_Z};
        double re[3], imv[3], rr[3], ri[3];
        for (int d = 0; d < DIM; d++) {
          re[d] = u[d][im];
          imv[d] = u[d][im + 1];
        }
        rr[0] = (ks[1] * imv[2] - ks[2] * imv[1]);
        ri[0] = -(ks[1] * re[2] - ks[2] * re[1]);
        rr[1] = (ks[2] * imv[0] - ks[0] * imv[2]);
        ri[1] = -(ks[2] * re[0] - ks[0] * re[2]);
        rr[2] = (ks[0] * imv[1] - ks[1] * imv[0]);
        ri[2] = -(ks[0] * re[1] - ks[1] * re[0]);
        for (int d = 0; d < DIM; d++) {
          u[d][im] = rr[d];
          u[d][im + 1] = ri[d];
        }
      }
}

/* Omega_k2zeta_k_OBL, fft_wrapper.cxx:175-198 (every interleaved slot) */
static void omega_k2zeta_k_obl(oracle_ctx *c, double *const omega[3], double *const zetak[2]) {
  for (size_t im = 0; im < c->nfield; im++) {
    if (c->KX_int[im] != 0) {
      zetak[0][im] = omega[1][im];
      zetak[1][im] = omega[2][im];
    } else if (c->KZ_int[im] != 0) {
      zetak[0][im] = omega[1][im];
      zetak[1][im] = omega[0][im];
    } else {
      zetak[0][im] = omega[0][im];
      zetak[1][im] = omega[2][im];
    }
  }
}

/* Zeta_k2advection_k_OBL, operate_omega.cxx:115-173 */
static void zeta_k2advection_k_obl(oracle_ctx *c, double *const adv[2]) {
  double *const *u = c->u;
  for (int d = 0; d < DIM - 1; d++) oracle_truncate(c, c->zeta[d]);
  /* Zeta_k2u_cpuky, operate_omega.h:144-156 */
  zeta_k2u_k_obl(c, c->zeta, c->uk_dc, u);
  memcpy(c->wv1, u[1], sizeof(double) * c->nfield);
  u_k2u(c, u);
  /* Zeta_k2omega_OBL, operate_omega.h:169-175 */
  zeta_k2omega_k_obl(c, c->zeta, c->wv3);
  u_k2u(c, c->wv3);
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++)
      for (int k = 0; k < c->NZ; k++) {
        const size_t im = IDX(c, i, j, k);
        const double u1 = u[0][im], u2 = u[1][im], u3 = u[2][im];
        const double w1 = c->wv3[0][im], w2 = c->wv3[1][im], w3 = c->wv3[2][im];
        u[0][im] = u2 * w3 - u3 * w2;
        u[1][im] = u3 * w1 - u1 * w3;
        u[2][im] = u1 * w2 - u2 * w1;
      }
  u2u_k(c, u);
  for (size_t im = 0; im < c->nfield; im++) {
    u[0][im] -= 2. * c->shear_rate_eff * c->wv1[im];
    u[1][im] -= 2. * c->shear_rate_eff * c->degree_oblique * c->wv1[im];
  }
  u_k2rotation_k(c, u);
  omega_k2zeta_k_obl(c, u, adv);
}

/* NS_solver_slavedEuler_Shear_OBL, fluid_solver.cxx:398-429 */
int oracle_ns_solver_shear_obl(oracle_ctx *c) {
  if (!c->le) return -1;
  zeta_k2advection_k_obl(c, c->f_ns0);
  const double dmy0 = -c->NU * c->dt;
  const double dmy1 = 1. / c->NU;
  int rng[4][6];
  ijk_blocks(c, rng);
  for (int n = 0; n < 4; n++)
    for (int i = rng[n][0]; i <= rng[n][1]; i++)
      for (int j = rng[n][2]; j <= rng[n][3]; j++)
        for (int k = rng[n][4]; k <= rng[n][5]; k++) {
          const size_t im = IDX(c, i, j, k);
          const double dmy = exp(dmy0 * c->K2[im]) - 1.;
          c->zeta[0][im] += dmy * (c->zeta[0][im] - dmy1 * c->IK2[im] * c->f_ns0[0][im]);
          c->zeta[1][im] += dmy * (c->zeta[1][im] - dmy1 * c->IK2[im] * c->f_ns0[1][im]);
        }
  return 0;
}

/* U_oblique2u, fft_wrapper.h:265-325: oblique contravariant components on
 * the sheared grid -> Cartesian components on the orthogonal grid, plus the
 * mean shear flow */
static void u_oblique2u(oracle_ctx *c, double *const uu[3]) {
  const int NX = c->NX, NY = c->NY, NZ = c->NZ;
  const double deg = c->degree_oblique;
  for (int d = 0; d < DIM; d++) memcpy(c->wv3[d], uu[d], sizeof(double) * c->nfield);
  for (int i = 0; i < NX; i++)
    for (int j = 0; j < NY; j++) {
      double sign = j - NY / 2;
      if (!(sign == 0)) sign = sign / fabs(sign);
      int i_oblique = (int)(sign * deg * (j - NY / 2.)) * sign + sign;
      const double alpha = (i_oblique - deg * (j - NY / 2.)) * sign;
      const double beta = 1. - alpha;
      i_oblique = (int)fmod(i + i_oblique + 4. * NX, NX);
      const int i_plus = (int)fmod(i + sign + 2 * NX, NX);
      for (int k = 0; k < NZ; k++) {
        const size_t im = IDX(c, i, j, k), im_ob = IDX(c, i_oblique, j, k), im_p = IDX(c, i_plus, j, k);
        uu[0][im_ob] = (beta * c->wv3[0][im] + alpha * c->wv3[0][im_p]) +
                       deg * (beta * c->wv3[1][im] + alpha * c->wv3[1][im_p]) + c->shear_rate_eff * (j - NY / 2.);
        uu[1][im_ob] = beta * c->wv3[1][im] + alpha * c->wv3[1][im_p];
        uu[2][im_ob] = beta * c->wv3[2][im] + alpha * c->wv3[2][im_p];
      }
    }
}

/* Calc_shear_rate_eff, sp_3d_ns.h:174-196 (ivolume = Ivolume DX^3).  The
 * reference sums the NX (NY-1) NZ differences in an OpenMP reduction, whose
 * order is not fixed; a plain loop over 10^6 and more terms carries a
 * rounding of up to ~1e-10 relative, so the oracle adds them with Kahan
 * compensation, which is within rounding of the exact sum that every order
 * approximates. */
static void calc_shear_rate_eff(oracle_ctx *c) {
  const double ivolume = (1. / (c->L[0] * c->L[1] * c->L[2])) * POW3(c->DX);
  double s = 0.0, comp = 0.0;
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY - 1; j++)
      for (int k = 0; k < c->NZ; k++) {
        const double y = (c->ucp[0][IDX(c, i, j + 1, k)] - c->ucp[0][IDX(c, i, j, k)]) / c->DX - comp;
        const double t = s + y;
        comp = (t - s) - y;
        s = t;
      }
  s *= ivolume * c->NY / (c->NY - 1);
  c->shear_rate_eff = s;
}

/* Solenoidal_uk_OBL, operate_omega.cxx:222-262 */
static void solenoidal_uk_obl(oracle_ctx *c, double *const u[3]) {
  const double deg = c->degree_oblique;
  for (size_t im = 0; im < c->nfield; im++) {
    const double kx = c->KX_int[im] * c->WAVE_X, ky = c->KY_int[im] * c->WAVE_Y, kz = c->KZ_int[im] * c->WAVE_Z;
    const double kxc = (1. + deg * deg) * c->KX_int[im] * c->WAVE_X - deg * c->KY_int[im] * c->WAVE_Y;
    const double kyc = -deg * c->KX_int[im] * c->WAVE_X + c->KY_int[im] * c->WAVE_Y;
    const double dmy = c->IK2[im] * (u[0][im] * kx + u[1][im] * ky + u[2][im] * kz);
    u[0][im] -= dmy * kxc;
    u[1][im] -= dmy * kyc;
    u[2][im] -= dmy * kz;
  }
}

/* contra2co, fft_wrapper.h:327-361 (every slot) */
static void contra2co_field(oracle_ctx *c, double *const u[3]) {
  const double deg = c->degree_oblique;
  for (size_t im = 0; im < c->nfield; im++) {
    const double w0 = u[0][im], w1 = u[1][im], w2 = u[2][im];
    u[0][im] = w0 + deg * w1;
    u[1][im] = deg * w0 + (1. + deg * deg) * w1;
    u[2][im] = w2;
  }
}

/* U_k2zeta_k_OBL, fft_wrapper.cxx:341-406 */
static void u_k2zeta_k_obl(oracle_ctx *c, double *const u[3], double *const zeta[2], double uk_dc[3]) {
  uk_dc[0] = u[0][0];
  uk_dc[1] = u[1][0];
  uk_dc[2] = u[2][0];
  co2contra_single(c->degree_oblique, uk_dc);
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++)
      for (int k = 0; k < c->HNZ_; k++) {
        const size_t im = IDX(c, i, j, 2 * k);
        const double ks[3] = {c->KX_int[im] * c->WAVE_X, c->KY_int[im] * c->WAVE_Y, c->KZ_int[im] * c->WAVE_Z};
        double re[3], iv[3];
        re[0] = (ks[1] * u[2][im + 1] - ks[2] * u[1][im + 1]);
        iv[0] = -(ks[1] * u[2][im] - ks[2] * u[1][im]);
        re[1] = (ks[2] * u[0][im + 1] - ks[0] * u[2][im + 1]);
        iv[1] = -(ks[2] * u[0][im] - ks[0] * u[2][im]);
        re[2] = (ks[0] * u[1][im + 1] - ks[1] * u[0][im + 1]);
        iv[2] = -(ks[0] * u[1][im] - ks[1] * u[0][im]);
        int a, b;
        if (c->KX_int[im] != 0) {
          a = 1; b = 2;
        } else if (c->KZ_int[im] != 0) {
          a = 1; b = 0;
        } else {
          a = 0; b = 2;
        }
        zeta[0][im] = re[a];
        zeta[0][im + 1] = iv[a];
        zeta[1][im] = re[b];
        zeta[1][im + 1] = iv[b];
      }
}

/* ---- spherical particles under Lees-Edwards shear ---- */

/* Relative_coord_check_stepover_Y (make_phi.h:197-236): the stencil point
 * `cell` of a particle anchored at x_int + residue; a point across the y
 * boundary belongs to the particle's image shifted by -sign degree_oblique
 * NY cells in x.  Returns sign (+1 / -1 across the upper / lower boundary,
 * else 0). */
static int relative_coord_obl(const oracle_ctx *c, const int *cell, const int *x_int, const double *residue,
                              int *r_mesh, double *r) {
  const int Ns[3] = {c->NX, c->NY, c->NZ};
  const double dx = c->DX;
  double signY = x_int[1] + cell[1];
  r_mesh[1] = (x_int[1] + cell[1] + Ns[1]) % Ns[1];
  signY -= r_mesh[1];
  int sign = (int)signY;
  if (!(sign == 0)) sign = sign / abs(sign);
  const double xo = fmod(x_int[0] - sign * c->degree_oblique * Ns[1] + residue[0] / dx + cell[0] + 4 * Ns[0], Ns[0]);
  r_mesh[0] = (int)xo;
  const double x_oblique_residue = xo - r_mesh[0];
  r_mesh[2] = (x_int[2] + cell[2] + Ns[2]) % Ns[2];
  r[0] = (double)cell[0] * dx - x_oblique_residue * dx;
  r[1] = (double)cell[1] * dx - residue[1];
  r[2] = (double)cell[2] * dx - residue[2];
  return sign;
}

/* Make_phi_u_primitive_OBL (make_phi.cxx:299-385) with SW_UP, after
 * Reset_phi_u: the x velocity of a point across the y boundary is taken in
 * the frame of the particle's image, v_x - sign Shear_rate_eff LY. */
static void make_phi_u_obl(oracle_ctx *c) {
  reset_phi_u(c);
  const double ly = c->L[1];
  for (int n = 0; n < c->Np; n++) {
    const spm_particle *pn = &c->p[n];
    int x_int[3], r_mesh[3];
    double residue[3], r[3], v_rot[3];
    particle_cell(c, pn->x, x_int, residue);
    for (int mesh = 0; mesh < c->np_domain; mesh++) {
      const int sign = relative_coord_obl(c, &c->sekibun[3 * mesh], x_int, residue, r_mesh, r);
      double dmy = 0.;
      for (int d = 0; d < DIM; d++) dmy += SQ(r[d]);
      dmy = sqrt(dmy);
      const double dmy_phi = Phi(c, dmy, c->RADIUS);
      const size_t im = IDX(c, r_mesh[0], r_mesh[1], r_mesh[2]);
      c->phi[im] += dmy_phi;
      angular2v(c, pn->omega, r, v_rot);
      c->up[0][im] += ((pn->v[0] - sign * c->shear_rate_eff * ly + v_rot[0]) * dmy_phi);
      c->up[1][im] += (pn->v[1] + v_rot[1]) * dmy_phi;
      c->up[2][im] += (pn->v[2] + v_rot[2]) * dmy_phi;
    }
  }
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++)
      for (int k = 0; k < c->NZ; k++) {
        const size_t im = IDX(c, i, j, k);
        if (c->phi[im] > 1.) {
          const double idmy_phi = 1. / c->phi[im];
          c->up[0][im] *= idmy_phi;
          c->up[1][im] *= idmy_phi;
          c->up[2][im] *= idmy_phi;
          c->phi[im] = 1.;
        }
      }
}

/* Calc_f_hydro_correct_precision_OBL (md_force.cxx:391-620) on the
 * Cartesian velocity `fld` (no rigid bodies).  The Hydro_force and
 * Hydro_force_new fields feed only Calc_hydro_stress (fluid_solver.h:
 * 195-221), which sums them over the grid, so their sums are kept, added
 * point by point in the reference's three particle loops. */
static void calc_f_hydro_obl(oracle_ctx *c, double *const fld[3]) {
  const double dmy0 = -c->DX3 * c->RHO;
  const double dmy = dmy0 / c->dt;
  const double ly = c->L[1];
  double *volume = (double *)calloc((size_t)(c->Np > 0 ? c->Np : 1), sizeof(double));
  double *itrace = (double *)calloc((size_t)(c->Np > 0 ? c->Np : 1), sizeof(double));
  double hf = 0.0, hfn = 0.0, sum_force = 0.0, sum_volume = 0.0;
  for (int n = 0; n < c->Np; n++) {
    spm_particle *pn = &c->p[n];
    const double rhop = c->P.mass_ratio[pn->spec] * c->RHO; /* RHO_particle, input.cxx:279 */
    int x_int[3], r_mesh[3];
    double residue[3], r[3], v_rot[3], dmy_fp[3];
    double force[3] = {0, 0, 0}, torque[3] = {0, 0, 0};
    particle_cell(c, pn->x, x_int, residue);
    for (int mesh = 0; mesh < c->np_domain; mesh++) {
      const int sign = relative_coord_obl(c, &c->sekibun[3 * mesh], x_int, residue, r_mesh, r);
      double dmyR = 0.;
      for (int d = 0; d < DIM; d++) dmyR += SQ(r[d]);
      dmyR = sqrt(dmyR);
      const double dmy_phi = Phi(c, dmyR, c->RADIUS);
      angular2v(c, pn->omega, r, v_rot);
      const size_t im = IDX(c, r_mesh[0], r_mesh[1], r_mesh[2]);
      dmy_fp[0] = ((pn->v[0] - sign * c->shear_rate_eff * ly + v_rot[0]) - fld[0][im]) * dmy_phi;
      dmy_fp[1] = ((pn->v[1] + v_rot[1]) - fld[1][im]) * dmy_phi;
      dmy_fp[2] = ((pn->v[2] + v_rot[2]) - fld[2][im]) * dmy_phi;
      for (int d = 0; d < DIM; d++) force[d] += dmy_fp[d];
      torque[0] += (r[1] * dmy_fp[2] - r[2] * dmy_fp[1]);
      torque[1] += (r[2] * dmy_fp[0] - r[0] * dmy_fp[2]);
      torque[2] += (r[0] * dmy_fp[1] - r[1] * dmy_fp[0]);
      const double dmy_ry = (r_mesh[1] + sign * ly); /* md_force.cxx:498 */
      hf += dmy_fp[0] * dmy_ry * rhop;
      hfn += dmy_fp[0] * dmy_ry * rhop;
      volume[n] += dmy_phi;
      itrace[n] += dmy_phi * SQ(dmyR);
    }
    itrace[n] *= 2.0 / 3.0;
    for (int d = 0; d < DIM; d++) pn->f_hydro[d] = (dmy * force[d]);
    if (c->P.rotation)
      for (int d = 0; d < DIM; d++) pn->torque_hydro[d] = (dmy * torque[d]);
    /* md_force.cxx:527-555 */
    for (int mesh = 0; mesh < c->np_domain; mesh++) {
      const int sign = relative_coord_obl(c, &c->sekibun[3 * mesh], x_int, residue, r_mesh, r);
      double dmyR = 0.;
      for (int d = 0; d < DIM; d++) dmyR += SQ(r[d]);
      dmyR = sqrt(dmyR);
      const double dmy_phi = Phi(c, dmyR, c->RADIUS);
      const double dmy_ry = (r_mesh[1] + sign * ly);
      const double dmy_stress_p = pn->momentum_depend_fr[0] / volume[n];
      hf -= dmy_stress_p * dmy_ry * dmy_phi;
      const double dmy_stress_v = -c->RHO * force[0] / volume[n];
      const double dmy_stress_w = -c->RHO * (torque[1] * r[2] - torque[2] * r[1]) / itrace[n];
      hfn += (dmy_stress_v + dmy_stress_w) * dmy_ry * dmy_phi;
    }
    sum_force += force[0];
    sum_volume += volume[n];
  }
  sum_volume /= c->RHO;
  /* md_force.cxx:580-620 */
  for (int n = 0; n < c->Np; n++) {
    int x_int[3], r_mesh[3];
    double residue[3], r[3];
    particle_cell(c, c->p[n].x, x_int, residue);
    for (int mesh = 0; mesh < c->np_domain; mesh++) {
      const int sign = relative_coord_obl(c, &c->sekibun[3 * mesh], x_int, residue, r_mesh, r);
      double dmyR = 0.;
      for (int d = 0; d < DIM; d++) dmyR += SQ(r[d]);
      dmyR = sqrt(dmyR);
      const double dmy_phi = Phi(c, dmyR, c->RADIUS);
      const double dmy_ry = (r_mesh[1] + sign * ly);
      hf -= sum_force * dmy_ry * dmy_phi / sum_volume;
    }
  }
  c->hydro_force_sum = hf;
  c->hydro_force_new_sum = hfn;
  free(volume);
  free(itrace);
}

/* PBC_OBL (periodic_boundary.h:22-44): a particle leaving through a y
 * boundary re-enters at the other one shifted by -sign degree_oblique LY
 * in x; returns sign and the x velocity change -sign Shear_rate_eff LY. */
static int pbc_obl(const oracle_ctx *c, double *x, double *delta_vx) {
  double signY = x[1];
  x[1] = fmod(x[1] + c->L[1], c->L[1]);
  signY -= x[1];
  int sign = (int)signY;
  if (!(sign == 0)) sign = sign / abs(sign);
  x[0] -= (double)sign * c->degree_oblique * c->L[1];
  x[0] = fmod(x[0] + c->L[0], c->L[0]);
  x[2] = fmod(x[2] + c->L[2], c->L[2]);
  *delta_vx = -(double)sign * c->shear_rate_eff * c->L[1];
  return sign;
}

/* MD_solver_position_Euler_OBL / _AB2_OBL (particle_solver.cxx:292-346). */
int oracle_md_solver_position_obl(oracle_ctx *c) {
  if (!c || !c->le) return -1;
  for (int n = 0; n < c->Np; n++) {
    spm_particle *p = &c->p[n];
    for (int d = 0; d < DIM; d++) {
      p->x_previous[d] = p->x[d];
      const double delta_x = (c->ts == 0) ? c->dt * p->v[d] : c->hdt * (3.0 * p->v[d] - p->v_old[d]);
      p->x_nopbc[d] += delta_x;
      p->x[d] += delta_x;
    }
    double delta_vx;
    pbc_obl(c, p->x, &delta_vx);
    p->v[0] += delta_vx;
    p->v_old[0] += delta_vx;
    if (c->P.rotation) {
      if (c->ts == 0)
        orientation_euler(p, c->dt);
      else
        orientation_ab2(p, c->hdt);
    }
  }
  return 0;
}

/* Distance0_OBL (interaction.h:39-65): the y separation to the nearest
 * image first, then x measured from the image of x1 shifted by
 * sign degree_oblique LY. */
static double distance0_obl(const oracle_ctx *c, const double *x1, const double *x2, double *x12) {
  double dmy = 0.0;
  double signY = x2[1] - x1[1];
  x12[1] = x2[1] - x1[1];
  x12[1] -= (double)Nint(x12[1] / c->L[1]) * c->L[1];
  signY -= x12[1];
  int sign = (int)signY;
  if (!(sign == 0)) sign = sign / abs(sign);
  dmy += SQ(x12[1]);
  x12[0] = x2[0] - (x1[0] + (double)sign * c->degree_oblique * c->L[1]);
  x12[0] -= (double)Nint(x12[0] / c->L[0]) * c->L[0];
  dmy += SQ(x12[0]);
  x12[2] = x2[2] - x1[2];
  x12[2] -= (double)Nint(x12[2] / c->L[2]) * c->L[2];
  dmy += SQ(x12[2]);
  return sqrt(dmy);
}

double oracle_distance0_obl(oracle_ctx *c, const double *x1, const double *x2, double *x12) {
  return distance0_obl(c, x1, x2, x12);
}

/* Force_OBL (particle_solver.h:111-127): Calc_f_Lennard_Jones_OBL, the
 * all-pairs loop of md_force.cxx:136-179 with Distance0_OBL, and
 * Add_f_gravity. */
int oracle_force_obl(oracle_ctx *c) {
  if (!c || !c->le) return -1;
  c->lj_shear_stress = 0.0; /* dev_shear_stress_lj = 0, particle_solver.h:113 */
  if (c->P.lj_truncate >= 0)
    for (int n = 0; n < c->Np; n++)
      for (int m = n + 1; m < c->Np; m++) {
        double r_ij_vec[3];
        const double r_ij = distance0_obl(c, c->p[n].x, c->p[m].x, r_ij_vec);
        if (r_ij < c->A_R_cutoff * c->LJ_dia) {
          const double dmy = DMIN(DBL_MAX / r_ij, lennard_jones_f(c, r_ij, c->LJ_dia));
          for (int d = 0; d < DIM; d++) {
            const double dmy1 = dmy * -r_ij_vec[d];
            c->p[n].fr[d] += dmy1;
            c->p[m].fr[d] -= dmy1;
          }
          c->lj_shear_stress += ((dmy * -r_ij_vec[0]) * (r_ij_vec[1])); /* md_force.cxx:174 */
        }
      }
  if (c->P.g != 0.0) {
    const double Gravity_on_fluid = c->P.g * c->RHO * 4. / 3. * M_PI * SQ(c->RADIUS) * c->RADIUS;
    for (int n = 0; n < c->Np; n++)
      c->p[n].fr[c->P.g_direction] -= Gravity_on_fluid * (c->P.mass_ratio[c->p[n].spec] - 1.0);
  }
  return 0;
}

/* MD_solver_velocity_Euler_OBL / _AB2_hydro_OBL (particle_solver.cxx:
 * 348-443): the Navier-Stokes update after Force_OBL. */
int oracle_md_solver_velocity_obl(oracle_ctx *c) {
  if (!c || !c->le) return -1;
  oracle_force_obl(c);
  /* the impulse of fr the stress output reads (particle_solver.cxx:373,422) */
  for (int n = 0; n < c->Np; n++)
    for (int d = 0; d < DIM; d++)
      c->p[n].momentum_depend_fr[d] =
          (c->ts == 0) ? c->dt * c->p[n].fr[d] : c->hdt * (c->p[n].fr[d] + c->p[n].fr_previous[d]);
  velocity_update(c);
  return 0;
}

/* Mean_shear_stress(SHOW, ...) for Lees-Edwards DC (sp_3d_ns.h:209-358):
 * time, Shear_rate, Shear_strain_realized, dev_shear_stress_lj,
 * hydro_stress[1][0] (Hydro_force), hydro_stress_new[1][0]
 * (Hydro_force_new), viscosity (Calc_hydro_stress, fluid_solver.h:195-221:
 * stress = -sum * Ivolume DX^3 / dt; Force_OBL's dev stress times
 * Ivolume). */
int oracle_le_shear_stress(oracle_ctx *c, double out[7]) {
  if (!c || !c->le) return -1;
  const double Ivolume = 1. / (c->L[0] * c->L[1] * c->L[2]);
  const double dmy = Ivolume * c->DX3 / c->dt;
  const double dev = c->lj_shear_stress * Ivolume;
  out[0] = c->time;
  out[1] = c->shear_rate;
  out[2] = c->strain;
  out[3] = dev;
  out[4] = -c->hydro_force_sum * dmy;
  out[5] = -c->hydro_force_new_sum * dmy;
  out[6] = (out[5] + dev) / c->shear_rate + c->P.eta;
  return 0;
}

/* U2u_oblique (fft_wrapper.h:204-263): Cartesian components on the
 * orthogonal grid -> oblique contravariant components on the sheared grid,
 * by linear interpolation along x. */
static void u2u_oblique(oracle_ctx *c, double *const uu[3]) {
  const int NX = c->NX, NY = c->NY, NZ = c->NZ;
  const double deg = c->degree_oblique;
  for (int d = 0; d < DIM; d++) memcpy(c->wv3[d], uu[d], sizeof(double) * c->nfield);
  for (int i = 0; i < NX; i++)
    for (int j = 0; j < NY; j++) {
      double sign = j - NY / 2;
      if (!(sign == 0)) sign = sign / fabs(sign);
      int i_oblique = (int)(sign * deg * (j - NY / 2)) * sign;
      const double alpha = (deg * (j - NY / 2) - i_oblique) * sign;
      const double beta = 1. - alpha;
      i_oblique = (int)fmod(i + i_oblique + 4. * NX, NX);
      const int i_oblique_plus = (int)fmod(i_oblique + sign + 4. * NX, NX);
      for (int k = 0; k < NZ; k++) {
        const size_t im = IDX(c, i, j, k), im_ob = IDX(c, i_oblique, j, k), im_ob_p = IDX(c, i_oblique_plus, j, k);
        uu[0][im] = (beta * c->wv3[0][im_ob] + alpha * c->wv3[0][im_ob_p]) -
                    deg * (beta * c->wv3[1][im_ob] + alpha * c->wv3[1][im_ob_p]);
        uu[1][im] = beta * c->wv3[1][im_ob] + alpha * c->wv3[1][im_ob_p];
        uu[2][im] = beta * c->wv3[2][im_ob] + alpha * c->wv3[2][im_ob_p];
      }
    }
}

/* hooks: the Cartesian velocity is taken from u0..u2 */
int oracle_make_phi_u_particle_obl(oracle_ctx *c) {
  if (!c || !c->le) return -1;
  make_phi_u_obl(c);
  return 0;
}
int oracle_calc_f_hydro_obl(oracle_ctx *c) {
  if (!c || !c->le) return -1;
  calc_f_hydro_obl(c, c->u);
  return 0;
}
int oracle_u2u_oblique(oracle_ctx *c) {
  if (!c || !c->le) return -1;
  u2u_oblique(c, c->u);
  return 0;
}
int oracle_u_oblique2u(oracle_ctx *c) {
  if (!c || !c->le) return -1;
  u_oblique2u(c, c->u);
  return 0;
}

/* The integer strain (sp_3d_ns.cxx:280-305): once degree_oblique reaches 1
 * the grid is remapped, u0 += u1 with the rows shifted by j - NY/2, the
 * copy ucp refreshed and degree_oblique reduced by 1. */
static void remap_obl(oracle_ctx *c) {
  const int NX = c->NX, NY = c->NY, NZ = c->NZ;
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++) {
        double sign = j - NY / 2;
        if (!(sign == 0)) sign = sign / fabs(sign);
        int i_oblique = (int)(sign * (j - NY / 2)) * sign;
        i_oblique = (int)fmod(i + i_oblique + 4. * NX, NX);
        for (int k = 0; k < NZ; k++) {
          const size_t im = IDX(c, i, j, k), im_obl = IDX(c, i_oblique, j, k);
          c->u[0][im_obl] = c->ucp[0][im] + c->ucp[1][im];
          c->u[1][im_obl] = c->ucp[1][im];
          c->u[2][im_obl] = c->ucp[2][im];
        }
      }
    c->degree_oblique -= 1.;
    for (int d = 0; d < DIM; d++) memcpy(c->ucp[d], c->u[d], sizeof(double) * c->nfield);
}

int oracle_remap_obl(oracle_ctx *c) {
  if (!c || !c->le) return -1;
  for (int d = 0; d < DIM; d++) memcpy(c->ucp[d], c->u[d], sizeof(double) * c->nfield);
  remap_obl(c);
  return 0;
}

/* Time_evolution_hydro_OBL (sp_3d_ns.cxx:214-390).  Without particles f is
 * zero and the particle block leaves u as it is. */
static void le_step(oracle_ctx *c) {
  oracle_ns_solver_shear_obl(c);
  for (int d = 0; d < DIM; d++)
    if (c->P.fix_cell[d]) c->uk_dc[d] = 0.0;
  zeta_k2u_k_obl(c, c->zeta, c->uk_dc, c->u);
  u_k2u(c, c->u);
  c->shear_rate_eff = c->shear_rate;
  for (int d = 0; d < DIM; d++) memcpy(c->ucp[d], c->u[d], sizeof(double) * c->nfield);
  c->degree_oblique += c->shear_rate_eff * c->dt;
  if (c->degree_oblique >= 1.) {
    remap_obl(c);
  }
  u_oblique2u(c, c->ucp);
  obl_update_k2(c);
  calc_shear_rate_eff(c);
  if (c->Np > 0) {
    oracle_md_solver_position_obl(c);
    reset_phi_u(c);
    calc_f_hydro_obl(c, c->ucp);
    oracle_md_solver_velocity_obl(c);
    if (c->kBT > 0.) oracle_add_random_force_thermostat(c);
    make_phi_u_obl(c);
    /* Make_f_particle_dt_nonsole (f_particle.h:41-62), U2u_oblique,
     * Add_f_particle */
    for (size_t im = 0; im < c->nfield; im++)
      for (int d = 0; d < DIM; d++) c->f[d][im] = c->up[d][im] - c->phi[im] * c->ucp[d][im];
    u2u_oblique(c, c->f);
    for (int i = 0; i < c->NX; i++)
      for (int j = 0; j < c->NY; j++)
        for (int k = 0; k < c->NZ; k++) {
          const size_t im = IDX(c, i, j, k);
          for (int d = 0; d < DIM; d++) c->u[d][im] += c->f[d][im];
        }
  }
  u2u_k(c, c->u);
  solenoidal_uk_obl(c, c->u);
  contra2co_field(c, c->u);
  u_k2zeta_k_obl(c, c->u, c->zeta, c->uk_dc);
  c->strain += c->shear_rate_eff * c->dt; /* sp_3d_ns.cxx:600 */
}

int oracle_set_lees_edwards(oracle_ctx *c, double shear_rate) {
  if (!c || c->le || c->electro) return -1;
  for (int d = 0; d < DIM; d++) {
    c->ucp[d] = (double *)calloc(c->nfield, sizeof(double));
    c->wv3[d] = (double *)calloc(c->nfield, sizeof(double));
    if (!c->ucp[d] || !c->wv3[d]) return -1;
  }
  c->wv1 = (double *)calloc(c->nfield, sizeof(double));
  if (!c->wv1) return -1;
  /* LJ_dia under shear, input.cxx:389-403 */
  {
    const double dummy_pow = c->P.lj_powers == 0 ? pow(2., 1. / 6.) : c->P.lj_powers == 1 ? pow(2., 1. / 12.)
                                                                                          : pow(2., 1. / 18.);
    c->LJ_dia = DMIN((c->SIGMA + c->XI) / dummy_pow, c->SIGMA);
  }
  c->le = 1;
  c->shear_rate = shear_rate;
  c->shear_rate_eff = 0.0;
  c->degree_oblique = 0.0;
  return 0;
}

int oracle_put_oblique(oracle_ctx *c, double degree_oblique, double shear_rate_eff) {
  if (!c || !c->le) return -1;
  c->degree_oblique = degree_oblique;
  c->shear_rate_eff = shear_rate_eff;
  obl_update_k2(c);
  return 0;
}

int oracle_get_oblique(oracle_ctx *c, double *degree_oblique, double *shear_rate_eff) {
  if (!c || !c->le) return -1;
  *degree_oblique = c->degree_oblique;
  *shear_rate_eff = c->shear_rate_eff;
  return 0;
}

/* the contravariant velocity spectrum of zeta (Zeta_k2u_k_OBL) into u */
int oracle_zeta_k2u_k_obl(oracle_ctx *c) {
  if (!c || !c->le) return -1;
  zeta_k2u_k_obl(c, c->zeta, c->uk_dc, c->u);
  return 0;
}

/* Solenoidal_uk_OBL, contra2co and U_k2zeta_k_OBL on the spectrum in u */
int oracle_u_k2zeta_k_obl(oracle_ctx *c) {
  if (!c || !c->le) return -1;
  solenoidal_uk_obl(c, c->u);
  contra2co_field(c, c->u);
  u_k2zeta_k_obl(c, c->u, c->zeta, c->uk_dc);
  return 0;
}

/* ------------------------------------------------------------------ */
/* lifetime and transfers                                              */
/* ------------------------------------------------------------------ */
/* powers of two from 8, and the mixed lengths 48, 80, 96, 160, 192 that the
 * GPU path adds (the reference's Ooura transform takes powers of two only) */
static int size_ok(int n) {
  return (n >= 8 && n <= 1024 && (n & (n - 1)) == 0) || n == 48 || n == 80 || n == 96 || n == 160 || n == 192;
}

oracle_ctx *oracle_create(const spm_params *prm) {
  if (!prm || !size_ok(prm->nx) || !size_ok(prm->ny) || !size_ok(prm->nz) || prm->n_particles < 0 ||
      prm->n_species < 1 || prm->n_species > SPM_MAX_SPECIES || !(prm->dt > 0))
    return NULL;
  oracle_ctx *c = (oracle_ctx *)calloc(1, sizeof(oracle_ctx));
  if (!c) return NULL;
  c->P = *prm;
  c->NX = prm->nx;
  c->NY = prm->ny;
  c->NZ = prm->nz;
  c->HNZ_ = c->NZ / 2 + 1;
  c->NZ_ = 2 * c->HNZ_;
  c->TRN_X = (c->NX + 2) / 3;
  c->TRN_Y = (c->NY + 2) / 3;
  c->TRN_Z = (c->NZ + 2) / 3;
  c->nfield = (size_t)c->NX * c->NY * c->NZ_;
  /* Set_global_parameters, input.cxx:274-433 */
  c->DX = prm->dx;
  c->RHO = prm->rho;
  c->ETA = prm->eta;
  c->kBT = prm->kbt;
  c->IRHO = 1. / c->RHO;
  c->NU = c->ETA * c->IRHO;
  c->DX3 = c->DX * c->DX * c->DX;
  c->RADIUS = prm->a * c->DX;
  c->XI = prm->a_xi * c->DX;
  c->HXI = c->XI * 0.5;
  c->SIGMA = 2.0 * c->RADIUS;
  const int Ns[3] = {c->NX, c->NY, c->NZ};
  for (int d = 0; d < DIM; d++) c->L[d] = Ns[d] * c->DX;
  const double PI2 = 2.0 * M_PI;
  c->WAVE_X = PI2 / c->L[0];
  c->WAVE_Y = PI2 / c->L[1];
  c->WAVE_Z = PI2 / c->L[2];
  for (int s = 0; s < prm->n_species; s++) {
    double rho_p = prm->mass_ratio[s] * c->RHO;
    c->MASS[s] = rho_p * 4. / 3. * M_PI * POW3(c->RADIUS);
    c->IMASS[s] = 1. / c->MASS[s];
    c->MOI[s] = 2. / 5. * c->MASS[s] * SQ(c->RADIUS);
    c->IMOI[s] = 1. / c->MOI[s];
  }
  c->LJ_dia = c->SIGMA;
  if (prm->lj_truncate > 0) {
    c->A_R_cutoff = prm->lj_powers == 0 ? pow(2., 1. / 6.) : prm->lj_powers == 1 ? pow(2., 1. / 12.) : pow(2., 1. / 18.);
  } else if (prm->lj_truncate == 0) {
    int nmax = c->NX > c->NY ? c->NX : c->NY;
    nmax = nmax > c->NZ ? nmax : c->NZ;
    c->A_R_cutoff = DMIN(nmax * c->DX * .5 / c->SIGMA, 2.5);
  } else {
    c->A_R_cutoff = 0.;
  }
  c->dt = prm->dt;
  c->hdt = prm->dt * 0.5;
  size_t nb = sizeof(double) * c->nfield;
  int fail = 0;
  c->KX_int = (int *)malloc(sizeof(int) * c->nfield);
  c->KY_int = (int *)malloc(sizeof(int) * c->nfield);
  c->KZ_int = (int *)malloc(sizeof(int) * c->nfield);
  c->K2 = (double *)malloc(nb);
  c->IK2 = (double *)malloc(nb);
  fail |= !c->KX_int || !c->KY_int || !c->KZ_int || !c->K2 || !c->IK2;
  for (int d = 0; d < 2; d++) {
    c->zeta[d] = (double *)calloc(c->nfield, sizeof(double));
    c->f_ns0[d] = (double *)calloc(c->nfield, sizeof(double));
    fail |= !c->zeta[d] || !c->f_ns0[d];
  }
  for (int d = 0; d < 3; d++) {
    c->u[d] = (double *)calloc(c->nfield, sizeof(double));
    c->up[d] = (double *)calloc(c->nfield, sizeof(double));
    c->f[d] = (double *)calloc(c->nfield, sizeof(double));
    fail |= !c->u[d] || !c->up[d] || !c->f[d];
  }
  c->phi = (double *)calloc(c->nfield, sizeof(double));
  c->ztw = (double *)malloc(sizeof(double) * (c->NZ + 2));
  fail |= !c->phi || !c->ztw;
  if (c->ztw)
    for (int k = 0; k <= c->NZ / 2; k++) {
      long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)c->NZ;
      c->ztw[2 * k] = (double)cosl(ang);
      c->ztw[2 * k + 1] = (double)sinl(ang);
    }
  fail |= plan_init(&c->px, c->NX) | plan_init(&c->py, c->NY) | plan_init(&c->pz, c->NZ / 2);
  fail |= particle_domain(c);
  c->Np = prm->n_particles;
  c->p = (spm_particle *)calloc(c->Np > 0 ? c->Np : 1, sizeof(spm_particle));
  fail |= !c->p;
  if (fail) {
    oracle_destroy(c);
    return NULL;
  }
  for (int n = 0; n < c->Np; n++) c->p[n].q.s = 1.0;
  init_k(c);
  init_genrand(c, prm->seed);
  return c;
}

void oracle_destroy(oracle_ctx *c) {
  if (!c) return;
  for (int d = 0; d < 3; d++) {
    free(c->ucp[d]);
    free(c->wv3[d]);
  }
  free(c->wv1);
  free(c->KX_int);
  free(c->KY_int);
  free(c->KZ_int);
  free(c->K2);
  free(c->IK2);
  for (int d = 0; d < 2; d++) {
    free(c->zeta[d]);
    free(c->f_ns0[d]);
  }
  for (int d = 0; d < 3; d++) {
    free(c->u[d]);
    free(c->up[d]);
    free(c->f[d]);
  }
  free(c->phi);
  for (int n = 0; n < 2; n++) {
    free(c->conc[n]);
    free(c->conc_rhs[n]);
    free(c->f_ns1[n]);
  }
  for (int d = 0; d < 3; d++) free(c->surface_normal[d]);
  free(c->ztw);
  plan_free(&c->px);
  plan_free(&c->py);
  plan_free(&c->pz);
  free(c->sekibun);
  free(c->p);
  free(c);
}

int oracle_stencil_size(const oracle_ctx *c) { return c->np_domain; }
int oracle_stencil(const oracle_ctx *c, int *out) {
  memcpy(out, c->sekibun, sizeof(int) * 3 * c->np_domain);
  return c->np_domain;
}

static double *field_ptr(oracle_ctx *c, int f) {
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
    default: return NULL;
  }
}

int oracle_put_field(oracle_ctx *c, int field, const double *h) {
  double *p = field_ptr(c, field);
  if (!p || !h) return SPM_ERR_INVALID;
  memcpy(p, h, sizeof(double) * c->nfield);
  return 0;
}
int oracle_get_field(oracle_ctx *c, int field, double *h) {
  double *p = field_ptr(c, field);
  if (!p || !h) return SPM_ERR_INVALID;
  memcpy(h, p, sizeof(double) * c->nfield);
  return 0;
}
int oracle_put_particles(oracle_ctx *c, const spm_particle *p, int n) {
  if (n != c->Np) return SPM_ERR_INVALID;
  memcpy(c->p, p, sizeof(spm_particle) * n);
  return 0;
}
int oracle_get_particles(oracle_ctx *c, spm_particle *p, int n) {
  if (n != c->Np) return SPM_ERR_INVALID;
  memcpy(p, c->p, sizeof(spm_particle) * n);
  return 0;
}
int oracle_put_uk_dc(oracle_ctx *c, const double u[3]) {
  memcpy(c->uk_dc, u, sizeof(double) * 3);
  return 0;
}
int oracle_get_uk_dc(oracle_ctx *c, double u[3]) {
  memcpy(u, c->uk_dc, sizeof(double) * 3);
  return 0;
}
int oracle_put_time(oracle_ctx *c, int ts, double time) {
  c->ts = ts;
  c->time = time;
  return 0;
}
int oracle_get_time(oracle_ctx *c, int *ts, double *time) {
  *ts = c->ts;
  *time = c->time;
  return 0;
}
int oracle_put_rng(oracle_ctx *c, const uint32_t mt[624], int mti) {
  memcpy(c->mt, mt, sizeof(uint32_t) * MT_N);
  c->mti = mti;
  return 0;
}
int oracle_get_rng(oracle_ctx *c, uint32_t mt[624], int *mti) {
  memcpy(mt, c->mt, sizeof(uint32_t) * MT_N);
  *mti = c->mti;
  return 0;
}
