// fused.hip -- the SPM step as fused FFT passes (FP64).
//
// The literal schedule (spm_api.cpp: one kernel per reference routine)
// moves about 187 field-sized arrays through HBM per step.  Here every
// pointwise operation rides on the FFT pass that already holds the data:
//
//   NS_solver_slavedEuler (fluid_solver.cxx:328-354)
//     x_zeta2u   : x pass (inverse) with Zeta_k2u_k + 2/3 mask as prologue
//     y pass x3  : inverse
//     z_advect   : z pass (inverse) x3 -> products of U2advection_k
//                  (operate_omega.cxx:33-49) -> z pass (forward) x5
//     y pass x5  : forward
//     x_update   : x pass (forward) x5 with the contraction + Omega_k2zeta_k
//                  (operate_omega.cxx:60-98) and the slaved-Euler update
//                  (fluid_solver.cxx:337-353) as epilogue
//   Zeta_k2u (sp_3d_ns.cxx:67-79): x_zeta2u (+FIX_CELL), y x3, z x3
//   Make_phi_u_particle + Make_f_particle_dt_nonsole (make_phi.cxx:221-298,
//   f_particle.h:41-64)
//     make_phi (mode f, particles.hip): phi and up evaluated per grid point
//                  over 8x8x8 tiles, f = up - phi u written directly
//     z pass x3, y pass x3: forward
//   Solenoidal_u + Add_f_particle + U2zeta_k (f_particle.h:71-99,
//   operate_omega.h:186-189)
//     x_project  : x pass (forward) x3 with 2/3 mask + projection of f, plus
//                  u_k = Zeta_k2u_k(zeta) and zeta = U_k2zeta_k(u_k + P f).
//
// The last kernel uses R2C(C2R(u_k)) = u_k for the truncated, Hermitian
// spectrum u_k: the reference transforms u + f back to k space, which is
// the same field up to rounding.  About 80 field-passes per step remain.
#include "fft_core.h"

namespace spm {

// Minimum waves per SIMD the compiler must fit (a VGPR cap: 4 waves of
// 512-thread blocks is two blocks per CU at 128 VGPRs), FP64 x passes.
#ifndef SPM_X_ZETA2U_WAVES_F64
#define SPM_X_ZETA2U_WAVES_F64 4
#endif
#ifndef SPM_X_UPDATE_WAVES_F64
#define SPM_X_UPDATE_WAVES_F64 1
#endif
#ifndef SPM_X_PROJECT_WAVES_F64
#define SPM_X_PROJECT_WAVES_F64 1
#endif
#ifndef SPM_ZADV_WAVES
#define SPM_ZADV_WAVES 1  // minimum waves per SIMD of z_advect_io (a VGPR cap)
#endif

namespace {

template <typename T>
__device__ __forceinline__ T sq(T x) { return x * x; }
__device__ __forceinline__ int nint(double a) { return a >= 0.0 ? (int)(a + .5) : (int)(a - .5); }

struct ModeK {
  int kx, ky, kz;
  bool keep;
};

__device__ __forceinline__ ModeK mode_k(const Grid &g, int i, int j, int kz) {
  ModeK m;
  m.kx = i > g.nx / 2 ? i - g.nx : i;
  m.ky = j > g.ny / 2 ? j - g.ny : j;
  m.kz = kz;
  m.keep = abs(m.kx) < g.trn_x && abs(m.ky) < g.trn_y && m.kz < g.trn_z;
  return m;
}

// Column (j, kz) of an x pass holds kept modes (2/3 rule in ky and kz).
__device__ __forceinline__ bool col_kept(const Grid &g, int j, int kz) {
  return (j < g.ny - j ? j : g.ny - j) < g.trn_y && kz < g.trn_z;
}

// Whether any of the C flat columns [c0, c0 + C) is kept; the same answer
// in every thread, so a block may return on it before its first barrier.
template <int C>
__device__ __forceinline__ bool tile_kept(const Grid &g, int c0) {
  const int H = g.hp, ncol = g.ny * H;
  bool any = false;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int col = c0 + c;
    if (col < ncol) any = any || col_kept(g, col / H, col - (col / H) * H);
  }
  return any;
}

// Tile sequence of a block of the x passes: with one block per tile
// (gridDim = ntiles) step 0 is the block's own tile; with fewer blocks
// (persistent) block b takes tiles b, b + G, b + 2G, ... or, with the XCD
// grouping on and G a multiple of 8, the y-th block of XCD x takes tiles
// lo_x + y, lo_x + y + G/8, ... of its XCD's contiguous range.  ntiles
// marks the end of the sequence.
template <typename T>
__device__ __forceinline__ int x_tile_seq(int step, int ntiles) {
  const int b = blockIdx.x, G = gridDim.x;
  if (G >= ntiles) return step == 0 ? fused_x_tile<T>() : ntiles;
  if ((sizeof(T) == 8 ? SPM_XREMAP_F64 : SPM_XREMAP_F32) && G % 8 == 0) {
    const int x = b % 8, y = b / 8, per = G / 8;
    const int lo = (int)((long)x * ntiles / 8), hi = (int)((long)(x + 1) * ntiles / 8);
    const long t = lo + y + (long)step * per;
    return t < hi ? (int)t : ntiles;
  }
  const long t = b + (long)step * G;
  return t < ntiles ? (int)t : ntiles;
}

// first step >= k of the block's sequence whose tile is kept (or the end)
template <typename T, int C>
__device__ __forceinline__ int x_next_kept(const Grid &g, int k, int ntiles) {
  for (;; ++k) {
    const int t = x_tile_seq<T>(k, ntiles);
    if (t >= ntiles || tile_kept<C>(g, t * C)) return k;
  }
}

// Zeta_k2u_k (fft_wrapper.cxx:408-511) on one mode; component d.
template <typename T>
__device__ __forceinline__ cplx<T> zeta2u_mode(const Grid &g, const ModeK &m, cplx<T> d1, cplx<T> d2, int d) {
  using C2 = cplx<T>;
  T kx = (T)(g.wave_x * m.kx), ky = (T)(g.wave_y * m.ky), kz = (T)(g.wave_z * m.kz);
  const T k2 = (T)(sq(g.wave_x * (double)m.kx) + sq(g.wave_y * (double)m.ky) + sq(g.wave_z * (double)m.kz));
  const T ik2 = k2 > (T)0 ? (T)1 / k2 : (T)0;
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
  if (d == 0) return C2{(ky * wi2 - kz * wi1), -(ky * wr2 - kz * wr1)};
  if (d == 1) return C2{(kz * wi0 - kx * wi2), -(kz * wr0 - kx * wr2)};
  return C2{(kx * wi1 - ky * wi0), -(kx * wr1 - ky * wr0)};
}

// Makes the values of two register arrays count as used and redefined here:
// what is computed from them before this point stays before it, and what is
// computed after it stays after it (a field's contribution to accumulators is
// then computed while its transform is in LDS rather than sunk, with the
// values it needs held in registers, to the end of the pass).
template <int R, typename C2>
__device__ __forceinline__ void pin_acc(C2 *acc0, C2 *acc1) {
#pragma unroll
  for (int q = 0; q < R; ++q) {
    asm volatile("" : "+v"(acc0[q].x), "+v"(acc0[q].y), "+v"(acc1[q].x), "+v"(acc1[q].y));
  }
}

// ---------------------------------------------------------------------
// x pass (inverse) with Zeta_k2u_k prologue: zeta (2 spectra) -> 3 spectra
// transformed along x.  Modes outside the 2/3 box read as zero (the in-place
// truncation of sp_3d_ns.cxx:75-77 / operate_omega.cxx:106-108).  The thread
// holding mode 0 sets u(0) = uk_dc, zeroing the FIX_CELL components first.
// ---------------------------------------------------------------------
template <typename T, int N, int R, int C, bool PERSIST>
__global__ __launch_bounds__(C *N / R, sizeof(T) == 8 ? SPM_X_ZETA2U_WAVES_F64 : 1) void x_zeta2u_kernel(Grid g, const T *__restrict__ z0,
                                                           const T *__restrict__ z1, T *u0, T *u1, T *u2,
                                                           double *uk_dc, int fix_mask,
                                                           const double2 *__restrict__ tw_g, int ntiles) {
  using C2 = cplx<T>;
  static_assert(N >= R, "one first-stage butterfly per thread");
  __shared__ C2 lds[N * C];
  const int tid = threadIdx.x;
  // twiddle tables in LDS: the transform's barriers then wait on LDS
  // only, and global loads issued ahead stay in flight across them
  __shared__ double2 tw[N];
  for (int t = tid; t < N; t += blockDim.x) tw[t] = tw_g[t];
  const int c = tid % C, j0 = tid / C;
  const long H = g.hp;
  const long es = (long)g.ny * H;
  const FieldBuf bz0 = field_buf(z0, 2 * (size_t)g.nx * es), bz1 = field_buf(z1, 2 * (size_t)g.nx * es);
  // Tiles of C columns (j, kz) numbered as one run j*(NZ/2+1) + kz (a tile
  // starts on a 128-byte boundary).  A block takes the kept tiles of its
  // sequence (x_tile_seq) in turn and loads the next one's zeta while the
  // current one is transformed; tiles outside the 2/3 box are zero and the
  // y pass never reads them.
  int k = x_next_kept<T, C>(g, 0, ntiles);
  int t = x_tile_seq<T>(k, ntiles);
  if (t >= ntiles) return;
  // zeta at the thread's first-stage points i = j0 + q N/R (the 2/3 mask
  // applied on load)
  auto fetch = [&](int tile, C2 *za, C2 *zb) {
    const int col = tile * C + c;
    const int j = col / (int)H, kz = col - j * (int)H;
    const bool col_ok = col < es && col_kept(g, j, kz);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = j0 + q * (N / R);
      const ModeK m = mode_k(g, i, j, kz);
      const bool ok = col_ok && m.keep;
      const long e = 2 * (col + i * es);
      za[q] = buf_load<T>(bz0, e, ok);
      zb[q] = buf_load<T>(bz1, e, ok);
    }
  };
  C2 za[R], zb[R], na[R], nb[R];
  fetch(t, za, zb);
  // uk_dc as it was on entry (the block holding mode 0 reads it before it
  // clears the FIX_CELL components), loaded outside the tile loop so that no
  // conditional load sits in the loop's pipeline of loads
  double dc_in[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) dc_in[d] = uk_dc[d];
  for (;;) {
    const int kn = x_next_kept<T, C>(g, k + 1, ntiles);
    const int tn = x_tile_seq<T>(kn, ntiles);
    if (PERSIST && tn < ntiles) fetch(tn, na, nb);
    const int c0 = t * C;
    const int col = c0 + c;
    const int j = col / (int)H, kz = col - j * (int)H;
    const bool col_ok = col < es && col_kept(g, j, kz);
    const bool has0 = c0 == 0;  // the tile holding column (j, kz) = (0, 0)
    double dc[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) dc[d] = (fix_mask & (1 << d)) ? 0.0 : dc_in[d];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      // each component's prologue is computed where its transform starts,
      // not hoisted (with its values held in registers) to the first one
      pin_acc<R>(za, zb);
      T *out = d == 0 ? u0 : (d == 1 ? u1 : u2);
      auto ld = [&](int i, int, int q) -> C2 {
        const ModeK m = mode_k(g, i, j, kz);
        C2 v = zeta2u_mode<T>(g, m, za[q], zb[q], d);
        if (has0 && i == 0 && c == 0) v.x = (T)dc[d];
        return v;
      };
      auto st = [&](int i, int, int, C2 v) {
        if (col_ok) *reinterpret_cast<C2 *>(out + 2 * (col + i * es)) = v;
      };
      fft_stages_io<T, N, R, C, C, -1, 1, true, true>(lds, tw, tid, ld, st);
    }
    if (has0) {  // the same in every thread of the block
      __syncthreads();  // every thread has read uk_dc
      if (tid == 0)
        for (int d = 0; d < 3; ++d)
          if (fix_mask & (1 << d)) uk_dc[d] = 0.0;
    }
    if (!PERSIST || tn >= ntiles) break;
    t = tn;
    k = kn;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      za[q] = na[q];
      zb[q] = nb[q];
    }
  }
}

// ---------------------------------------------------------------------
// z pass: C2R of (u0, u1, u2) -> real products of U2advection_k -> R2C of
// the five products into (p0 .. p4).  p0..p2 may alias u0..u2.
// ---------------------------------------------------------------------
// The ROWS x (M+1) complex values of a block's rows, KF per thread
// (element f = tid + t NT is (k = f % (M+1), row = f / (M+1))).
template <int M, int R, int ROWS>
struct ZRows {
  static constexpr int NT = ROWS * M / R, H = M + 1, CP = ROWS + 1, KF = (ROWS * H + NT - 1) / NT;
};

template <typename T, int M, int R, int ROWS>
__device__ __forceinline__ void z_rows_fetch(const T *a, long row0, int nrows, int nz_, int kz_cut, int tid,
                                             cplx<T> *dst) {
  using Z = ZRows<M, R, ROWS>;
#pragma unroll
  for (int t = 0; t < Z::KF; ++t) {
    const int f = tid + t * Z::NT, k = f % Z::H, r = f / Z::H;
    const long row = row0 + r;
    dst[t] = (f < ROWS * Z::H && row < nrows && k < kz_cut)
                 ? *reinterpret_cast<const cplx<T> *>(a + row * nz_ + 2 * k)
                 : cplx<T>{0, 0};
  }
}

// fetched rows -> LDS, then the C2R pre-twiddle that packs each row's half
// spectrum into M complex values for the inverse complex transform
template <typename T, int M, int R, int ROWS>
__device__ __forceinline__ void z_c2r_regs_to_lds(cplx<T> *lds, const cplx<T> *src,
                                                  const double2 *__restrict__ twr, int tid) {
  using C2 = cplx<T>;
  using Z = ZRows<M, R, ROWS>;
  constexpr int NT = Z::NT, CP = Z::CP, H = Z::H, HP = M / 2 + 1;
#pragma unroll
  for (int t = 0; t < Z::KF; ++t) {
    const int f = tid + t * NT;
    if (f < ROWS * H) lds[(f % H) * CP + f / H] = src[t];
  }
  __syncthreads();
  for (int f = tid; f < ROWS * HP; f += NT) {
    const int p = f % HP, r = f / HP;
    const int k = p, kk = M - p;
    const C2 xk = lds[k * CP + r], xkk = lds[kk * CP + r];
    {
      const C2 cj{xkk.x, -xkk.y};
      const C2 e = cadd(xk, cj), d = csub(xk, cj);
      const C2 w = C2{(T)twr[k].x, (T)twr[k].y};
      const C2 o = cmul(d, C2{w.x, -w.y});
      lds[k * CP + r] = C2{e.x - o.y, e.y + o.x};
    }
    if (p > 0 && kk != k) {
      const C2 cj{xk.x, -xk.y};
      const C2 e = cadd(xkk, cj), d = csub(xkk, cj);
      const C2 w = C2{(T)twr[kk].x, (T)twr[kk].y};
      const C2 o = cmul(d, C2{w.x, -w.y});
      lds[kk * CP + r] = C2{e.x - o.y, e.y + o.x};
    }
  }
  __syncthreads();
}

// writes the forward z transform of the LDS rows (packed real pairs) to `a`
template <typename T, int M, int R, int ROWS>
__device__ __forceinline__ void z_r2c_lds_to_row(cplx<T> *lds, T *a, long row0, int nrows, int nz_, int kz_cut,
                                                 const double2 *__restrict__ tw, const double2 *__restrict__ twr,
                                                 int tid) {
  using C2 = cplx<T>;
  constexpr int NT = ROWS * M / R, CP = ROWS + 1, H = M + 1;
  fft_stages<T, M, R, ROWS, CP, +1, 1>(lds, tw, tid);
  for (int f = tid; f < ROWS * H; f += NT) {
    const int k = f % H, r = f / H;
    const long row = row0 + r;
    const C2 z = lds[(k % M) * CP + r];
    C2 zc = lds[((M - k) % M) * CP + r];
    zc.y = -zc.y;
    const C2 e{(T)0.5 * (z.x + zc.x), (T)0.5 * (z.y + zc.y)};
    const C2 o{(T)0.5 * (z.y - zc.y), (T)-0.5 * (z.x - zc.x)};
    const C2 w = C2{(T)twr[k].x, (T)twr[k].y};
    const C2 x = cadd(e, cmul(o, w));
    if (row < nrows && k < kz_cut) *reinterpret_cast<C2 *>(a + row * nz_ + 2 * k) = x;
  }
  __syncthreads();
}

template <typename T, int M, int R, int ROWS>
__global__ __launch_bounds__(ROWS *M / R) void z_advect_kernel(Grid g, T *u0, T *u1, T *u2, T *p3,
                                                              T *p4, const double2 *__restrict__ tw_g,
                                                              const double2 *__restrict__ twr_g, T scale) {
  using C2 = cplx<T>;
  constexpr int NT = ROWS * M / R, CP = ROWS + 1, E = ROWS * M / NT;
  __shared__ C2 lds[(M + 1) * CP];
  const int tid = threadIdx.x;
  // twiddle tables in LDS: the transform's barriers then wait on LDS
  // only, and global loads issued ahead stay in flight across them
  __shared__ double2 tw[M];
  for (int t = tid; t < M; t += blockDim.x) tw[t] = tw_g[t];
  __shared__ double2 twr[M + 1];
  for (int t = tid; t < M + 1; t += blockDim.x) twr[t] = twr_g[t];
  const long row0 = (long)blockIdx.x * ROWS;
  const int nrows = g.nx * g.ny;
  C2 v[3][E];
  // component d+1 is fetched while component d is transformed
  constexpr int KF = ZRows<M, R, ROWS>::KF;
  C2 cur[KF], nxt[KF];
  // the inputs hold kept modes only below kz = TRN_Z, and only those
  // outputs are read on
  const int kz_cut = g.trn_z;
  z_rows_fetch<T, M, R, ROWS>(u0, row0, nrows, g.nz_, kz_cut, tid, cur);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d < 2) z_rows_fetch<T, M, R, ROWS>(d == 0 ? u1 : u2, row0, nrows, g.nz_, kz_cut, tid, nxt);
    z_c2r_regs_to_lds<T, M, R, ROWS>(lds, cur, twr, tid);
    if (d < 2)
#pragma unroll
      for (int t = 0; t < KF; ++t) cur[t] = nxt[t];
    fft_stages<T, M, R, ROWS, CP, -1, 1>(lds, tw, tid);
#pragma unroll
    for (int t = 0; t < E; ++t) {
      const int f = tid + t * NT, m = f % M, r = f / M;
      const C2 x = lds[m * CP + r];
      v[d][t] = C2{x.x * scale, x.y * scale};
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    T *out = q == 0 ? u0 : q == 1 ? u1 : q == 2 ? u2 : q == 3 ? p3 : p4;
#pragma unroll
    for (int t = 0; t < E; ++t) {
      const int f = tid + t * NT, m = f % M, r = f / M;
      C2 pr;
      const T a1 = v[0][t].x, a2 = v[1][t].x, a3 = v[2][t].x;
      const T b1 = v[0][t].y, b2 = v[1][t].y, b3 = v[2][t].y;
      if (q == 0) pr = C2{a2 * a3, b2 * b3};
      else if (q == 1) pr = C2{a3 * a1, b3 * b1};
      else if (q == 2) pr = C2{a1 * a2, b1 * b2};
      else if (q == 3) pr = C2{(a2 - a3) * (a2 + a3), (b2 - b3) * (b2 + b3)};
      else pr = C2{(a3 - a1) * (a3 + a1), (b3 - b1) * (b3 + b1)};
      lds[m * CP + r] = pr;
    }
    __syncthreads();
    z_r2c_lds_to_row<T, M, R, ROWS>(lds, out, row0, nrows, g.nz_, kz_cut, tw, twr, tid);
  }
}

// ---------------------------------------------------------------------
// z pass of the sheared solver (Zeta_k2advection_k_OBL, operate_omega.cxx:
// 115-146): C2R of the velocity (u0, u1, u2) and the vorticity (w0, w1,
// w2), the real-space u x w (operate_omega.cxx:123-146), R2C of its three
// components into u0..u2.  kz < TRN_Z entries only are read and written
// (the spectra are truncated; the update reads kept modes only).
// ---------------------------------------------------------------------
template <typename T, int M, int R, int ROWS>
__global__ __launch_bounds__(ROWS *M / R) void z_cross_kernel(Grid g, T *u0, T *u1, T *u2, const T *w0, const T *w1,
                                                             const T *w2, const double2 *__restrict__ tw_g,
                                                             const double2 *__restrict__ twr_g, T scale) {
  using C2 = cplx<T>;
  constexpr int NT = ROWS * M / R, CP = ROWS + 1, E = ROWS * M / NT;
  __shared__ C2 lds[(M + 1) * CP];
  const int tid = threadIdx.x;
  __shared__ double2 tw[M];
  for (int t = tid; t < M; t += blockDim.x) tw[t] = tw_g[t];
  __shared__ double2 twr[M + 1];
  for (int t = tid; t < M + 1; t += blockDim.x) twr[t] = twr_g[t];
  const long row0 = (long)blockIdx.x * ROWS;
  const int nrows = g.nx * g.ny;
  C2 v[6][E];
  constexpr int KF = ZRows<M, R, ROWS>::KF;
  C2 cur[KF], nxt[KF];
  const int kz_cut = g.trn_z;
  const T *const in[6] = {u0, u1, u2, w0, w1, w2};
  z_rows_fetch<T, M, R, ROWS>(in[0], row0, nrows, g.nz_, kz_cut, tid, cur);
#pragma unroll
  for (int d = 0; d < 6; ++d) {
    if (d < 5) z_rows_fetch<T, M, R, ROWS>(in[d + 1], row0, nrows, g.nz_, kz_cut, tid, nxt);
    z_c2r_regs_to_lds<T, M, R, ROWS>(lds, cur, twr, tid);
    if (d < 5)
#pragma unroll
      for (int t = 0; t < KF; ++t) cur[t] = nxt[t];
    fft_stages<T, M, R, ROWS, CP, -1, 1>(lds, tw, tid);
#pragma unroll
    for (int t = 0; t < E; ++t) {
      const int f = tid + t * NT, m = f % M, r = f / M;
      const C2 x = lds[m * CP + r];
      v[d][t] = C2{x.x * scale, x.y * scale};
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    T *out = q == 0 ? u0 : q == 1 ? u1 : u2;
#pragma unroll
    for (int t = 0; t < E; ++t) {
      const int f = tid + t * NT, m = f % M, r = f / M;
      // each complex slot packs two real points: the cross product of each
      const int a = (q + 1) % 3, b = (q + 2) % 3;  // (u x w)_q = u_a w_b - u_b w_a
      const C2 pr{v[a][t].x * v[3 + b][t].x - v[b][t].x * v[3 + a][t].x,
                  v[a][t].y * v[3 + b][t].y - v[b][t].y * v[3 + a][t].y};
      lds[m * CP + r] = pr;
    }
    __syncthreads();
    z_r2c_lds_to_row<T, M, R, ROWS>(lds, out, row0, nrows, g.nz_, kz_cut, tw, twr, tid);
  }
}

// Slot of the last Stockham stage (radix N / NS_last) that holds point
// j0 + off of a thread whose first-stage points are j0 + q N / R: the last
// stage's points are those same R points in another order.
template <int N, int R>
constexpr int last_slot_of(int off) {
  int ns = 1;
  while (ns * R < N) ns *= R;
  const int rs = N / ns;
  for (int q = 0; q < R; ++q)
    if ((q / rs) * (N / R) + (q % rs) * ns == off) return q;
  return -1;
}

// z_advect with the C2R pre-twiddle computed in the first Stockham stage of
// each inverse transform (from the pair X[k], X[M-k] in LDS), the inverse's
// last stage handing the real-space values to registers, and the forward's
// first stage taking the products from those registers: per block 3 x 9 +
// 5 x 9 LDS passes instead of 3 x 13 + 5 x 11, and fewer barriers.  Same
// arithmetic as z_advect_kernel.
template <typename T, int M, int R, int ROWS>
__global__ __launch_bounds__(ROWS *M / R, SPM_ZADV_WAVES) void z_advect_io_kernel(Grid g, T *u0, T *u1, T *u2, T *p3, T *p4,
                                                                 const double2 *__restrict__ tw_g,
                                                                 const double2 *__restrict__ twr_g, T scale) {
  using C2 = cplx<T>;
  using Z = ZRows<M, R, ROWS>;
  constexpr int NT = Z::NT, CP = Z::CP, H = Z::H;
  // the kept entries k < TRN_Z <= (2M + 2) / 3 + 1 of the block's rows
  constexpr int KF = (ROWS * ((2 * M + 2) / 3 + 1) + NT - 1) / NT;
  static_assert(M >= R, "one first-stage butterfly per thread");
  __shared__ C2 lds[H * CP];
  const int tid = threadIdx.x;
  __shared__ double2 tw[M];
  for (int t = tid; t < M; t += blockDim.x) tw[t] = tw_g[t];
  __shared__ double2 twr[M + 1];
  for (int t = tid; t < M + 1; t += blockDim.x) twr[t] = twr_g[t];
  const long row0 = (long)blockIdx.x * ROWS;
  const int nrows = g.nx * g.ny;
  const int kz_cut = g.trn_z;
  C2 v[3][R];  // real-space u (packed pairs) at the thread's R points of its row
  C2 cur[KF], nxt[KF];
  // element t of a thread: f = tid + t NT is entry k = f % kz_cut of row f / kz_cut
  auto fetch = [&](const T *a, C2 *dst) {
#pragma unroll
    for (int t = 0; t < KF; ++t) {
      const int f = tid + t * NT, k = f % kz_cut, r = f / kz_cut;
      const long row = row0 + r;
      dst[t] = (r < ROWS && row < nrows) ? *reinterpret_cast<const C2 *>(a + row * g.nz_ + 2 * k) : C2{0, 0};
    }
  };
  fetch(u0, cur);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d < 2) fetch(d == 0 ? u1 : u2, nxt);
#pragma unroll
    for (int t = 0; t < KF; ++t) {
      const int f = tid + t * NT, k = f % kz_cut, r = f / kz_cut;
      if (r < ROWS) lds[k * CP + r] = cur[t];
    }
    __syncthreads();
    if (d < 2)
#pragma unroll
      for (int t = 0; t < KF; ++t) cur[t] = nxt[t];
    // the C2R pre-twiddle that packs the half spectrum into M complex values
    // (entries k >= TRN_Z are zero and not held in LDS)
    auto ld = [&](int k, int c, int) -> C2 {
      const C2 xk = k < kz_cut ? lds[k * CP + c] : C2{0, 0};
      const C2 xkk = M - k < kz_cut ? lds[(M - k) * CP + c] : C2{0, 0};
      const C2 cj{xkk.x, -xkk.y};
      const C2 e = cadd(xk, cj), dd = csub(xk, cj);
      const C2 w = C2{(T)twr[k].x, (T)twr[k].y};
      const C2 o = cmul(dd, C2{w.x, -w.y});
      return C2{e.x - o.y, e.y + o.x};
    };
    auto st = [&](int, int, int q, C2 x) { v[d][q] = C2{x.x * scale, x.y * scale}; };
    fft_stages_io<T, M, R, ROWS, CP, -1, 1, true, true>(lds, tw, tid, ld, st);
    __syncthreads();  // every read of this transform's LDS precedes the next rows' writes
  }
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    T *out = q == 0 ? u0 : q == 1 ? u1 : q == 2 ? u2 : q == 3 ? p3 : p4;
    // first-stage slot s holds point j0 + s M / R: the inverse's last-stage
    // slot of that point
    auto ld = [&](int, int, int s) -> C2 {
      const int p = last_slot_of<M, R>(s * (M / R));
      const T a1 = v[0][p].x, a2 = v[1][p].x, a3 = v[2][p].x;
      const T b1 = v[0][p].y, b2 = v[1][p].y, b3 = v[2][p].y;
      if (q == 0) return C2{a2 * a3, b2 * b3};
      if (q == 1) return C2{a3 * a1, b3 * b1};
      if (q == 2) return C2{a1 * a2, b1 * b2};
      if (q == 3) return C2{(a2 - a3) * (a2 + a3), (b2 - b3) * (b2 + b3)};
      return C2{(a3 - a1) * (a3 + a1), (b3 - b1) * (b3 + b1)};
    };
    auto none = [](int, int, int, C2) {};
    fft_stages_io<T, M, R, ROWS, CP, +1, 1, true, false>(lds, tw, tid, ld, none);
    // R2C post-twiddle: split the M packed values into the half spectrum
    for (int f = tid; f < ROWS * H; f += NT) {
      const int k = f % H, r = f / H;
      const long row = row0 + r;
      const C2 z = lds[(k % M) * CP + r];
      C2 zc = lds[((M - k) % M) * CP + r];
      zc.y = -zc.y;
      const C2 e{(T)0.5 * (z.x + zc.x), (T)0.5 * (z.y + zc.y)};
      const C2 o{(T)0.5 * (z.y - zc.y), (T)-0.5 * (z.x - zc.x)};
      const C2 w = C2{(T)twr[k].x, (T)twr[k].y};
      const C2 x = cadd(e, cmul(o, w));
      if (row < nrows && k < kz_cut) *reinterpret_cast<C2 *>(out + row * g.nz_ + 2 * k) = x;
    }
    // the next product's first stage reads registers only and waits at its
    // own barrier before it writes LDS
  }
}

// ---------------------------------------------------------------------
// x pass (forward) of the five product spectra with the advection
// contraction, Omega_k2zeta_k and the slaved-Euler update as epilogue.
// The transformed products are not stored.
// ---------------------------------------------------------------------
template <typename T, int N, int R, int C, bool PERSIST>
__global__ __launch_bounds__(C *N / R, sizeof(T) == 8 ? SPM_X_UPDATE_WAVES_F64 : 1) void x_update_kernel(Grid g, const T *p0, const T *p1, const T *p2,
                                                           const T *p3, const T *p4, T *z0, T *z1, double nu,
                                                           double dt, const double2 *__restrict__ tw_g, int ntiles) {
  using C2 = cplx<T>;
  __shared__ C2 lds[N * C];
  const int tid = threadIdx.x;
  // twiddle tables in LDS: the transform's barriers then wait on LDS
  // only, and global loads issued ahead stay in flight across them
  __shared__ double2 tw[N];
  for (int t = tid; t < N; t += blockDim.x) tw[t] = tw_g[t];
  const int c = tid % C;
  const long H = g.hp;
  const long es = (long)g.ny * H;
  // kept tiles of the block's sequence (a column outside the 2/3 box gives
  // only discarded modes, and zeta is zero there already)
  int k = x_next_kept<T, C>(g, 0, ntiles);
  int t = x_tile_seq<T>(k, ntiles);
  if (t >= ntiles) return;
  const int j0 = tid / C;
  const size_t nel = 2 * (size_t)g.nx * es;  // elements per field
  const FieldBuf bp[5] = {field_buf(p0, nel), field_buf(p1, nel), field_buf(p2, nel), field_buf(p3, nel),
                          field_buf(p4, nel)};
  auto fetch = [&](int tile, int f, C2 *dst) {
    const int col = tile * C + c;
    const bool ok = col < es && col_kept(g, col / (int)H, col - (col / (int)H) * (int)H);
#pragma unroll
    for (int q = 0; q < R; ++q)
      dst[q] = buf_load<T>(bp[f], 2 * (col + (long)(j0 + q * (N / R)) * es), ok);
  };
  // Software pipeline over the five spectra and over tiles: the first-stage
  // points of the next spectrum (the next tile's first one after the
  // fifth) are loaded while the current one is transformed in LDS; the
  // barriers of the transform wait on LDS only.
  C2 cur[R], nxt[R];
  fetch(t, 0, cur);
  for (;;) {
    const int kn = x_next_kept<T, C>(g, k + 1, ntiles);
    const int tn = x_tile_seq<T>(kn, ntiles);
    const int col = t * C + c;  // flat column j*(NZ/2+1) + kz
    const int j = col / (int)H, kz = col - j * (int)H;
    const bool col_ok = col < es && col_kept(g, j, kz);
    // The contraction (operate_omega.cxx:60-97) is linear in the five
    // product spectra P0..P4 = u2u3, u3u1, u1u2, u2^2-u3^2, u3^2-u1^2:
    // accumulate the two components Omega_k2zeta_k keeps as each product is
    // transformed.
    C2 acc0[R], acc1[R];
    int rows[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      acc0[q] = acc1[q] = C2{0, 0};
      rows[q] = 0;
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      if (f < 4)
        fetch(t, f + 1, nxt);
      else if (PERSIST && tn < ntiles)
        fetch(tn, 0, nxt);
      auto ld = [&](int, int, int q) -> C2 { return cur[q]; };
      auto st = [&](int i, int, int q, C2 P) {
        rows[q] = i;
        const ModeK m = mode_k(g, i, j, kz);
        const T k1 = (T)(m.kx * g.wave_x), k2 = (T)(m.ky * g.wave_y), k3 = (T)(m.kz * g.wave_z);
        // coefficient of P_f in -Omega_s, s = 0, 1, 2
        T cf[3];
        if (f == 0) {
          cf[0] = (k3 - k2) * (k3 + k2); cf[1] = k1 * k2; cf[2] = -(k3 * k1);
        } else if (f == 1) {
          cf[0] = -(k1 * k2); cf[1] = (k1 - k3) * (k1 + k3); cf[2] = k2 * k3;
        } else if (f == 2) {
          cf[0] = k3 * k1; cf[1] = -(k2 * k3); cf[2] = (k2 - k1) * (k2 + k1);
        } else if (f == 3) {
          cf[0] = k2 * k3; cf[1] = 0; cf[2] = -(k1 * k2);
        } else {
          cf[0] = 0; cf[1] = k3 * k1; cf[2] = -(k1 * k2);
        }
        const T a = m.kx != 0 ? cf[1] : (m.ky != 0 ? cf[2] : cf[0]);
        const T b = m.kx != 0 ? cf[2] : (m.ky != 0 ? cf[0] : cf[1]);
        acc0[q].x -= a * P.x;
        acc0[q].y -= a * P.y;
        acc1[q].x -= b * P.x;
        acc1[q].y -= b * P.y;
      };
      fft_stages_io<T, N, R, C, C, +1, 1, true, true>(lds, tw, tid, ld, st);
      pin_acc<R>(acc0, acc1);
#pragma unroll
      for (int q = 0; q < R; ++q) cur[q] = nxt[q];
    }
    if (col_ok) {
      const T dmy0 = (T)(-nu * dt), dmy1 = (T)(1. / nu);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = rows[q];
        const ModeK m = mode_k(g, i, j, kz);
        const long e = 2 * (col + i * es);
        if (!m.keep) continue;  // zeta stays zero outside the 2/3 box
        const T kk2 = (T)(sq(g.wave_x * (double)m.kx) + sq(g.wave_y * (double)m.ky) + sq(g.wave_z * (double)m.kz));
        const T ik2 = kk2 > (T)0 ? (T)1 / kk2 : (T)0;
        const T dmy = exp(dmy0 * kk2) - (T)1;
        C2 za = *reinterpret_cast<const C2 *>(z0 + e);
        C2 zb = *reinterpret_cast<const C2 *>(z1 + e);
        za.x += dmy * (za.x - dmy1 * ik2 * acc0[q].x);
        za.y += dmy * (za.y - dmy1 * ik2 * acc0[q].y);
        zb.x += dmy * (zb.x - dmy1 * ik2 * acc1[q].x);
        zb.y += dmy * (zb.y - dmy1 * ik2 * acc1[q].y);
        *reinterpret_cast<C2 *>(z0 + e) = za;
        *reinterpret_cast<C2 *>(z1 + e) = zb;
      }
    }
    if (!PERSIST || tn >= ntiles) break;
    t = tn;
    k = kn;
  }
}

// Coefficients (a, b) of the product spectrum P_f in the two components
// Omega_k2zeta_k keeps of -Omega (operate_omega.cxx:60-98) at mode m.
template <typename T>
__device__ __forceinline__ void advect_coef(const Grid &g, const ModeK &m, int f, T &a, T &b) {
  const T k1 = (T)(m.kx * g.wave_x), k2 = (T)(m.ky * g.wave_y), k3 = (T)(m.kz * g.wave_z);
  T cf[3];
  if (f == 0) {
    cf[0] = (k3 - k2) * (k3 + k2); cf[1] = k1 * k2; cf[2] = -(k3 * k1);
  } else if (f == 1) {
    cf[0] = -(k1 * k2); cf[1] = (k1 - k3) * (k1 + k3); cf[2] = k2 * k3;
  } else if (f == 2) {
    cf[0] = k3 * k1; cf[1] = -(k2 * k3); cf[2] = (k2 - k1) * (k2 + k1);
  } else if (f == 3) {
    cf[0] = k2 * k3; cf[1] = 0; cf[2] = -(k1 * k2);
  } else {
    cf[0] = 0; cf[1] = k3 * k1; cf[2] = -(k1 * k2);
  }
  a = m.kx != 0 ? cf[1] : (m.ky != 0 ? cf[2] : cf[0]);
  b = m.kx != 0 ? cf[2] : (m.ky != 0 ? cf[0] : cf[1]);
}

// Byte offsets of a thread's direct-to-LDS tile loads (the same in every
// field): FP64, its R first-stage points; FP32, its 2R dwords (dma_tile4).
// Columns outside the field or the 2/3 box read as zero.
template <typename T, int N, int R, int C>
__device__ __forceinline__ void tile_offsets(const Grid &g, int t, int tid, long es, unsigned *off) {
  const long H = g.hp;
  auto col_ok = [&](int col) {
    const int j = col / (int)H, kz = col - j * (int)H;
    return col < es && col_kept(g, j, kz);
  };
  if constexpr (sizeof(T) == 8) {
    const int col = t * C + tid % C, j0 = tid / C;
    const bool ok = col_ok(col);
#pragma unroll
    for (int q = 0; q < R; ++q)
      off[q] = opaque(ok ? (unsigned)(2 * (col + (long)(j0 + q * (N / R)) * es) * (long)sizeof(T)) : 0xfffffff0u);
  } else {
    constexpr int NT = C * N / R;
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) {
      const int f = k * NT + tid, row = f / (2 * C), cc = (f % (2 * C)) / 2, part = f & 1;
      const int col = t * C + cc;
      off[k] = opaque(col_ok(col) ? (unsigned)((2 * (col + (long)row * es) + part) * (long)sizeof(T)) : 0xfffffff0u);
    }
  }
}

template <typename T, int N, int R, int C>
__device__ __forceinline__ void dma_any(const FieldBuf &b, cplx<T> *buf, int tid, const unsigned *off) {
  if constexpr (sizeof(T) == 8)
    dma_tile<N, R, C>(b, buf, tid, off);
  else
    dma_tile4<N, R, C>(b, buf, tid, off);
}

// Before a transform reads a tile an FP32 block loaded: every wave's loads
// have landed (its loads in flight are this tile's only) and every wave has
// passed the barrier.  FP64 tiles are read by the lanes that loaded them.
template <typename T>
__device__ __forceinline__ void tile_landed() {
  if constexpr (sizeof(T) == 4) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    lds_barrier();
  }
}

// x_update with direct-to-LDS loads: one block per tile, two LDS tile
// buffers, the next product spectrum in flight into one buffer while the
// current one is transformed in the other; no VGPRs hold loads in flight,
// so two blocks (FP64) fit a CU.  Same arithmetic as x_update_kernel.
template <typename T, int N, int R, int C, bool ZU>
__global__ __launch_bounds__(C *N / R, 4) void x_update_dma_kernel(Grid g, const T *p0, const T *p1, const T *p2,
                                                                   const T *p3, const T *p4, T *z0, T *z1,
                                                                   double nu, double dt,
                                                                   const double2 *__restrict__ tw_g, T *u0, T *u1,
                                                                   T *u2, double *uk_dc, int fix_mask) {
  using C2 = cplx<T>;
  __shared__ C2 bufA[N * C];
  __shared__ C2 bufB[N * C];
  __shared__ double2 tw[N];
  const int tid = threadIdx.x;
  const int t = fused_x_tile<T>();
  if (!tile_kept<C>(g, t * C)) return;  // zeta stays zero outside the 2/3 box
  for (int i = tid; i < N; i += blockDim.x) tw[i] = tw_g[i];
  // uk_dc as it was on entry (only the block holding mode 0 uses it, and
  // it clears the FIX_CELL components after every thread has read them)
  double dc_in[3] = {0.0, 0.0, 0.0};
  if constexpr (ZU)
#pragma unroll
    for (int d = 0; d < 3; ++d) dc_in[d] = uk_dc[d];
  const int c = tid % C, j0 = tid / C;
  const long H = g.hp;
  const long es = (long)g.ny * H;
  const int col = t * C + c;  // flat column j*(NZ/2+1) + kz
  const int j = col / (int)H, kz = col - j * (int)H;
  const bool col_ok = col < es && col_kept(g, j, kz);
  const size_t nel = 2 * (size_t)g.nx * es;
  const FieldBuf bp[5] = {field_buf(p0, nel), field_buf(p1, nel), field_buf(p2, nel), field_buf(p3, nel),
                          field_buf(p4, nel)};
  unsigned off[sizeof(T) == 8 ? R : 2 * R];
  tile_offsets<T, N, R, C>(g, t, tid, es, off);
  dma_any<T, N, R, C>(bp[0], bufA, tid, off);
  const FieldBuf bz0 = field_buf<T>(z0, nel), bz1 = field_buf<T>(z1, nel);
  C2 acc0[R], acc1[R], za[R], zb[R];
#pragma unroll
  for (int q = 0; q < R; ++q) acc0[q] = acc1[q] = C2{0, 0};
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    C2 *buf = f % 2 == 0 ? bufA : bufB;
    C2 *oth = f % 2 == 0 ? bufB : bufA;
    auto hook = [&]() {
      if (f < 4) {
        dma_any<T, N, R, C>(bp[f + 1], oth, tid, off);
      } else {
        // zeta at the modes this thread updates, loaded while the last
        // spectrum is transformed
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int i = last_stage_point<N, R>(j0, q);
          const bool ok = col_ok && mode_k(g, i, j, kz).keep;
          const long e = 2 * (col + (long)i * es);
          za[q] = buf_load<T>(bz0, e, ok);
          zb[q] = buf_load<T>(bz1, e, ok);
        }
      }
    };
    auto st = [&](int i, int, int q, C2 P) {
      T a, b;
      advect_coef<T>(g, mode_k(g, opaque(i), j, kz), f, a, b);
      acc0[q].x -= a * P.x;
      acc0[q].y -= a * P.y;
      acc1[q].x -= b * P.x;
      acc1[q].y -= b * P.y;
      __builtin_amdgcn_sched_barrier(0);  // points in turn: fewer live registers
    };
    tile_landed<T>();
    fft_stages_lds<T, N, R, C, +1, 1>(buf, tw, tid, hook, st);
    pin_acc<R>(acc0, acc1);
  }
  if constexpr (ZU) {
    lds_barrier();  // the last transform has read bufA
  } else {
    if (!col_ok) return;
  }
  const T dmy0 = (T)(-nu * dt), dmy1 = (T)(1. / nu);
#pragma unroll
  for (int q = 0; q < R; ++q) {
    // one mode after the other (each takes an exponential and a division)
    __builtin_amdgcn_sched_barrier(0);
    const int i = opaque_after_memory(last_stage_point<N, R>(j0, q));
    const ModeK m = mode_k(g, i, j, kz);
    if (ZU && !(col_ok && m.keep)) {  // zero, as Zeta_k2u_k reads it there
      bufA[i * C + c] = bufB[i * C + c] = C2{0, 0};
      continue;
    }
    if (!m.keep) continue;  // zeta stays zero outside the 2/3 box
    const long e = 2 * (col + (long)i * es);
    const T kk2 = (T)(sq(g.wave_x * (double)m.kx) + sq(g.wave_y * (double)m.ky) + sq(g.wave_z * (double)m.kz));
    const T ik2 = kk2 > (T)0 ? (T)1 / kk2 : (T)0;
    const T dmy = exp(dmy0 * kk2) - (T)1;
    C2 a = za[q], b = zb[q];
    a.x += dmy * (a.x - dmy1 * ik2 * acc0[q].x);
    a.y += dmy * (a.y - dmy1 * ik2 * acc0[q].y);
    b.x += dmy * (b.x - dmy1 * ik2 * acc1[q].x);
    b.y += dmy * (b.y - dmy1 * ik2 * acc1[q].y);
    *reinterpret_cast<C2 *>(z0 + e) = a;
    *reinterpret_cast<C2 *>(z1 + e) = b;
    if constexpr (ZU) {
      bufA[i * C + c] = a;
      bufB[i * C + c] = b;
    }
  }
  if constexpr (ZU) {
    // FIX_CELL, truncation and Zeta_k2u_k on the new zeta, then the inverse
    // x transforms of the three velocity spectra, as x_zeta2u_kernel does
    // from zeta in memory: the new zeta moves from the forward transform's
    // last-stage points to the inverse one's first-stage points through LDS
    lds_barrier();
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = j0 + q * (N / R);
      za[q] = bufA[i * C + c];
      zb[q] = bufB[i * C + c];
    }
    lds_barrier();
    const bool has0 = t * C == 0;  // the tile holding column (j, kz) = (0, 0)
    double dc[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) dc[d] = (fix_mask & (1 << d)) ? 0.0 : dc_in[d];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      pin_acc<R>(za, zb);
      T *out = d == 0 ? u0 : (d == 1 ? u1 : u2);
      auto ld = [&](int i, int, int q) -> C2 {
        const ModeK m = mode_k(g, i, j, kz);
        C2 v = zeta2u_mode<T>(g, m, za[q], zb[q], d);
        if (has0 && i == 0 && c == 0) v.x = (T)dc[d];
        return v;
      };
      auto st = [&](int i, int, int, C2 v) {
        if (col_ok) *reinterpret_cast<C2 *>(out + 2 * (col + i * es)) = v;
      };
      fft_stages_io<T, N, R, C, C, -1, 1, true, true>(bufA, tw, tid, ld, st);
    }
    if (has0) {  // the same in every thread of the block
      __syncthreads();  // every thread has read uk_dc
      if (tid == 0)
        for (int d = 0; d < 3; ++d)
          if (fix_mask & (1 << d)) uk_dc[d] = 0.0;
    }
  }
}

// ---------------------------------------------------------------------
// x pass (forward) of f, then zeta += U_k2zeta_k(mask(f)) and
// uk_dc += Re f(0).  The reference projects f (Solenoidal_u), adds it to u
// in real space and takes U2zeta_k of the sum; U_k2zeta_k is linear, maps
// the spectrum of u back to zeta exactly (u = Zeta_k2u(zeta)) and sends the
// gradient part removed by the projection to zero, while the projection
// keeps the k = 0 mode (IK2(0) = 0).  So the result is the same field.
// ---------------------------------------------------------------------
template <typename T, int N, int R, int C, bool PERSIST>
__global__ __launch_bounds__(C *N / R, sizeof(T) == 8 ? SPM_X_PROJECT_WAVES_F64 : 1) void x_project_kernel(Grid g, const T *f0, const T *f1, const T *f2, T *z0,
                                                            T *z1, double *uk_dc, const double2 *__restrict__ tw_g,
                                                            int ntiles) {
  using C2 = cplx<T>;
  __shared__ C2 lds[N * C];
  const int tid = threadIdx.x;
  // twiddle tables in LDS: the transform's barriers then wait on LDS
  // only, and global loads issued ahead stay in flight across them
  __shared__ double2 tw[N];
  for (int t = tid; t < N; t += blockDim.x) tw[t] = tw_g[t];
  const int c = tid % C;
  const long H = g.hp;
  const long es = (long)g.ny * H;
  // kept tiles of the block's sequence (a column outside the 2/3 box gives
  // only discarded modes, and zeta is zero there already)
  int k = x_next_kept<T, C>(g, 0, ntiles);
  int t = x_tile_seq<T>(k, ntiles);
  if (t >= ntiles) return;
  const int j0 = tid / C;
  const size_t nel = 2 * (size_t)g.nx * es;  // elements per field
  const FieldBuf bf[3] = {field_buf(f0, nel), field_buf(f1, nel), field_buf(f2, nel)};
  auto fetch = [&](int tile, int f, C2 *dst) {
    const int col = tile * C + c;
    const bool ok = col < es && col_kept(g, col / (int)H, col - (col / (int)H) * (int)H);
#pragma unroll
    for (int q = 0; q < R; ++q)
      dst[q] = buf_load<T>(bf[f], 2 * (col + (long)(j0 + q * (N / R)) * es), ok);
  };
  // software pipeline over the three components and over tiles, as in
  // x_update_kernel
  C2 cur[R], nxt[R];
  fetch(t, 0, cur);
  for (;;) {
    const int kn = x_next_kept<T, C>(g, k + 1, ntiles);
    const int tn = x_tile_seq<T>(kn, ntiles);
    const int c0 = t * C;
    const int col = c0 + c;  // flat column j*(NZ/2+1) + kz
    const int j = col / (int)H, kz = col - j * (int)H;
    const bool col_ok = col < es && col_kept(g, j, kz);
    C2 acc0[R], acc1[R];
    int rows[R];
    double f00[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < R; ++q) {
      acc0[q] = acc1[q] = C2{0, 0};
      rows[q] = 0;
    }
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      if (f < 2)
        fetch(t, f + 1, nxt);
      else if (PERSIST && tn < ntiles)
        fetch(tn, 0, nxt);
      auto ld = [&](int, int, int q) -> C2 { return cur[q]; };
      auto st = [&](int i, int, int q, C2 F) {
        rows[q] = i;
        if (i == 0 && c == 0) f00[f] = F.x;
        const ModeK m = mode_k(g, i, j, kz);
        if (!m.keep) return;
        const T kx = (T)(m.kx * g.wave_x), ky = (T)(m.ky * g.wave_y), kz = (T)(m.kz * g.wave_z);
        // omega_s of U_k2zeta_k's curl gets (ks[a] F_b - ks[b] F_a), (s, a, b)
        // cyclic; omega_re from Im F, omega_im from -Re F.  Coefficient of F_f in
        // omega_s: (0,2) ky, (0,1) -kz, (1,0) kz, (1,2) -kx, (2,1) kx, (2,0) -ky.
        auto coef = [&](int sidx) -> T {
          if (sidx == 0) return f == 2 ? ky : (f == 1 ? -kz : (T)0);
          if (sidx == 1) return f == 0 ? kz : (f == 2 ? -kx : (T)0);
          return f == 1 ? kx : (f == 0 ? -ky : (T)0);
        };
        const int s0 = m.kx != 0 ? 1 : (m.ky != 0 ? 2 : 0);
        const int s1 = m.kx != 0 ? 2 : (m.ky != 0 ? 0 : 1);
        const T a = coef(s0), b = coef(s1);
        acc0[q].x += a * F.y;
        acc0[q].y -= a * F.x;
        acc1[q].x += b * F.y;
        acc1[q].y -= b * F.x;
      };
      fft_stages_io<T, N, R, C, C, +1, 1, true, true>(lds, tw, tid, ld, st);
      pin_acc<R>(acc0, acc1);
#pragma unroll
      for (int q = 0; q < R; ++q) cur[q] = nxt[q];
    }
    const bool has0 = c0 == 0;
    if (col_ok) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = rows[q];
        const ModeK m = mode_k(g, i, j, kz);
        const long e = 2 * (col + i * es);
        if (has0 && i == 0 && c == 0)
          for (int d = 0; d < 3; ++d) uk_dc[d] += f00[d];
        if (!m.keep) continue;  // zeta stays zero outside the 2/3 box
        *reinterpret_cast<C2 *>(z0 + e) = cadd(*reinterpret_cast<const C2 *>(z0 + e), acc0[q]);
        *reinterpret_cast<C2 *>(z1 + e) = cadd(*reinterpret_cast<const C2 *>(z1 + e), acc1[q]);
      }
    }
    if (!PERSIST || tn >= ntiles) break;
    t = tn;
    k = kn;
  }
}

// Coefficients (a, b) of the component F_f of f in the two components of
// U_k2zeta_k's curl kept for zeta at mode m: omega_s gets (ks[a] F_b -
// ks[b] F_a), (s, a, b) cyclic, so F_f enters omega_s with (0,2) ky,
// (0,1) -kz, (1,0) kz, (1,2) -kx, (2,1) kx, (2,0) -ky.
template <typename T>
__device__ __forceinline__ void project_coef(const Grid &g, const ModeK &m, int f, T &a, T &b) {
  const T kx = (T)(m.kx * g.wave_x), ky = (T)(m.ky * g.wave_y), kz = (T)(m.kz * g.wave_z);
  auto coef = [&](int sidx) -> T {
    if (sidx == 0) return f == 2 ? ky : (f == 1 ? -kz : (T)0);
    if (sidx == 1) return f == 0 ? kz : (f == 2 ? -kx : (T)0);
    return f == 1 ? kx : (f == 0 ? -ky : (T)0);
  };
  const int s0 = m.kx != 0 ? 1 : (m.ky != 0 ? 2 : 0);
  const int s1 = m.kx != 0 ? 2 : (m.ky != 0 ? 0 : 1);
  a = coef(s0);
  b = coef(s1);
}

// x_project for FP64 fields with direct-to-LDS loads (see
// x_update_dma_kernel): the three components of f through two LDS buffers,
// one block per tile.  Same arithmetic as x_project_kernel.
template <typename T, int N, int R, int C>
__global__ __launch_bounds__(C *N / R, 4) void x_project_dma_kernel(Grid g, const T *f0, const T *f1, const T *f2,
                                                                    T *z0, T *z1, double *uk_dc,
                                                                    const double2 *__restrict__ tw_g) {
  using C2 = cplx<T>;
  __shared__ C2 bufA[N * C];
  __shared__ C2 bufB[N * C];
  __shared__ double2 tw[N];
  const int tid = threadIdx.x;
  const int t = fused_x_tile<T>();
  if (!tile_kept<C>(g, t * C)) return;  // zeta stays zero outside the 2/3 box
  for (int i = tid; i < N; i += blockDim.x) tw[i] = tw_g[i];
  const int c = tid % C, j0 = tid / C;
  const long H = g.hp;
  const long es = (long)g.ny * H;
  const int c0 = t * C;
  const int col = c0 + c;  // flat column j*(NZ/2+1) + kz
  const int j = col / (int)H, kz = col - j * (int)H;
  const bool col_ok = col < es && col_kept(g, j, kz);
  const size_t nel = 2 * (size_t)g.nx * es;
  const FieldBuf bf[3] = {field_buf(f0, nel), field_buf(f1, nel), field_buf(f2, nel)};
  unsigned off[sizeof(T) == 8 ? R : 2 * R];
  tile_offsets<T, N, R, C>(g, t, tid, es, off);
  dma_any<T, N, R, C>(bf[0], bufA, tid, off);
  const FieldBuf bz0 = field_buf<T>(z0, nel), bz1 = field_buf<T>(z1, nel);
  C2 acc0[R], acc1[R], za[R], zb[R];
  double f00[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < R; ++q) acc0[q] = acc1[q] = C2{0, 0};
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    C2 *buf = f % 2 == 0 ? bufA : bufB;
    C2 *oth = f % 2 == 0 ? bufB : bufA;
    auto hook = [&]() {
      if (f < 2) {
        dma_any<T, N, R, C>(bf[f + 1], oth, tid, off);
      } else {
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int i = last_stage_point<N, R>(j0, q);
          const bool ok = col_ok && mode_k(g, i, j, kz).keep;
          const long e = 2 * (col + (long)i * es);
          za[q] = buf_load<T>(bz0, e, ok);
          zb[q] = buf_load<T>(bz1, e, ok);
        }
      }
    };
    auto st = [&](int i, int, int q, C2 F) {
      // slot 0 of thread 0 is mode 0 of the tile's first column
      if (q == 0) f00[f] = F.x;
      T a, b;  // not used where the mode is discarded
      project_coef<T>(g, mode_k(g, opaque(i), j, kz), f, a, b);
      acc0[q].x += a * F.y;
      acc0[q].y -= a * F.x;
      acc1[q].x += b * F.y;
      acc1[q].y -= b * F.x;
    };
    tile_landed<T>();
    fft_stages_lds<T, N, R, C, +1, 1>(buf, tw, tid, hook, st);
    pin_acc<R>(acc0, acc1);
  }
  if (!col_ok) return;
  if (c0 == 0 && tid == 0)  // mode (0, 0, 0)
    for (int d = 0; d < 3; ++d) uk_dc[d] += f00[d];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = last_stage_point<N, R>(j0, q);
    if (!mode_k(g, i, j, kz).keep) continue;  // zeta stays zero outside the 2/3 box
    const long e = 2 * (col + (long)i * es);
    *reinterpret_cast<C2 *>(z0 + e) = cadd(za[q], acc0[q]);
    *reinterpret_cast<C2 *>(z1 + e) = cadd(zb[q], acc1[q]);
  }
}

#define SPM_SIZE_SWITCH(n, ...)                \
  switch (n) {                                  \
    case 8: { constexpr int S = 8; __VA_ARGS__; } break;     \
    case 16: { constexpr int S = 16; __VA_ARGS__; } break;   \
    case 32: { constexpr int S = 32; __VA_ARGS__; } break;   \
    case 64: { constexpr int S = 64; __VA_ARGS__; } break;   \
    case 128: { constexpr int S = 128; __VA_ARGS__; } break; \
    case 256: { constexpr int S = 256; __VA_ARGS__; } break; \
    case 512: { constexpr int S = 512; __VA_ARGS__; } break; \
    case 1024: { constexpr int S = 1024; __VA_ARGS__; } break; \
    default: return SPM_ERR_UNSUPPORTED;        \
  }

#define SPM_ZSIZE_SWITCH(n, ...)               \
  switch (n) {                                  \
    case 4: { constexpr int S = 4; __VA_ARGS__; } break;     \
    case 8: { constexpr int S = 8; __VA_ARGS__; } break;     \
    case 16: { constexpr int S = 16; __VA_ARGS__; } break;   \
    case 32: { constexpr int S = 32; __VA_ARGS__; } break;   \
    case 64: { constexpr int S = 64; __VA_ARGS__; } break;   \
    case 128: { constexpr int S = 128; __VA_ARGS__; } break; \
    case 256: { constexpr int S = 256; __VA_ARGS__; } break; \
    case 512: { constexpr int S = 512; __VA_ARGS__; } break; \
    default: return SPM_ERR_UNSUPPORTED;        \
  }

inline double field_bytes(const spm_ctx *c) {
  return (double)ref_count(c->grid) * (c->prm.precision == SPM_FP32 ? sizeof(float) : sizeof(double));
}

template <typename T>
inline T *F(double *p) {
  return reinterpret_cast<T *>(p);
}

// Blocks of a persistent x pass: `per_cu` per compute unit (0: one block
// per tile).
#ifndef SPM_X_ZETA2U_BLOCKS_F64
#define SPM_X_ZETA2U_BLOCKS_F64 0
#endif
#ifndef SPM_X_ZETA2U_BLOCKS_F32
#define SPM_X_ZETA2U_BLOCKS_F32 0
#endif
#ifndef SPM_X_UPDATE_BLOCKS_F64
#define SPM_X_UPDATE_BLOCKS_F64 1
#endif
#ifndef SPM_X_UPDATE_BLOCKS_F32
#define SPM_X_UPDATE_BLOCKS_F32 0
#endif
#ifndef SPM_X_PROJECT_BLOCKS_F64
#define SPM_X_PROJECT_BLOCKS_F64 0
#endif
#ifndef SPM_X_PROJECT_BLOCKS_F32
#define SPM_X_PROJECT_BLOCKS_F32 0
#endif
// z_advect with the pre- and post-twiddles and the products fused into
// the Stockham stages (0: the kernel with separate LDS passes)
#ifndef SPM_ZADV_IO
#define SPM_ZADV_IO 1
#endif
// x_update / x_project through direct-to-LDS loads (0: the register-prefetch
// kernels)
#ifndef SPM_X_DMA_F64
#define SPM_X_DMA_F64 1
#endif
#ifndef SPM_X_DMA_F32
#define SPM_X_DMA_F32 1
#endif
// the fused step runs x_update and the following x_zeta2u as one launch
#ifndef SPM_X_UPDATE_ZU
#define SPM_X_UPDATE_ZU 1
#endif
inline int x_grid(const spm_ctx *c, int ntiles, int per_cu) {
  const int n = per_cu > 0 ? per_cu * c->n_cu : ntiles;
  return n < ntiles ? n : ntiles;
}

template <typename T>
int fx_zeta2u_t(spm_ctx *c, int fix_mask) {
  const Grid &g = c->grid;
  SPM_SIZE_SWITCH(g.nx, {
    constexpr int R = kFusedR(S), C = kFusedC(S);
    const int ntiles = (g.ny * g.hp + C - 1) / C;
    constexpr int per_cu = sizeof(T) == 8 ? SPM_X_ZETA2U_BLOCKS_F64 : SPM_X_ZETA2U_BLOCKS_F32;
    hipLaunchKernelGGL((x_zeta2u_kernel<T, S, R, C, (per_cu > 0)>), dim3(x_grid(c, ntiles, per_cu)), dim3(C * S / R), 0, c->stream, g, F<T>(c->zeta[0]), F<T>(c->zeta[1]), F<T>(c->u[0]), F<T>(c->u[1]), F<T>(c->u[2]),
                       c->uk_dc, fix_mask, c->tw_x, ntiles);
  });
  return 0;
}

template <typename T>
int fz_advect_t(spm_ctx *c) {
  const Grid &g = c->grid;
  const T scale = (T)(1.0 / ((double)g.nx * g.ny * g.nz));
  SPM_ZSIZE_SWITCH(g.nz / 2, {
    constexpr int R = kZadvR(S), ROWS = kZadvRows<T>(S);
    if constexpr (SPM_ZADV_IO && S >= R)
      hipLaunchKernelGGL((z_advect_io_kernel<T, S, R, ROWS>), dim3((g.nx * g.ny + ROWS - 1) / ROWS),
                         dim3(ROWS * S / R), 0, c->stream, g, F<T>(c->u[0]), F<T>(c->u[1]), F<T>(c->u[2]),
                         F<T>(c->work[0]), F<T>(c->work[1]), c->tw_z, c->tw_rz, scale);
    else
      hipLaunchKernelGGL((z_advect_kernel<T, S, R, ROWS>), dim3((g.nx * g.ny + ROWS - 1) / ROWS), dim3(ROWS * S / R),
                         0, c->stream, g, F<T>(c->u[0]), F<T>(c->u[1]), F<T>(c->u[2]), F<T>(c->work[0]),
                         F<T>(c->work[1]), c->tw_z, c->tw_rz, scale);
  });
  return 0;
}

template <typename T>
int fx_update_t(spm_ctx *c, double nu, double dt) {
  const Grid &g = c->grid;
  SPM_SIZE_SWITCH(g.nx, {
    constexpr int R = kFusedR(S), C = kFusedC(S);
    const int ntiles = (g.ny * g.hp + C - 1) / C;
    constexpr int per_cu = sizeof(T) == 8 ? SPM_X_UPDATE_BLOCKS_F64 : SPM_X_UPDATE_BLOCKS_F32;
    if constexpr (sizeof(T) == 8 ? SPM_X_DMA_F64 : SPM_X_DMA_F32) {
      hipLaunchKernelGGL((x_update_dma_kernel<T, S, R, C, false>), dim3(ntiles), dim3(C * S / R), 0, c->stream, g,
                         F<T>(c->u[0]), F<T>(c->u[1]), F<T>(c->u[2]), F<T>(c->work[0]), F<T>(c->work[1]),
                         F<T>(c->zeta[0]), F<T>(c->zeta[1]), nu, dt, c->tw_x, nullptr, nullptr, nullptr,
                         nullptr, 0);
      return 0;
    }
    hipLaunchKernelGGL((x_update_kernel<T, S, R, C, (per_cu > 0)>), dim3(x_grid(c, ntiles, per_cu)), dim3(C * S / R), 0, c->stream, g, F<T>(c->u[0]),
                       F<T>(c->u[1]), F<T>(c->u[2]), F<T>(c->work[0]), F<T>(c->work[1]), F<T>(c->zeta[0]),
                       F<T>(c->zeta[1]), nu, dt, c->tw_x, ntiles);
  });
  return 0;
}

// x_update followed by x_zeta2u(fix_mask) in one launch: the new zeta
// stays on chip (the direct-to-LDS path only)
template <typename T>
int fx_update_zeta2u_t(spm_ctx *c, double nu, double dt, int fix_mask) {
  const Grid &g = c->grid;
  SPM_SIZE_SWITCH(g.nx, {
    constexpr int R = kFusedR(S), C = kFusedC(S);
    const int ntiles = (g.ny * g.hp + C - 1) / C;
    hipLaunchKernelGGL((x_update_dma_kernel<T, S, R, C, true>), dim3(ntiles), dim3(C * S / R), 0, c->stream, g,
                       F<T>(c->u[0]), F<T>(c->u[1]), F<T>(c->u[2]), F<T>(c->work[0]), F<T>(c->work[1]),
                       F<T>(c->zeta[0]), F<T>(c->zeta[1]), nu, dt, c->tw_x, F<T>(c->u[0]), F<T>(c->u[1]),
                       F<T>(c->u[2]), c->uk_dc, fix_mask);
  });
  return 0;
}

template <typename T>
int fx_project_t(spm_ctx *c) {
  const Grid &g = c->grid;
  SPM_SIZE_SWITCH(g.nx, {
    constexpr int R = kFusedR(S), C = kFusedC(S);
    const int ntiles = (g.ny * g.hp + C - 1) / C;
    constexpr int per_cu = sizeof(T) == 8 ? SPM_X_PROJECT_BLOCKS_F64 : SPM_X_PROJECT_BLOCKS_F32;
    if constexpr (sizeof(T) == 8 ? SPM_X_DMA_F64 : SPM_X_DMA_F32) {
      hipLaunchKernelGGL((x_project_dma_kernel<T, S, R, C>), dim3(ntiles), dim3(C * S / R), 0, c->stream, g,
                         F<T>(c->f[0]), F<T>(c->f[1]), F<T>(c->f[2]), F<T>(c->zeta[0]), F<T>(c->zeta[1]), c->uk_dc,
                         c->tw_x);
      return 0;
    }
    hipLaunchKernelGGL((x_project_kernel<T, S, R, C, (per_cu > 0)>), dim3(x_grid(c, ntiles, per_cu)), dim3(C * S / R), 0, c->stream, g, F<T>(c->f[0]),
                       F<T>(c->f[1]), F<T>(c->f[2]), F<T>(c->zeta[0]), F<T>(c->zeta[1]), c->uk_dc, c->tw_x, ntiles);
  });
  return 0;
}

#define SPM_PREC_CALL(c, fn, ...) \
  ((c)->prm.precision == SPM_FP32 ? fn<float>(c, ##__VA_ARGS__) : fn<double>(c, ##__VA_ARGS__))

// Fractions of a field the band-limited passes touch: kept (ky, kz)
// columns, kept kx within them, kept kz entries of a row.
struct Band {
  double col, kx, kz;
};
inline Band band(const spm_ctx *c) {
  const Grid &g = c->grid;
  return Band{(double)(2 * g.trn_y - 1) / g.ny * g.trn_z / g.hnz, (double)(2 * g.trn_x - 1) / g.nx,
              (double)g.trn_z / g.hnz};
}

}  // namespace

// Algorithmic bytes below count what each pass must move: kept columns and
// modes only.
int fx_zeta2u(spm_ctx *c, int fix_mask) {
  tbeg(c);
  const int rc = SPM_PREC_CALL(c, fx_zeta2u_t, fix_mask);
  const Band b = band(c);
  tend(c, SPM_K_FFT_X, (2.0 * b.col * b.kx + 3.0 * b.col) * field_bytes(c));
  if (rc) return rc;
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int fz_advect(spm_ctx *c) {
  tbeg(c);
  const int rc = SPM_PREC_CALL(c, fz_advect_t);
  tend(c, SPM_K_FFT_Z_C2R, 8.0 * band(c).kz * field_bytes(c));
  if (rc) return rc;
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int fx_update(spm_ctx *c, double nu, double dt) {
  tbeg(c);
  const int rc = SPM_PREC_CALL(c, fx_update_t, nu, dt);
  const Band b = band(c);
  tend(c, SPM_K_FFT_X, (5.0 * b.col + 4.0 * b.col * b.kx) * field_bytes(c));
  if (rc) return rc;
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int fx_update_zeta2u(spm_ctx *c, double nu, double dt, int fix_mask) {
  const bool fp32 = c->prm.precision == SPM_FP32;
  if (!SPM_X_UPDATE_ZU || !(fp32 ? SPM_X_DMA_F32 : SPM_X_DMA_F64)) {
    const int rc = fx_update(c, nu, dt);
    return rc ? rc : fx_zeta2u(c, fix_mask);
  }
  tbeg(c);
  const int rc = SPM_PREC_CALL(c, fx_update_zeta2u_t, nu, dt, fix_mask);
  const Band b = band(c);
  tend(c, SPM_K_FFT_X, (8.0 * b.col + 4.0 * b.col * b.kx) * field_bytes(c));
  if (rc) return rc;
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int fx_project(spm_ctx *c) {
  tbeg(c);
  const int rc = SPM_PREC_CALL(c, fx_project_t);
  const Band b = band(c);
  tend(c, SPM_K_FFT_X, (3.0 * b.col + 4.0 * b.col * b.kx) * field_bytes(c));
  if (rc) return rc;
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int fz_cross_obl(spm_ctx *c, double *const u[3], double *const w[3]) {
  using T = double;
  const Grid &g = c->grid;
  const T scale = (T)(1.0 / ((double)g.nx * g.ny * g.nz));
  tbeg(c);
  SPM_ZSIZE_SWITCH(g.nz / 2, {
    constexpr int R = kZadvR(S), ROWS = kZadvRows<T>(S);
    hipLaunchKernelGGL((z_cross_kernel<T, S, R, ROWS>), dim3((g.nx * g.ny + ROWS - 1) / ROWS), dim3(ROWS * S / R), 0,
                       c->stream, g, u[0], u[1], u[2], w[0], w[1], w[2], c->tw_z, c->tw_rz, scale);
  });
  // six truncated spectra read, three written, kz < TRN_Z entries
  tend(c, SPM_K_FFT_Z_C2R, 9.0 * ref_count(g) * sizeof(T) * ((double)g.trn_z / g.hnz));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

}  // namespace spm
