// particles.hip -- particle <-> grid coupling and particle dynamics.
//
//   make_phi   Make_phi_u_primitive (make_phi.cxx:221-298) as an
//              owner-computes gather: the grid is cut into cubic tiles,
//              each tile holds the sorted list of particles whose stencil
//              cube reaches it, and each grid point sums its particles in
//              ascending index -- the order in which the reference's
//              particle loop adds into the same point.  No atomics on
//              floating point, so results are run-to-run identical.
//   f_hydro    Calc_f_hydro_correct_precision (md_force.cxx:263-389): one
//              wavefront per particle strides the stencil table and
//              reduces force and torque with cross-lane shuffles in FP64.
//   force      Force (particle_solver.h:97-109): LJ/WCA on a linked-cell
//              list whose cells are at least one cutoff wide (a correct
//              list also for boxes under 3 cells per axis, where the
//              reference's md_force.cxx:78-92 visits a cell twice), plus
//              gravity (md_force.cxx:181-191).
//   integrators MD_solver_position_* / MD_solver_velocity_* and the
//              quaternion updates, one thread per particle, FP64.
#include "phi_core.h"
#include "spm_internal.h"

namespace spm {

namespace {

using phi::kTile;
using phi::PhiAcc;
using phi::distance0;
using phi::phi_fast;

__device__ __forceinline__ double sq(double x) { return x * x; }
__device__ __forceinline__ int nint(double a) { return a >= 0.0 ? (int)(a + .5) : (int)(a - .5); }

// H and Phi, profile.h:29-49.
__device__ __forceinline__ double hfun(double dx, double x) { return x > 0 ? exp(-sq(dx / x)) : 0.0; }
__device__ __forceinline__ double phi_of(const PModel &pm, double x) {
  const double dmy = hfun(pm.dx, pm.radius + pm.hxi - x);
  return dmy / (dmy + hfun(pm.dx, x - pm.radius + pm.hxi));
}

__device__ __forceinline__ double P(const double *pb, int np, int slot, int n) { return pb[(long)slot * np + n]; }
__device__ __forceinline__ double &PR(double *pb, int np, int slot, int n) { return pb[(long)slot * np + n]; }

// ---------------------------------------------------------------------
// anchors and per-tile lists
// ---------------------------------------------------------------------
// Particle_cell (make_phi.h:102-130) and the number of tiles each
// particle's stencil cube [x_int - nh, x_int + nh] reaches.
__device__ __forceinline__ void tile_span(int xint, int nh, int ntile, int &t0, int &cnt) {
  const int lo = xint - nh, hi = xint + nh;
  const int a = (lo >= 0) ? lo / kTile : -((-lo + kTile - 1) / kTile);
  const int b = (hi >= 0) ? hi / kTile : -((-hi + kTile - 1) / kTile);
  cnt = b - a + 1;
  if (cnt > ntile) cnt = ntile;
  t0 = ((a % ntile) + ntile) % ntile;
}

// Thread (n, s) handles slot s = (a, b, c) of the k0 x k1 x k2 box of tiles
// particle n's stencil cube can reach: one atomic per thread rather than a
// loop of up to k0*k1*k2 atomics per particle.  Slot (0, 0, 0) also stores
// the anchor of Particle_cell (make_phi.h:102-130).
__device__ __forceinline__ int tile_slot(int s, const int k[3], const int t0[3], const int cn[3], const int nt[3]) {
  const int a = s / (k[1] * k[2]), b = (s / k[2]) % k[1], cc = s % k[2];
  if (a >= cn[0] || b >= cn[1] || cc >= cn[2]) return -1;
  return ((t0[0] + a) % nt[0] * nt[1] + (t0[1] + b) % nt[1]) * nt[2] + (t0[2] + cc) % nt[2];
}

// Whether a particle with anchor xi and residue res reaches tile ti: some
// point of the tile's box lies within the stencil cube (+-nh) and within
// rcut + sqrt(3) dx of the particle (plus 0.01 dx, far above the rounding
// of the residue form of the distance).  The lists keep a particle only in
// the tiles it reaches, so every per-point sum over a list (phi, up, the
// shell charge and normals, all zero beyond rcut) is that over the cube's
// tiles, and so is phi at an off-grid point taken from the list of the
// tile of the grid point below it (that point lies within sqrt(3) dx).  A
// box whose offsets cross the periodic wrap is kept.
__device__ __forceinline__ bool tile_reached(const Grid &g, const PModel &pm, const int xi[3], const double res[3],
                                             int ti, int ty, int tz) {
  const int tc[3] = {ti / (tz * ty), (ti / tz) % ty, ti % tz};
  const int N[3] = {g.nx, g.ny, g.nz};
  const double reach = pm.radius + pm.hxi + (1.7320508075688772 + 1e-2) * pm.dx;
  double d2 = 0.0;
  bool cross = false, out = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int o0 = tc[d] * kTile - xi[d];
    if (o0 >= N[d] / 2) o0 -= N[d];
    if (o0 < -(N[d] / 2)) o0 += N[d];
    int o1 = o0 + kTile - 1;
    cross = cross || o1 >= N[d] / 2;
    out = out || o1 < -pm.nh || o0 > pm.nh;
    o0 = o0 < -pm.nh ? -pm.nh : o0;
    o1 = o1 > pm.nh ? pm.nh : o1;
    const double r0 = (double)o0 * pm.dx - res[d], r1 = (double)o1 * pm.dx - res[d];
    const double r = r0 > 0.0 ? r0 : (r1 < 0.0 ? -r1 : 0.0);
    d2 += r * r;
  }
  return cross || (!out && d2 < reach * reach);
}

__global__ void anchor_count_kernel(Grid g, PModel pm, const double *__restrict__ pb, int *pxint, double *pres,
                                    int tx, int ty, int tz, int k0, int k1, int k2, int *tile_count) {
  const int per = k0 * k1 * k2;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int n = (int)(t / per), s = (int)(t % per);
  if (n >= pm.np) return;
  const double idx = 1. / pm.dx;
  const int N[3] = {g.nx, g.ny, g.nz};
  int xi[3];
  double res[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double dmy = P(pb, pm.np, P_X + d, n) * idx;
    // positions are kept in [0, L) by PBC; the clamp only keeps indices in
    // range if a non-finite or escaped position ever arrives
    xi[d] = (dmy >= 0.0 && dmy < (double)N[d]) ? (int)dmy : 0;
    res[d] = (dmy - (double)xi[d]) * pm.dx;
    if (s == 0) {
      pxint[3 * n + d] = xi[d];
      pres[3 * n + d] = res[d];
    }
  }
  int t0[3], cn[3];
  const int nt[3] = {tx, ty, tz}, k[3] = {k0, k1, k2};
#pragma unroll
  for (int d = 0; d < 3; ++d) tile_span(xi[d], pm.nh, nt[d], t0[d], cn[d]);
  const int ti = tile_slot(s, k, t0, cn, nt);
  if (ti >= 0 && tile_reached(g, pm, xi, res, ti, ty, tz)) atomicAdd(&tile_count[ti], 1);
}

__global__ void tile_fill_kernel(Grid g, PModel pm, const int *__restrict__ pxint, const double *__restrict__ pres,
                                 int tx, int ty, int tz, int k0, int k1, int k2, int *tile_cursor, int *tile_list) {
  const int per = k0 * k1 * k2;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int n = (int)(t / per), s = (int)(t % per);
  if (n >= pm.np) return;
  int t0[3], cn[3];
  const int nt[3] = {tx, ty, tz}, k[3] = {k0, k1, k2};
  const int xi[3] = {pxint[3 * n], pxint[3 * n + 1], pxint[3 * n + 2]};
  const double res[3] = {pres[3 * n], pres[3 * n + 1], pres[3 * n + 2]};
#pragma unroll
  for (int d = 0; d < 3; ++d) tile_span(xi[d], pm.nh, nt[d], t0[d], cn[d]);
  const int ti = tile_slot(s, k, t0, cn, nt);
  if (ti >= 0 && tile_reached(g, pm, xi, res, ti, ty, tz)) tile_list[atomicAdd(&tile_cursor[ti], 1)] = n;
}

// Exclusive scan of count[0..n) into start[] and cursor[], one block per
// 4096 counts.  Each block first sums all the counts before its chunk (at
// most n ints, read from L2), so blocks never wait on one another; then
// each thread takes 4 consecutive counts, the wavefronts scan their 64
// partial sums with shuffles, and one LDS exchange combines the 16
// wavefronts.  Integer sums: the result does not depend on the order.
__global__ __launch_bounds__(1024) void scan_kernel(const int *__restrict__ count, int *start, int *cursor, int n) {
  __shared__ int wsum[16];
  __shared__ int carry_s;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int base = blockIdx.x * 4096;
  int acc = 0;
  for (int i = t; i < base && i < n; i += 1024) acc += count[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) wsum[w] = acc;
  __syncthreads();
  if (t == 0) {
    int c = 0;
    for (int k = 0; k < 16; ++k) c += wsum[k];
    carry_s = c;
  }
  __syncthreads();
  const int carry = carry_s;
  const int i0 = base + 4 * t;
  int v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = (i0 + q < n) ? count[i0 + q] : 0;
  const int s = v[0] + v[1] + v[2] + v[3];
  int x = s;  // inclusive scan over the wavefront
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  if (w == 0) {
    int y = lane < 16 ? wsum[lane] : 0;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      const int z = __shfl_up(y, off, 64);
      if (lane >= off) y += z;
    }
    if (lane < 16) wsum[lane] = y;  // inclusive over wavefronts
  }
  __syncthreads();
  int run = carry + (w > 0 ? wsum[w - 1] : 0) + x - s;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (i0 + q < n) {
      start[i0 + q] = run;
      cursor[i0 + q] = run;
    }
    run += v[q];
  }
}

inline unsigned scan_blocks(int n) { return (unsigned)((n + 4095) / 4096 > 0 ? (n + 4095) / 4096 : 1); }

// Sort each list segment ascending, 64 segments per 64-thread block.  A
// lane sorts its own segment of up to kSortLane entries in place
// (insertion sort); the wavefront then takes the block's longer segments
// one at a time: up to kSortLds entries are copied to LDS and every entry
// goes straight to its rank (the number of smaller entries, equal ones
// ranked by position, so repeated entries keep their count), a longer one
// is insertion-sorted by one lane.
constexpr int kSortLane = 16;
constexpr int kSortLds = 1024;
__global__ __launch_bounds__(64) void sort_segments_kernel(const int *__restrict__ start, const int *__restrict__ count,
                                                           int *list, int n) {
  const int lane = threadIdx.x;
  const int s = blockIdx.x * 64 + lane;
  const int m = s < n ? count[s] : 0;
  const int s0 = s < n ? start[s] : 0;
  if (m > 1 && m <= kSortLane) {
    int *a = list + s0;
    for (int i = 1; i < m; ++i) {
      const int v = a[i];
      int j = i - 1;
      while (j >= 0 && a[j] > v) {
        a[j + 1] = a[j];
        --j;
      }
      a[j + 1] = v;
    }
  }
  __shared__ int buf[kSortLds];
  unsigned long long big = __ballot(m > kSortLane);
  while (big) {
    const int l = __ffsll((long long)big) - 1;
    big &= big - 1;
    const int mm = __shfl(m, l);
    int *a = list + __shfl(s0, l);
    if (mm <= kSortLds) {
      for (int i = lane; i < mm; i += 64) buf[i] = a[i];
      __syncthreads();
      for (int i = lane; i < mm; i += 64) {
        const int v = buf[i];
        int r = 0;
        for (int j = 0; j < mm; ++j) {
          const int w = buf[j];
          r += (w < v) || (w == v && j < i);
        }
        a[r] = v;
      }
      __syncthreads();  // buf is refilled for the next segment
    } else if (lane == 0) {
      for (int i = 1; i < mm; ++i) {
        const int v = a[i];
        int j = i - 1;
        while (j >= 0 && a[j] > v) {
          a[j + 1] = a[j];
          --j;
        }
        a[j + 1] = v;
      }
    }
  }
}

// The whole tile-list build of a small box in one block: counts in LDS,
// the exclusive scan, the fill with LDS cursors and the per-tile sort, with
// barriers between the phases (the lists equal those of the multi-kernel
// build: each tile's members sorted ascending).  For boxes of at most
// kSmallTiles tiles, where the five launches of the general build each cost
// more than their work.  With more than kSmallSlots (particle, tile) pairs
// the one block's serial share outweighs the launches saved.
constexpr int kSmallTiles = 4096;
constexpr long kSmallSlots = 4096;
__global__ __launch_bounds__(1024) void tiles_small_kernel(Grid g, PModel pm, const double *__restrict__ pb,
                                                           int *pxint, double *pres, int tx, int ty, int tz, int k0,
                                                           int k1, int k2, int ntiles, int *tile_count,
                                                           int *tile_start, int *tile_list) {
  __shared__ int cnt[kSmallTiles], cur[kSmallTiles];
  __shared__ int wsum[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (int i = t; i < ntiles; i += 1024) cnt[i] = 0;
  __syncthreads();
  const int per = k0 * k1 * k2;
  const long slots = (long)pm.np * per;
  const int N[3] = {g.nx, g.ny, g.nz};
  const int nt[3] = {tx, ty, tz}, k[3] = {k0, k1, k2};
  const double idx = 1. / pm.dx;
  for (long q = t; q < slots; q += 1024) {
    const int n = (int)(q / per), sl = (int)(q % per);
    int xi[3], t0[3], cn[3];
    double res[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double dmy = P(pb, pm.np, P_X + d, n) * idx;
      xi[d] = (dmy >= 0.0 && dmy < (double)N[d]) ? (int)dmy : 0;
      res[d] = (dmy - (double)xi[d]) * pm.dx;
      if (sl == 0) {
        pxint[3 * n + d] = xi[d];
        pres[3 * n + d] = res[d];
      }
      tile_span(xi[d], pm.nh, nt[d], t0[d], cn[d]);
    }
    const int ti = tile_slot(sl, k, t0, cn, nt);
    if (ti >= 0 && tile_reached(g, pm, xi, res, ti, ty, tz)) atomicAdd(&cnt[ti], 1);
  }
  __syncthreads();
  // exclusive scan: 4 consecutive counts per thread, wavefront shuffles,
  // one LDS exchange across the 16 wavefronts
  int v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = (4 * t + q < ntiles) ? cnt[4 * t + q] : 0;
  int x = v[0] + v[1] + v[2] + v[3];
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  int before = 0;
  for (int k2w = 0; k2w < w; ++k2w) before += wsum[k2w];
  int run = before + x - (v[0] + v[1] + v[2] + v[3]);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i = 4 * t + q;
    if (i < ntiles) {
      cur[i] = run;
      tile_start[i] = run;
      tile_count[i] = v[q];
    }
    run += v[q];
  }
  __syncthreads();
  for (long q = t; q < slots; q += 1024) {
    const int n = (int)(q / per), sl = (int)(q % per);
    int t0[3], cn[3];
    const int xi[3] = {pxint[3 * n], pxint[3 * n + 1], pxint[3 * n + 2]};
    const double res[3] = {pres[3 * n], pres[3 * n + 1], pres[3 * n + 2]};
#pragma unroll
    for (int d = 0; d < 3; ++d) tile_span(xi[d], pm.nh, nt[d], t0[d], cn[d]);
    const int ti = tile_slot(sl, k, t0, cn, nt);
    if (ti >= 0 && tile_reached(g, pm, xi, res, ti, ty, tz)) tile_list[atomicAdd(&cur[ti], 1)] = n;
  }
  __syncthreads();
  for (int i = t; i < ntiles; i += 1024) {
    int *a = tile_list + (cur[i] - cnt[i]);
    const int m = cnt[i];
    for (int p = 1; p < m; ++p) {
      const int val = a[p];
      int j = p - 1;
      while (j >= 0 && a[j] > val) {
        a[j + 1] = a[j];
        --j;
      }
      a[j + 1] = val;
    }
  }
}

__global__ void izero_kernel(int *a, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = 0;
}

// ---------------------------------------------------------------------
// make_phi: owner-computes over cubic tiles (one point per thread)
// ---------------------------------------------------------------------
// The stencil point of particle n that lands on grid index gi along an axis
// of N points has offset cell = gi - x_int taken into [-N/2, N/2); it
// belongs to the stencil only if |cell| <= nh (the stencil cube).  Any
// point with distance >= a + xi/2 has phi = 0 exactly and adds nothing, and
// every point closer than that lies inside the stencil sphere
// (a + xi/2 + |residue| < a + xi/2 + 1.7321 dx), so the distance test alone
// reproduces the reference's sum over Sekibun_cell.
//   mode 0: phi and up (Make_phi_u_particle, with the clamp)
//   mode 1: phi only (Make_phi_particle, no clamp)
//   mode 2: f = up - phi u (Make_f_particle_dt_nonsole on the fly; phi and
//           up are not stored)
template <typename T>
__global__ __launch_bounds__(kPhiTilePoints) void make_phi_kernel(Grid g, PModel pm, const double *__restrict__ pb,
                                                       const int *__restrict__ pxint, const double *__restrict__ pres,
                                                       const int *__restrict__ tile_start,
                                                       const int *__restrict__ tile_count,
                                                       const int *__restrict__ tile_list, int tx, int ty, int tz,
                                                       T *phi, T *up0, T *up1, T *up2, const T *__restrict__ u0,
                                                       const T *__restrict__ u1, const T *__restrict__ u2, int mode) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);  // z-neighbour tiles share 128-byte lines
  const int tk = tile % tz, tj = (tile / tz) % ty, ti = tile / (tz * ty);
  const int lk = threadIdx.x % kTile, lj = (threadIdx.x / kTile) % kTile, li = threadIdx.x / (kTile * kTile);
  const int gi[3] = {ti * kTile + li, tj * kTile + lj, tk * kTile + lk};
  const bool sw_up = mode != 1;
  using A = typename PhiAcc<T>::type;
  A ph, a0, a1, a2;
  phi::phi_point<T>(g, pm, pb, pxint, pres, tile_list + tile_start[tile], tile_count[tile], gi, sw_up, ph, a0, a1,
                    a2);
  if (sw_up) phi::phi_clamp(ph, a0, a1, a2);
  const long im = ((long)gi[0] * g.ny + gi[1]) * g.nz_ + gi[2];
  const bool pad = gi[2] == g.nz - 1;  // also clear the two padding slots
  if (mode == 2) {
    // f_particle.h:55-57 with this point's phi and up; where phi = 0 (and so
    // up = 0) f = 0 - 0 u = 0 exactly and u is not read
    const T p = (T)ph;
    const bool touch = ph != (A)0;
    up0[im] = touch ? (T)a0 - p * u0[im] : (T)0;
    up1[im] = touch ? (T)a1 - p * u1[im] : (T)0;
    up2[im] = touch ? (T)a2 - p * u2[im] : (T)0;
    if (pad) up0[im + 1] = up0[im + 2] = up1[im + 1] = up1[im + 2] = up2[im + 1] = up2[im + 2] = (T)0;
    return;
  }
  phi[im] = (T)ph;
  if (pad) phi[im + 1] = phi[im + 2] = (T)0;
  if (sw_up) {
    up0[im] = (T)a0;
    up1[im] = (T)a1;
    up2[im] = (T)a2;
    if (pad) up0[im + 1] = up0[im + 2] = up1[im + 1] = up1[im + 2] = up2[im + 1] = up2[im + 2] = (T)0;
  }
}

// ---------------------------------------------------------------------
// electrolyte: charge profile and surface normals, owner-computes over the
// same tiles as make_phi
// ---------------------------------------------------------------------
// DPhi_compact_sin (profile.h:168-181): the surface-shell profile.
__device__ __forceinline__ double dphi_compact_sin(const PModel &pm, double x) {
  const double xi = 2.0 * pm.hxi, hxi = xi * .5, ixi = 1. / xi;
  const double dmy_x = pm.radius - x;
  return fabs(dmy_x) < hxi ? M_PI * .5 * ixi * cos(M_PI * dmy_x * ixi) : 0.;
}

// The tile and offset arithmetic of make_phi_kernel: cell offset of grid
// index gi from particle n's anchor, in [-N/2, N/2), and whether it lies in
// the stencil cube.
__device__ __forceinline__ bool stencil_offset(const int gi[3], const int *xi, const int N[3], int nh, int cell[3]) {
  bool in = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int o = gi[d] - xi[d];
    if (o >= N[d] / 2) o -= N[d];
    if (o < -(N[d] / 2)) o += N[d];
    cell[d] = o;
    in = in && (o <= nh && o >= -nh);
  }
  return in;
}

// Make_phi_qq_particle (operate_electrolyte.cxx:179-227) before its
// normalisation: phi (no clamp) and sum_n q_n DPhi(d_n), particles in
// ascending order at every point as the reference's particle loop adds.
__global__ __launch_bounds__(kPhiTilePoints) void make_phi_qq_kernel(Grid g, PModel pm, const double *__restrict__ pb,
                                                          const int *__restrict__ spec, const int *__restrict__ pxint,
                                                          const int *__restrict__ tile_start,
                                                          const int *__restrict__ tile_count,
                                                          const int *__restrict__ tile_list, int tx, int ty, int tz,
                                                          QTable qt, double *phi, double *surface) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tk = tile % tz, tj = (tile / tz) % ty, ti = tile / (tz * ty);
  const int lk = threadIdx.x % kTile, lj = (threadIdx.x / kTile) % kTile, li = threadIdx.x / (kTile * kTile);
  const int gi[3] = {ti * kTile + li, tj * kTile + lj, tk * kTile + lk};
  const int N[3] = {g.nx, g.ny, g.nz};
  const double x[3] = {gi[0] * pm.dx, gi[1] * pm.dx, gi[2] * pm.dx};
  const int s0 = tile_start[tile], cnt = tile_count[tile];
  const double rcut = pm.radius + pm.hxi;
  double ph = 0.0, sf = 0.0;
  for (int q = 0; q < cnt; ++q) {
    const int n = tile_list[s0 + q];
    int cell[3];
    if (!stencil_offset(gi, pxint + 3 * n, N, pm.nh, cell)) continue;
    const double xp[3] = {P(pb, pm.np, P_X, n), P(pb, pm.np, P_X + 1, n), P(pb, pm.np, P_X + 2, n)};
    double x12[3];
    const double dist = distance0(g, x, xp, x12);
    if (!(dist < rcut)) continue;  // Phi and DPhi vanish
    ph += phi_of(pm, dist);
    sf += qt.q[spec[n]] * dphi_compact_sin(pm, dist);
  }
  const long im = ((long)gi[0] * g.ny + gi[1]) * g.nz_ + gi[2];
  phi[im] = ph;
  surface[im] = sf;
  if (gi[2] == g.nz - 1) phi[im + 1] = phi[im + 2] = surface[im + 1] = surface[im + 2] = 0.0;
}

// Make_surface_normal (make_phi.cxx:17-58): sum of the unit radial vectors
// r/|r| of the particles whose shell |r| - a| < xi/2 holds the point, with
// r = offset * dx - residue (Relative_coord).  Any such point lies in the
// particle's stencil, so the shell test alone selects the reference's points.
__global__ __launch_bounds__(kPhiTilePoints) void surface_normal_kernel(Grid g, PModel pm, const int *__restrict__ pxint,
                                                             const double *__restrict__ pres,
                                                             const int *__restrict__ tile_start,
                                                             const int *__restrict__ tile_count,
                                                             const int *__restrict__ tile_list, int tx, int ty, int tz,
                                                             double *n0, double *n1, double *n2) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tk = tile % tz, tj = (tile / tz) % ty, ti = tile / (tz * ty);
  const int lk = threadIdx.x % kTile, lj = (threadIdx.x / kTile) % kTile, li = threadIdx.x / (kTile * kTile);
  const int gi[3] = {ti * kTile + li, tj * kTile + lj, tk * kTile + lk};
  const int N[3] = {g.nx, g.ny, g.nz};
  const int s0 = tile_start[tile], cnt = tile_count[tile];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int q = 0; q < cnt; ++q) {
    const int n = tile_list[s0 + q];
    int cell[3];
    if (!stencil_offset(gi, pxint + 3 * n, N, pm.nh, cell)) continue;
    double r[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) r[d] = (double)cell[d] * pm.dx - pres[3 * n + d];
    const double dmy_r = sqrt(sq(r[0]) + sq(r[1]) + sq(r[2]));
    if (fabs(dmy_r - pm.radius) < pm.hxi) {
      const double ir = 1. / dmy_r;
      a0 += r[0] * ir;
      a1 += r[1] * ir;
      a2 += r[2] * ir;
    }
  }
  const long im = ((long)gi[0] * g.ny + gi[1]) * g.nz_ + gi[2];
  n0[im] = a0;
  n1[im] = a1;
  n2[im] = a2;
  if (gi[2] == g.nz - 1) n0[im + 1] = n0[im + 2] = n1[im + 1] = n1[im + 2] = n2[im + 1] = n2[im + 2] = 0.0;
}

// copies three consecutive particle slots (f_hydro1 = f_hydro, ...)
__global__ void slot_copy_kernel(double *pb, int np, int dst, int src) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * np) return;
  const int d = t / np, n = t % np;
  PR(pb, np, dst + d, n) = P(pb, np, src + d, n);
}

// ---------------------------------------------------------------------
// Calc_f_hydro_correct_precision: one wavefront per particle
// ---------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// u is gathered in the field precision T; sums run in FP64.
// W wavefronts per particle: wave w of a particle takes stencil points
// w*64 + lane, stepping 64 W; the W wave sums are added in wave order.  W = 1
// (many particles) keeps every wave busy with its own particle; few
// particles (small boxes) take W up to 16 so a stencil of up to 5185 points
// is not walked by one latency-bound wave.  Blocks hold 256 threads, or one
// particle of 64 W threads when that is more.
template <int W>
constexpr int f_hydro_block() { return 64 * W > 256 ? 64 * W : 256; }

template <typename T, int W>
__global__ __launch_bounds__(f_hydro_block<W>()) void f_hydro_kernel(Grid g, PModel pm, double *pb,
                                                                    const int *__restrict__ pxint,
                                                                    const double *__restrict__ pres,
                                                                    const int *__restrict__ stencil, int nst,
                                                                    const T *__restrict__ u0,
                                                                    const T *__restrict__ u1,
                                                                    const T *__restrict__ u2, double dmy) {
  constexpr int BS = f_hydro_block<W>(), PPB = BS / (64 * W);
  __shared__ double part[BS / 64][6];
  const int lane = threadIdx.x & 63;
  const int wv = (threadIdx.x >> 6) % W;
  const int n = blockIdx.x * PPB + (int)(threadIdx.x >> 6) / W;
  const bool active = n < pm.np;
  if (W == 1 && !active) return;
  const int nn = active ? n : 0;
  const int N[3] = {g.nx, g.ny, g.nz};
  const int xi[3] = {pxint[3 * nn], pxint[3 * nn + 1], pxint[3 * nn + 2]};
  const double res[3] = {pres[3 * nn], pres[3 * nn + 1], pres[3 * nn + 2]};
  const double xp[3] = {P(pb, pm.np, P_X, nn), P(pb, pm.np, P_X + 1, nn), P(pb, pm.np, P_X + 2, nn)};
  const double vp[3] = {P(pb, pm.np, P_V, nn), P(pb, pm.np, P_V + 1, nn), P(pb, pm.np, P_V + 2, nn)};
  const double wp[3] = {P(pb, pm.np, P_OMEGA, nn), P(pb, pm.np, P_OMEGA + 1, nn), P(pb, pm.np, P_OMEGA + 2, nn)};
  const double rcut = pm.radius + pm.hxi;
  // single-precision screen as in make_phi: a point whose distance is below
  // rcut has a float distance below rcut + 1e-3 dx, so the FP64 distance
  // and the gather run only near the sphere (phi = 0 adds nothing)
  const float rscreen2 = (float)((rcut + 1e-3 * pm.dx) * (rcut + 1e-3 * pm.dx));
  const float fres[3] = {(float)res[0], (float)res[1], (float)res[2]};
  double fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0;
  for (int s = active ? wv * 64 + lane : nst; s < nst; s += 64 * W) {
    {
      const float f0 = (float)stencil[3 * s] * (float)pm.dx - fres[0];
      const float f1 = (float)stencil[3 * s + 1] * (float)pm.dx - fres[1];
      const float f2 = (float)stencil[3 * s + 2] * (float)pm.dx - fres[2];
      if (!(f0 * f0 + f1 * f1 + f2 * f2 < rscreen2)) continue;
    }
    int rm[3];
    double r[3], x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int cd = stencil[3 * s + d];
      int q = xi[d] + cd;  // in (-N, 2N)
      if (q >= N[d]) q -= N[d];
      if (q < 0) q += N[d];
      rm[d] = q;
      r[d] = (double)cd * pm.dx - res[d];
      x[d] = rm[d] * pm.dx;
    }
    double x12[3];
    const double dist = distance0(g, x, xp, x12);
    if (!(dist < rcut)) continue;  // phi = 0: adds nothing
    const double dphi = phi_fast<double>(pm, dist);
    double vr[3] = {0.0, 0.0, 0.0};
    if (pm.rotation) {
      vr[0] = wp[1] * r[2] - wp[2] * r[1];
      vr[1] = wp[2] * r[0] - wp[0] * r[2];
      vr[2] = wp[0] * r[1] - wp[1] * r[0];
    }
    const long im = ((long)rm[0] * g.ny + rm[1]) * g.nz_ + rm[2];
    const double f0 = ((vp[0] + vr[0]) - (double)u0[im]) * dphi;
    const double f1 = ((vp[1] + vr[1]) - (double)u1[im]) * dphi;
    const double f2 = ((vp[2] + vr[2]) - (double)u2[im]) * dphi;
    fx += f0;
    fy += f1;
    fz += f2;
    tx += (r[1] * f2 - r[2] * f1);
    ty += (r[2] * f0 - r[0] * f2);
    tz += (r[0] * f1 - r[1] * f0);
  }
  fx = wave_sum(fx);
  fy = wave_sum(fy);
  fz = wave_sum(fz);
  tx = wave_sum(tx);
  ty = wave_sum(ty);
  tz = wave_sum(tz);
  if constexpr (W > 1) {
    const int slot = threadIdx.x >> 6;
    if (lane == 0) {
      part[slot][0] = fx; part[slot][1] = fy; part[slot][2] = fz;
      part[slot][3] = tx; part[slot][4] = ty; part[slot][5] = tz;
    }
    __syncthreads();
    if (wv == 0 && lane == 0) {
      fx = fy = fz = tx = ty = tz = 0.0;
      for (int w = 0; w < W; ++w) {
        fx += part[slot + w][0]; fy += part[slot + w][1]; fz += part[slot + w][2];
        tx += part[slot + w][3]; ty += part[slot + w][4]; tz += part[slot + w][5];
      }
    }
  }
  if (active && wv == 0 && lane == 0) {
    PR(pb, pm.np, P_F_HYDRO, n) = (dmy * fx) * 1.0;
    PR(pb, pm.np, P_F_HYDRO + 1, n) = (dmy * fy) * 1.0;
    PR(pb, pm.np, P_F_HYDRO + 2, n) = (dmy * fz) * 1.0;
    for (int d = 0; d < 3; ++d) PR(pb, pm.np, P_F_SLIP + d, n) = 0.0;
    if (pm.rotation) {
      PR(pb, pm.np, P_TORQUE_HYDRO, n) = (dmy * tx) * 1.0;
      PR(pb, pm.np, P_TORQUE_HYDRO + 1, n) = (dmy * ty) * 1.0;
      PR(pb, pm.np, P_TORQUE_HYDRO + 2, n) = (dmy * tz) * 1.0;
      for (int d = 0; d < 3; ++d) PR(pb, pm.np, P_TORQUE_SLIP + d, n) = 0.0;
    }
  }
}

// ---------------------------------------------------------------------
// LJ / WCA on linked cells, gravity
// ---------------------------------------------------------------------
__device__ __forceinline__ int cell_coord(double x, double L, int nc) {
  int c = (int)(x / (L / nc));
  return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

__global__ void cell_count_kernel(Grid g, PModel pm, const double *__restrict__ pb, int cx, int cy, int cz,
                                  int *cell_count, int *pcell) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= pm.np) return;
  const int c = (cell_coord(P(pb, pm.np, P_X, n), g.lx, cx) * cy + cell_coord(P(pb, pm.np, P_X + 1, n), g.ly, cy)) * cz +
                cell_coord(P(pb, pm.np, P_X + 2, n), g.lz, cz);
  pcell[n] = c;
  atomicAdd(&cell_count[c], 1);
}

__global__ void cell_fill_kernel(PModel pm, const int *__restrict__ pcell, int *cell_cursor, int *cell_list) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= pm.np) return;
  const int slot = atomicAdd(&cell_cursor[pcell[n]], 1);
  cell_list[slot] = n;
}

// Lennard_Jones_f, interaction.h:156-181.
__device__ __forceinline__ double lj_f(const PModel &pm, double x) {
  double dmy = pm.lj_dia / x;
  double coeff;
  if (pm.lj_powers == 0) {
    dmy = sq(dmy) * sq(dmy) * sq(dmy);
    coeff = 24. * pm.epsilon;
  } else if (pm.lj_powers == 1) {
    dmy = sq(dmy) * sq(dmy) * sq(dmy) * sq(dmy) * sq(dmy) * sq(dmy);
    coeff = 48. * pm.epsilon;
  } else {
    dmy = sq(dmy) * sq(dmy) * sq(dmy) * sq(dmy) * sq(dmy) * sq(dmy) * sq(dmy) * sq(dmy) * sq(dmy);
    coeff = 72. * pm.epsilon;
  }
  return coeff / sq(x) * (2.0 * sq(dmy) - dmy);
}

// One wavefront per particle: lane l takes the l-th neighbour cell (the
// cells around the particle's own, each visited once even when fewer than
// three span an axis) and sums its pairs in list order; the 64 lane sums are
// combined by a fixed shuffle tree, so the result is run-to-run identical.
// With all_pairs (few particles) the lanes stride over every other particle
// instead of the neighbour cells: the same pairs within the cutoff (the
// nearest image, each once), without the cell lists' five launches.
__global__ __launch_bounds__(256) void force_kernel(Grid g, PModel pm, double *pb, const int *__restrict__ spec,
                                                    const int *__restrict__ pcell, const int *__restrict__ cell_start,
                                                    const int *__restrict__ cell_count,
                                                    const int *__restrict__ cell_list, int cx, int cy, int cz,
                                                    int all_pairs) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= pm.np) return;  // whole wavefronts leave together
  double F[3] = {0.0, 0.0, 0.0};
  if (pm.lj_on && all_pairs) {
    const double xi[3] = {P(pb, pm.np, P_X, i), P(pb, pm.np, P_X + 1, i), P(pb, pm.np, P_X + 2, i)};
    for (int j = lane; j < pm.np; j += 64) {
      if (j == i) continue;
      const double xj[3] = {P(pb, pm.np, P_X, j), P(pb, pm.np, P_X + 1, j), P(pb, pm.np, P_X + 2, j)};
      double rv[3];
      const double r = distance0(g, xi, xj, rv);
      if (r < pm.lj_cutoff) {
        const double ljf = lj_f(pm, r);
        const double cap = 1.79769313486231570815e+308 / r;
        const double dmy = cap < ljf ? cap : ljf;
#pragma unroll
        for (int d = 0; d < 3; ++d) F[d] += dmy * -rv[d];
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) F[d] = wave_sum(F[d]);
  } else if (pm.lj_on) {
    const int ci = pcell[i];
    const int c[3] = {ci / (cy * cz), (ci / cz) % cy, ci % cz};
    const int nc[3] = {cx, cy, cz};
    int lo[3], span[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = nc[d] >= 3 ? -1 : 0;
      span[d] = (nc[d] >= 2 ? 1 : 0) - lo[d] + 1;
    }
    if (lane < span[0] * span[1] * span[2]) {
      const int a = lo[0] + lane / (span[1] * span[2]), b = lo[1] + (lane / span[2]) % span[1],
                e = lo[2] + lane % span[2];
      const double xi[3] = {P(pb, pm.np, P_X, i), P(pb, pm.np, P_X + 1, i), P(pb, pm.np, P_X + 2, i)};
      const int nb = (((c[0] + a + cx) % cx) * cy + (c[1] + b + cy) % cy) * cz + (c[2] + e + cz) % cz;
      const int s0 = cell_start[nb], cnt = cell_count[nb];
      for (int q = 0; q < cnt; ++q) {
        const int j = cell_list[s0 + q];
        if (j == i) continue;
        const double xj[3] = {P(pb, pm.np, P_X, j), P(pb, pm.np, P_X + 1, j), P(pb, pm.np, P_X + 2, j)};
        double rv[3];
        const double r = distance0(g, xi, xj, rv);
        if (r < pm.lj_cutoff) {
          const double ljf = lj_f(pm, r);
          const double cap = 1.79769313486231570815e+308 / r;
          const double dmy = cap < ljf ? cap : ljf;
#pragma unroll
          for (int d = 0; d < 3; ++d) F[d] += dmy * -rv[d];
        }
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) F[d] = wave_sum(F[d]);
  }
  if (lane != 0) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) PR(pb, pm.np, P_FR + d, i) += F[d];
  if (pm.g_on_fluid != 0.0) PR(pb, pm.np, P_FR + pm.g_direction, i) -= pm.g_on_fluid * (pm.mass_ratio[spec[i]] - 1.0);
}

// ---------------------------------------------------------------------
// integrators (particle_solver.cxx, particle_rotation_solver.h)
// ---------------------------------------------------------------------
struct Q {
  double s, v0, v1, v2;
};

__device__ __forceinline__ Q qprod(const Q &a, const Q &b, double alpha) {
  Q r;
  r.s = alpha * (a.s * b.s - (a.v0 * b.v0 + a.v1 * b.v1 + a.v2 * b.v2));
  r.v0 = alpha * (a.s * b.v0 + b.s * a.v0 + a.v1 * b.v2 - a.v2 * b.v1);
  r.v1 = alpha * (a.s * b.v1 + b.s * a.v1 + a.v2 * b.v0 - a.v0 * b.v2);
  r.v2 = alpha * (a.s * b.v2 + b.s * a.v2 + a.v0 * b.v1 - a.v1 * b.v0);
  return r;
}
__device__ __forceinline__ Q qnorm(Q q) {
  double nn = sqrt(q.s * q.s + q.v0 * q.v0 + q.v1 * q.v1 + q.v2 * q.v2);
  nn = 1.0 / nn;
  q.s *= nn;
  q.v0 *= nn;
  q.v1 *= nn;
  q.v2 *= nn;
  return q;
}
__device__ __forceinline__ Q qload(const double *pb, int np, int slot, int n) {
  return Q{P(pb, np, slot, n), P(pb, np, slot + 1, n), P(pb, np, slot + 2, n), P(pb, np, slot + 3, n)};
}
__device__ __forceinline__ void qstore(double *pb, int np, int slot, int n, const Q &q) {
  PR(pb, np, slot, n) = q.s;
  PR(pb, np, slot + 1, n) = q.v0;
  PR(pb, np, slot + 2, n) = q.v1;
  PR(pb, np, slot + 3, n) = q.v2;
}
// rigid_body_rotation SPACE2BODY (rigid_body.cxx:10-34): conj(q) (0,v) q.
__device__ __forceinline__ void space2body(double *out, const double *v, const Q &q) {
  Q qx{0.0, v[0], v[1], v[2]};
  Q qc{q.s, -q.v0, -q.v1, -q.v2};
  qx = qprod(qx, q, 1.0);
  qc = qprod(qc, qx, 1.0);
  out[0] = qc.v0;
  out[1] = qc.v1;
  out[2] = qc.v2;
}

__device__ __forceinline__ void orientation_step(const PModel &pm, double *pb, int n, double dt, double hdt, int euler);

__global__ void position_kernel(Grid g, PModel pm, double *pb, double dt, double hdt, int euler) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= pm.np) return;
  const double L[3] = {g.lx, g.ly, g.lz};
  for (int d = 0; d < 3; ++d) {
    const double x = P(pb, pm.np, P_X + d, n);
    PR(pb, pm.np, P_X_PREV + d, n) = x;
    const double v = P(pb, pm.np, P_V + d, n);
    const double delta = euler ? dt * v : hdt * (3.0 * v - P(pb, pm.np, P_V_OLD + d, n));
    PR(pb, pm.np, P_X_NOPBC + d, n) += delta;
    PR(pb, pm.np, P_X + d, n) = fmod((x + delta) + L[d], L[d]);  // PBC, periodic_boundary.h:9-15
  }
  if (pm.rotation) orientation_step(pm, pb, n, dt, hdt, euler);
}

// MD_solver_orientation_Euler / _AB2 (particle_rotation_solver.h:15-39)
__device__ __forceinline__ void orientation_step(const PModel &pm, double *pb, int n, double dt, double hdt,
                                                 int euler) {
  {
    const Q q = qload(pb, pm.np, P_Q, n);
    Q dq;
    if (euler) {  // MD_solver_orientation_Euler
      const Q qw{0.0, P(pb, pm.np, P_OMEGA, n), P(pb, pm.np, P_OMEGA + 1, n), P(pb, pm.np, P_OMEGA + 2, n)};
      dq = qprod(qw, q, 0.5);
    } else {  // MD_solver_orientation_AB2 (Simo & Wong), step hdt
      const double w[3] = {P(pb, pm.np, P_OMEGA, n), P(pb, pm.np, P_OMEGA + 1, n), P(pb, pm.np, P_OMEGA + 2, n)};
      const double wo[3] = {P(pb, pm.np, P_OMEGA_OLD, n), P(pb, pm.np, P_OMEGA_OLD + 1, n),
                            P(pb, pm.np, P_OMEGA_OLD + 2, n)};
      double wb[3], wbo[3];
      space2body(wb, w, q);
      space2body(wbo, wo, qload(pb, pm.np, P_Q_OLD, n));
      const Q qw{0.0, 3.0 * wb[0] - wbo[0], 3.0 * wb[1] - wbo[1], 3.0 * wb[2] - wbo[2]};
      dq = qprod(q, qw, 0.5);
    }
    const double h = euler ? dt : hdt;
    Q qn{q.s + dq.s * h, q.v0 + dq.v0 * h, q.v1 + dq.v1 * h, q.v2 + dq.v2 * h};
    qstore(pb, pm.np, P_Q_OLD, n, q);
    qstore(pb, pm.np, P_Q, n, qnorm(qn));
  }
}

// MD_solver_velocity_Euler / _AB2_hydro + reset_Forces.
__global__ void velocity_kernel(PModel pm, double *pb, const int *__restrict__ spec, double dt, double hdt,
                                int euler, int obl = 0) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= pm.np) return;
  const int s = spec[n];
  for (int d = 0; d < 3; ++d) {
    const double v = P(pb, pm.np, P_V + d, n), w = P(pb, pm.np, P_OMEGA + d, n);
    const double fh = P(pb, pm.np, P_F_HYDRO + d, n), fr = P(pb, pm.np, P_FR + d, n);
    const double th = P(pb, pm.np, P_TORQUE_HYDRO + d, n);
    PR(pb, pm.np, P_V_OLD + d, n) = v;
    PR(pb, pm.np, P_OMEGA_OLD + d, n) = w;
    if (euler) {
      const double dmy = dt * pm.imass[s], dmy_rot = dt * pm.imoi[s];
      PR(pb, pm.np, P_V + d, n) = v + (dmy * (fh + fr + 0.0));
      PR(pb, pm.np, P_OMEGA + d, n) = w + (dmy_rot * (th + 0.0));
    } else {
      const double dmy = hdt * pm.imass[s], dmy_rot = hdt * pm.imoi[s];
      PR(pb, pm.np, P_V + d, n) = v + dmy * (2.0 * (fh + 0.0) + fr + P(pb, pm.np, P_FR_PREV + d, n));
      PR(pb, pm.np, P_OMEGA + d, n) = w + dmy_rot * 2.0 * (th + 0.0);
    }
    // the sheared solvers keep the impulse of fr for the stress output
    // (momentum_depend_fr, particle_solver.cxx:373,422)
    if (obl)
      PR(pb, pm.np, P_MOMENTUM_DEPEND_FR + d, n) = euler ? dt * fr : hdt * (fr + P(pb, pm.np, P_FR_PREV + d, n));
    // reset_Forces, particle_solver.cxx:11-34
    PR(pb, pm.np, P_FR_PREV + d, n) = fr;
    PR(pb, pm.np, P_FR + d, n) = 0.0;
    PR(pb, pm.np, P_F_HYDRO_PREV + d, n) = fh;
    PR(pb, pm.np, P_F_HYDRO + d, n) = 0.0;
    PR(pb, pm.np, P_F_SLIP_PREV + d, n) = P(pb, pm.np, P_F_SLIP + d, n);
    PR(pb, pm.np, P_F_SLIP + d, n) = 0.0;
    PR(pb, pm.np, P_TORQUE_HYDRO_PREV + d, n) = th;
    PR(pb, pm.np, P_TORQUE_HYDRO + d, n) = 0.0;
    PR(pb, pm.np, P_TORQUE_SLIP_PREV + d, n) = P(pb, pm.np, P_TORQUE_SLIP + d, n);
    PR(pb, pm.np, P_TORQUE_SLIP + d, n) = 0.0;
  }
}

// Add_random_force_thermostat (fluct.cxx:10-36): noise[6n..6n+5] holds the
// particle's three Gauss2 pairs, drawn on the host in the reference order.
__global__ void thermostat_kernel(PModel pm, double *pb, const int *__restrict__ spec, const double *__restrict__ noise,
                                  double nv, double no) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= pm.np) return;
  const int s = spec[n];
  const double imass = pm.imass[s], imoi = pm.imoi[s];
  for (int d = 0; d < 3; ++d) {
    PR(pb, pm.np, P_V + d, n) += noise[6 * n + d] * nv * imass;
    if (pm.rotation) PR(pb, pm.np, P_OMEGA + d, n) += noise[6 * n + d + 3] * no * imoi;
  }
}

// ---------------------------------------------------------------------
// spherical particles under Lees-Edwards shear (sp_3d_ns.cxx:343-373).
// The oblique geometry is written with the reference's expression order and
// without contraction into fused multiply-adds, so that the mesh index a
// floor() picks does not depend on rounding the reference does not have.
// ---------------------------------------------------------------------
#pragma clang fp contract(off)

// Relative_coord_check_stepover_Y (make_phi.h:197-236): stencil point
// `cell` of the particle anchored at xi + res; returns sign (+1 / -1 when
// the point lies across the upper / lower y boundary, where it belongs to
// the particle's image shifted by -sign deg NY cells in x).
__device__ __forceinline__ int rel_coord_obl(const Grid &g, const PModel &pm, const int cell[3], const int xi[3],
                                             const double res[3], double deg, int rm[3], double r[3]) {
  const int NX = g.nx, NY = g.ny, NZ = g.nz;
  double signY = xi[1] + cell[1];
  rm[1] = (xi[1] + cell[1] + NY) % NY;
  signY -= rm[1];
  int sign = (int)signY;
  if (!(sign == 0)) sign = sign / abs(sign);
  double xo = fmod(xi[0] - sign * deg * NY + res[0] / pm.dx + cell[0] + 4 * NX, (double)NX);
  if (xo < 0.0) xo += NX;  // only for NY > 4 NX - stencil (the reference asserts there)
  rm[0] = (int)xo;
  if (rm[0] >= NX) rm[0] = NX - 1;
  const double xres = xo - rm[0];
  rm[2] = (xi[2] + cell[2] + NZ) % NZ;
  r[0] = (double)cell[0] * pm.dx - xres * pm.dx;
  r[1] = (double)cell[1] * pm.dx - res[1];
  r[2] = (double)cell[2] * pm.dx - res[2];
  return sign;
}

// Make_phi_u_primitive_OBL (make_phi.cxx:299-364, clamp :366-384) as an
// owner-computes gather, so every sum has a fixed order and the sheared
// FP64 path is run-to-run identical.
//
// Lists: each particle is entered in the tiles its stencil cube reaches
// (region 0) and, when the cube crosses the upper / lower y face, in the
// tiles of the same cube moved by the image shift -sign deg NY along x
// (regions 1 / 2, one tile wider for the fractional part).  Entering a
// few tiles too many only costs time: the gather keeps a point only where
// the stencil map of the scatter lands on it.
__device__ __forceinline__ int obl_region_tiles(const Grid &g, const PModel &pm, const double *pb, int n,
                                                const int xi[3], double deg, int region, int t0[3], int cn[3]) {
  const int nt[3] = {g.nx / kTile, g.ny / kTile, g.nz / kTile};
  if (region == 1 && xi[1] + pm.nh < g.ny) return 0;
  if (region == 2 && xi[1] - pm.nh >= 0) return 0;
  // one cell of margin along x: the scatter's floor() of the shifted x
  // coordinate can land one past the cube where a residue rounds to 1
  int cx = xi[0], wx = pm.nh + 1;
  if (region > 0) {
    const double sign = region == 1 ? 1.0 : -1.0;
    const double xo = (double)xi[0] - sign * deg * g.ny + (P(pb, pm.np, P_X, n) / pm.dx - (double)xi[0]);
    cx = (int)floor(xo);
  }
  const int lo[3] = {cx - wx, xi[1] - pm.nh, xi[2] - pm.nh};
  const int hi[3] = {cx + wx, xi[1] + pm.nh, xi[2] + pm.nh};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int a = (lo[d] >= 0) ? lo[d] / kTile : -((-lo[d] + kTile - 1) / kTile);
    const int b = (hi[d] >= 0) ? hi[d] / kTile : -((-hi[d] + kTile - 1) / kTile);
    cn[d] = b - a + 1;
    if (cn[d] > nt[d]) cn[d] = nt[d];
    t0[d] = ((a % nt[d]) + nt[d]) % nt[d];
  }
  return 1;
}

// thread (n, region, slot); mode 0 counts, mode 1 fills
__global__ void obl_tiles_kernel(Grid g, PModel pm, const double *__restrict__ pb, const int *__restrict__ pxint,
                                 double deg, int k0, int k1, int k2, int *tile_count, int *tile_cursor, int *tile_list,
                                 int mode) {
  const int per = k0 * k1 * k2;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int n = (int)(t / (3L * per)), region = (int)((t / per) % 3), s = (int)(t % per);
  if (n >= pm.np) return;
  const int xi[3] = {pxint[3 * n], pxint[3 * n + 1], pxint[3 * n + 2]};
  int t0[3], cn[3];
  if (!obl_region_tiles(g, pm, pb, n, xi, deg, region, t0, cn)) return;
  const int nt[3] = {g.nx / kTile, g.ny / kTile, g.nz / kTile}, k[3] = {k0, k1, k2};
  const int ti = tile_slot(s, k, t0, cn, nt);
  if (ti < 0) return;
  if (mode == 0)
    atomicAdd(&tile_count[ti], 1);
  else
    tile_list[atomicAdd(&tile_cursor[ti], 1)] = n;
}

// One block per tile, one point per thread.  For every particle of
// the tile's sorted list (each once) the stencil cells that can map to this
// point are the y offsets c1 = j - xi1 (+NY across the upper face, -NY
// across the lower one) within the cube, and for each the x offsets next
// to the inverse of the sheared x map; a cell counts where the scatter's
// own map (rel_coord_obl) lands on this point and its distance is below
// a + xi/2, so the sums are the scatter's, added in ascending particle
// order as the reference's particle loop adds them.
__global__ __launch_bounds__(kPhiTilePoints) void make_phi_obl_kernel(Grid g, PModel pm, const double *__restrict__ pb,
                                                           const int *__restrict__ pxint,
                                                           const double *__restrict__ pres,
                                                           const int *__restrict__ tile_start,
                                                           const int *__restrict__ tile_count,
                                                           const int *__restrict__ tile_list, double deg,
                                                           const double *__restrict__ sre_p, double *phi,
                                                           double *up0, double *up1, double *up2) {
  const int tz = g.nz / kTile, ty = g.ny / kTile;
  const int tile = blockIdx.x;
  const int tk = tile % tz, tj = (tile / tz) % ty, ti = tile / (tz * ty);
  const int lk = threadIdx.x % kTile, lj = (threadIdx.x / kTile) % kTile, li = threadIdx.x / (kTile * kTile);
  const int gi[3] = {ti * kTile + li, tj * kTile + lj, tk * kTile + lk};
  const int NX = g.nx, NY = g.ny, NZ = g.nz;
  const double shift = *sre_p * g.ly;
  const double rcut = pm.radius + pm.hxi;
  const int *lst = tile_list + tile_start[tile];
  const int cnt = tile_count[tile];
  double ph = 0.0, a0 = 0.0, a1 = 0.0, a2 = 0.0;
  int prev = -1;
  for (int q = 0; q < cnt; ++q) {
    const int n = lst[q];
    if (n == prev) continue;  // a particle entered through two regions
    prev = n;
    const int xi[3] = {pxint[3 * n], pxint[3 * n + 1], pxint[3 * n + 2]};
    int c2 = gi[2] - xi[2];
    if (c2 >= NZ / 2) c2 -= NZ;
    if (c2 < -(NZ / 2)) c2 += NZ;
    if (c2 > pm.nh || c2 < -pm.nh) continue;
    const double res[3] = {pres[3 * n], pres[3 * n + 1], pres[3 * n + 2]};
    // screen: cells farther than a + xi/2 (plus a rounding margin) from the
    // particle add nothing, so only the survivors run the exact map
    const double rz = c2 * pm.dx - res[2];
    const double rc2 = (rcut + 1e-6 * pm.dx) * (rcut + 1e-6 * pm.dx);
    for (int yc = 0; yc < 3; ++yc) {
      const int c1 = gi[1] - xi[1] + (yc == 1 ? NY : (yc == 2 ? -NY : 0));
      if (c1 > pm.nh || c1 < -pm.nh) continue;
      const double ryz = sq(c1 * pm.dx - res[1]) + sq(rz);
      if (!(ryz < rc2)) continue;
      const int sign = yc == 1 ? 1 : (yc == 2 ? -1 : 0);
      const double base = (double)xi[0] - sign * deg * NY + res[0] / pm.dx;
      const double fb = floor(base);
      int c0 = gi[0] - (int)fb;
      c0 = ((c0 % NX) + NX + NX / 2) % NX - NX / 2;
      for (int dc = -1; dc <= 1; ++dc) {
        const int cell[3] = {c0 + dc, c1, c2};
        if (cell[0] > pm.nh || cell[0] < -pm.nh) continue;
        if (!(sq((cell[0] - (base - fb)) * pm.dx) + ryz < rc2)) continue;
        int rm[3];
        double r[3];
        const int sg = rel_coord_obl(g, pm, cell, xi, res, deg, rm, r);
        if (rm[0] != gi[0] || rm[1] != gi[1] || rm[2] != gi[2]) continue;
        double dist = 0.0;
        for (int d = 0; d < 3; ++d) dist += sq(r[d]);
        dist = sqrt(dist);
        if (!(dist < rcut)) continue;  // phi = 0 adds nothing
        const double dphi = phi_fast<double>(pm, dist);
        const double vp[3] = {P(pb, pm.np, P_V, n), P(pb, pm.np, P_V + 1, n), P(pb, pm.np, P_V + 2, n)};
        double vr[3] = {0.0, 0.0, 0.0};
        if (pm.rotation) {
          const double wp[3] = {P(pb, pm.np, P_OMEGA, n), P(pb, pm.np, P_OMEGA + 1, n),
                                P(pb, pm.np, P_OMEGA + 2, n)};
          vr[0] = wp[1] * r[2] - wp[2] * r[1];
          vr[1] = wp[2] * r[0] - wp[0] * r[2];
          vr[2] = wp[0] * r[1] - wp[1] * r[0];
        }
        ph += dphi;
        a0 += (vp[0] - sg * shift + vr[0]) * dphi;
        a1 += (vp[1] + vr[1]) * dphi;
        a2 += (vp[2] + vr[2]) * dphi;
      }
    }
  }
  if (ph > 1.) {  // make_phi.cxx:366-384
    const double inv = 1. / ph;
    a0 *= inv;
    a1 *= inv;
    a2 *= inv;
    ph = 1.;
  }
  const long im = ((long)gi[0] * g.ny + gi[1]) * g.nz_ + gi[2];
  phi[im] = ph;
  up0[im] = a0;
  up1[im] = a1;
  up2[im] = a2;
  if (gi[2] == NZ - 1)
    phi[im + 1] = phi[im + 2] = up0[im + 1] = up0[im + 2] = up1[im + 1] = up1[im + 2] = up2[im + 1] = up2[im + 2] =
        0.0;
}

// Calc_f_hydro_correct_precision_OBL (md_force.cxx:391-470, 497-506) on the
// Cartesian velocity u: one wavefront per particle, sums in a fixed order.
__global__ __launch_bounds__(256) void f_hydro_obl_kernel(Grid g, PModel pm, double *pb, const int *__restrict__ pxint,
                                                          const double *__restrict__ pres,
                                                          const int *__restrict__ stencil, int nst,
                                                          const double *__restrict__ u0,
                                                          const double *__restrict__ u1,
                                                          const double *__restrict__ u2, double deg,
                                                          const double *__restrict__ sre_p, double dmy,
                                                          double *__restrict__ spart) {
  const int lane = threadIdx.x & 63;
  const int n = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (n >= pm.np) return;
  const int xi[3] = {pxint[3 * n], pxint[3 * n + 1], pxint[3 * n + 2]};
  const double res[3] = {pres[3 * n], pres[3 * n + 1], pres[3 * n + 2]};
  const double vp[3] = {P(pb, pm.np, P_V, n), P(pb, pm.np, P_V + 1, n), P(pb, pm.np, P_V + 2, n)};
  const double wp[3] = {P(pb, pm.np, P_OMEGA, n), P(pb, pm.np, P_OMEGA + 1, n), P(pb, pm.np, P_OMEGA + 2, n)};
  const double shift = *sre_p * g.ly;
  const double rcut = pm.radius + pm.hxi;
  double fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0;
  // moments of the Hydro_force fields of md_force.cxx:432-620 (only their
  // sums reach the output, Calc_hydro_stress, fluid_solver.h:195-221):
  // ry is the point's y index unwrapped across the y faces,
  // r_mesh[1] + sign L_particle[1] (md_force.cxx:498)
  double sA = 0, sB = 0, sC = 0, sD = 0, sV = 0, sI = 0;
  for (int s = lane; s < nst; s += 64) {
    const int cell[3] = {stencil[3 * s], stencil[3 * s + 1], stencil[3 * s + 2]};
    int rm[3];
    double r[3];
    const int sign = rel_coord_obl(g, pm, cell, xi, res, deg, rm, r);
    double dist = 0.0;
    for (int d = 0; d < 3; ++d) dist += sq(r[d]);
    dist = sqrt(dist);
    if (!(dist < rcut)) continue;
    const double dphi = phi_fast<double>(pm, dist);
    double vr[3] = {0.0, 0.0, 0.0};
    if (pm.rotation) {
      vr[0] = wp[1] * r[2] - wp[2] * r[1];
      vr[1] = wp[2] * r[0] - wp[0] * r[2];
      vr[2] = wp[0] * r[1] - wp[1] * r[0];
    }
    const long im = ((long)rm[0] * g.ny + rm[1]) * g.nz_ + rm[2];
    const double f0 = ((vp[0] - sign * shift + vr[0]) - u0[im]) * dphi;
    const double f1 = ((vp[1] + vr[1]) - u1[im]) * dphi;
    const double f2 = ((vp[2] + vr[2]) - u2[im]) * dphi;
    fx += f0;
    fy += f1;
    fz += f2;
    tx += (r[1] * f2 - r[2] * f1);
    ty += (r[2] * f0 - r[0] * f2);
    tz += (r[0] * f1 - r[1] * f0);
    if (spart) {
      const double ry = (double)rm[1] + sign * g.ly;
      sA += f0 * ry;
      sB += ry * dphi;
      sC += ry * dphi * r[2];
      sD += ry * dphi * r[1];
      sV += dphi;
      sI += dphi * (dist * dist);
    }
  }
  fx = wave_sum(fx);
  fy = wave_sum(fy);
  fz = wave_sum(fz);
  tx = wave_sum(tx);
  ty = wave_sum(ty);
  tz = wave_sum(tz);
  if (spart) {
    sA = wave_sum(sA);
    sB = wave_sum(sB);
    sC = wave_sum(sC);
    sD = wave_sum(sD);
    sV = wave_sum(sV);
    sI = wave_sum(sI);
  }
  if (lane == 0) {
    if (spart) {  // with the raw sums and the momentum_depend_fr[0] the step began with
      double *o = spart + (long)kStressSlots * n;
      o[0] = sA; o[1] = sB; o[2] = sC; o[3] = sD; o[4] = sV; o[5] = sI;
      o[6] = fx; o[7] = ty; o[8] = tz;
      o[9] = P(pb, pm.np, P_MOMENTUM_DEPEND_FR, n);
    }
    PR(pb, pm.np, P_F_HYDRO, n) = dmy * fx;
    PR(pb, pm.np, P_F_HYDRO + 1, n) = dmy * fy;
    PR(pb, pm.np, P_F_HYDRO + 2, n) = dmy * fz;
    if (pm.rotation) {
      PR(pb, pm.np, P_TORQUE_HYDRO, n) = dmy * tx;
      PR(pb, pm.np, P_TORQUE_HYDRO + 1, n) = dmy * ty;
      PR(pb, pm.np, P_TORQUE_HYDRO + 2, n) = dmy * tz;
    }
  }
}

// PBC_OBL (periodic_boundary.h:22-44) after the step of
// MD_solver_position_*_OBL (particle_solver.cxx:292-346).
__global__ void position_obl_kernel(Grid g, PModel pm, double *pb, double dt, double hdt, int euler, double deg,
                                    const double *__restrict__ sre_p) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= pm.np) return;
  double x[3];
  for (int d = 0; d < 3; ++d) {
    x[d] = P(pb, pm.np, P_X + d, n);
    PR(pb, pm.np, P_X_PREV + d, n) = x[d];
    const double v = P(pb, pm.np, P_V + d, n);
    const double delta = euler ? dt * v : hdt * (3.0 * v - P(pb, pm.np, P_V_OLD + d, n));
    PR(pb, pm.np, P_X_NOPBC + d, n) += delta;
    x[d] += delta;
  }
  double signY = x[1];
  x[1] = fmod(x[1] + g.ly, g.ly);
  signY -= x[1];
  int sign = (int)signY;
  if (!(sign == 0)) sign = sign / abs(sign);
  x[0] -= (double)sign * deg * g.ly;
  x[0] = fmod(x[0] + g.lx, g.lx);
  if (x[0] < 0.0) {  // LY > LX: the shift can exceed one box (the reference asserts)
    x[0] -= floor(x[0] / g.lx) * g.lx;
    if (!(x[0] < g.lx)) x[0] = 0.0;
  }
  x[2] = fmod(x[2] + g.lz, g.lz);
  for (int d = 0; d < 3; ++d) PR(pb, pm.np, P_X + d, n) = x[d];
  const double dvx = -(double)sign * *sre_p * g.ly;
  PR(pb, pm.np, P_V, n) += dvx;
  PR(pb, pm.np, P_V_OLD, n) += dvx;
  if (pm.rotation) orientation_step(pm, pb, n, dt, hdt, euler);
}

// Distance0_OBL (interaction.h:39-65)
__device__ __forceinline__ double distance0_obl(const Grid &g, double deg, const double *x1, const double *x2,
                                                double *x12) {
  double dmy = 0.0;
  double signY = x2[1] - x1[1];
  x12[1] = x2[1] - x1[1];
  x12[1] -= (double)nint(x12[1] / g.ly) * g.ly;
  signY -= x12[1];
  int sign = (int)signY;
  if (!(sign == 0)) sign = sign / abs(sign);
  dmy += sq(x12[1]);
  x12[0] = x2[0] - (x1[0] + (double)sign * deg * g.ly);
  x12[0] -= (double)nint(x12[0] / g.lx) * g.lx;
  dmy += sq(x12[0]);
  x12[2] = x2[2] - x1[2];
  x12[2] -= (double)nint(x12[2] / g.lz) * g.lz;
  dmy += sq(x12[2]);
  return sqrt(dmy);
}

// Force_OBL (particle_solver.h:111-127): the all-pairs LJ of
// md_force.cxx:136-179 with Distance0_OBL, then gravity.  One wavefront
// per particle i; each pair is measured in the reference's orientation
// (from the lower index to the higher), and the lane sums are combined by a
// fixed shuffle tree.
__global__ __launch_bounds__(256) void force_obl_kernel(Grid g, PModel pm, double *pb, const int *__restrict__ spec,
                                                        double deg, double *__restrict__ vir) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= pm.np) return;
  double F[3] = {0.0, 0.0, 0.0};
  double V = 0.0;  // the pair virial (dmy -r_x) r_y of md_force.cxx:174 over pairs (i, j > i)
  if (pm.lj_on) {
    const double xi[3] = {P(pb, pm.np, P_X, i), P(pb, pm.np, P_X + 1, i), P(pb, pm.np, P_X + 2, i)};
    for (int j = lane; j < pm.np; j += 64) {
      if (j == i) continue;
      const double xj[3] = {P(pb, pm.np, P_X, j), P(pb, pm.np, P_X + 1, j), P(pb, pm.np, P_X + 2, j)};
      double rv[3];
      const double r = i < j ? distance0_obl(g, deg, xi, xj, rv) : distance0_obl(g, deg, xj, xi, rv);
      if (r < pm.lj_cutoff) {
        const double ljf = lj_f(pm, r);
        const double cap = 1.79769313486231570815e+308 / r;
        const double dmy = cap < ljf ? cap : ljf;
        const double sgn = i < j ? 1.0 : -1.0;  // fr[n] += dmy1, fr[m] -= dmy1
#pragma unroll
        for (int d = 0; d < 3; ++d) F[d] += sgn * (dmy * -rv[d]);
        if (i < j) V += (dmy * -rv[0]) * rv[1];
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) F[d] = wave_sum(F[d]);
    V = wave_sum(V);
  }
  if (lane != 0) return;
  if (vir) vir[i] = V;
#pragma unroll
  for (int d = 0; d < 3; ++d) PR(pb, pm.np, P_FR + d, i) += F[d];
  if (pm.g_on_fluid != 0.0) PR(pb, pm.np, P_FR + pm.g_direction, i) -= pm.g_on_fluid * (pm.mass_ratio[spec[i]] - 1.0);
}

// Make_f_particle_dt_nonsole (f_particle.h:41-62), U2u_oblique
// (fft_wrapper.h:204-263, already a gather) and Add_f_particle in one pass:
// u(i, j) += the oblique components of f = up - phi ucp interpolated
// between x rows i_oblique and i_oblique + sign.  ucp is the Cartesian
// velocity.  Each source row serves two destination rows, so 56 B of f
// sources and 24 B of u are read and 24 B written per point where a
// particle reaches; elsewhere only phi is read.
__global__ void f_add_obl_kernel(Grid g, const double *__restrict__ up0, const double *__restrict__ up1,
                                 const double *__restrict__ up2, const double *__restrict__ phi,
                                 const double *__restrict__ c0, const double *__restrict__ c1,
                                 const double *__restrict__ c2, double *u0, double *u1, double *u2, double deg) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  const int NX = g.nx, NY = g.ny;
  const int k = (int)(t % g.nz), j = (int)((t / g.nz) % NY), i = (int)(t / ((long)g.nz * NY));
  double sign = j - NY / 2;
  if (!(sign == 0)) sign = sign / fabs(sign);
  int io = (int)(sign * deg * (j - NY / 2)) * sign;
  const double alpha = (deg * (j - NY / 2) - io) * sign;
  const double beta = 1. - alpha;
  // (i + io) mod NX and the next row toward sign, as fmod(... + 4 NX, NX)
  // gives them, in integer arithmetic that stays in range for any NY
  io = ((i + io) % NX + NX) % NX;
  const int iop = ((io + (int)sign) % NX + NX) % NX;
  const long a = ((long)io * NY + j) * g.nz_ + k, b = ((long)iop * NY + j) * g.nz_ + k;
  // phi = 0 means no particle reached the point, so up is zero there too and
  // f vanishes: u stays as it is (adding zeros would not change it)
  if (phi[a] == 0.0 && phi[b] == 0.0) return;
  const long im = ((long)i * NY + j) * g.nz_ + k;
  const double fa0 = up0[a] - phi[a] * c0[a], fb0 = up0[b] - phi[b] * c0[b];
  const double fa1 = up1[a] - phi[a] * c1[a], fb1 = up1[b] - phi[b] * c1[b];
  const double fa2 = up2[a] - phi[a] * c2[a], fb2 = up2[b] - phi[b] * c2[b];
  u0[im] += (beta * fa0 + alpha * fb0) - deg * (beta * fa1 + alpha * fb1);
  u1[im] += beta * fa1 + alpha * fb1;
  u2[im] += beta * fa2 + alpha * fb2;
}

#pragma clang fp contract(on)

inline unsigned blocks(long n, int b) { return (unsigned)((n + b - 1) / b); }

#define RCP(x)              \
  do {                      \
    int rc_ = (x);          \
    if (rc_) return rc_;    \
  } while (0)

}  // namespace

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------
int p_position(spm_ctx *c, int euler) {
  if (c->np == 0) return 0;
  tbeg(c);
  hipLaunchKernelGGL(position_kernel, dim3(blocks(c->np, 128)), dim3(128), 0, c->stream, c->grid, c->pm, c->pbuf,
                     c->prm.dt, 0.5 * c->prm.dt, euler);
  tend(c, SPM_K_PARTICLE, 8.0 * 40 * c->np);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int p_build_tiles(spm_ctx *c) {
  if (c->np == 0) return 0;
  tbeg(c);
  // tiles a stencil cube of 2 nh + 1 points can reach per axis (the same
  // bound sizes tile_list)
  int k[3];
  for (int d = 0; d < 3; ++d) k[d] = std::min(c->tiles[d], (2 * c->pm.nh + kTile) / kTile + 1);
  const long slots = (long)c->np * k[0] * k[1] * k[2];
  if (c->ntiles <= kSmallTiles && slots <= kSmallSlots) {
    hipLaunchKernelGGL(tiles_small_kernel, dim3(1), dim3(1024), 0, c->stream, c->grid, c->pm, c->pbuf, c->pxint,
                       c->pres, c->tiles[0], c->tiles[1], c->tiles[2], k[0], k[1], k[2], c->ntiles, c->tile_count,
                       c->tile_start, c->tile_list);
    tend(c, SPM_K_PARTICLE, 16.0 * c->ntiles);
    SPM_LAUNCH_CHECK(c);
    return 0;
  }
  hipLaunchKernelGGL(izero_kernel, dim3(blocks(c->ntiles, 256)), dim3(256), 0, c->stream, c->tile_count, c->ntiles);
  hipLaunchKernelGGL(anchor_count_kernel, dim3(blocks(slots, 256)), dim3(256), 0, c->stream, c->grid, c->pm, c->pbuf,
                     c->pxint, c->pres, c->tiles[0], c->tiles[1], c->tiles[2], k[0], k[1], k[2], c->tile_count);
  hipLaunchKernelGGL(scan_kernel, dim3(scan_blocks(c->ntiles)), dim3(1024), 0, c->stream, c->tile_count, c->tile_start, c->tile_cursor,
                     c->ntiles);
  hipLaunchKernelGGL(tile_fill_kernel, dim3(blocks(slots, 256)), dim3(256), 0, c->stream, c->grid, c->pm, c->pxint,
                     c->pres, c->tiles[0], c->tiles[1], c->tiles[2], k[0], k[1], k[2], c->tile_cursor, c->tile_list);
  hipLaunchKernelGGL(sort_segments_kernel, dim3(blocks(c->ntiles, 64)), dim3(64), 0, c->stream, c->tile_start,
                     c->tile_count, c->tile_list, c->ntiles);
  tend(c, SPM_K_PARTICLE, 16.0 * c->ntiles);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

static int launch_make_phi(spm_ctx *c, int mode, double *o0, double *o1, double *o2, double *o3) {
  tbeg(c);
  if (c->prm.precision == SPM_FP32) {
    using T = float;
    hipLaunchKernelGGL(make_phi_kernel<T>, dim3(c->ntiles), dim3(kPhiTilePoints), 0,
                       c->stream, c->grid, c->pm, c->pbuf, c->pxint, c->pres, c->tile_start, c->tile_count,
                       c->tile_list, c->tiles[0], c->tiles[1], c->tiles[2], (T *)o0, (T *)o1, (T *)o2, (T *)o3,
                       (const T *)c->u[0], (const T *)c->u[1], (const T *)c->u[2], mode);
  } else {
    using T = double;
    hipLaunchKernelGGL(make_phi_kernel<T>, dim3(c->ntiles), dim3(kPhiTilePoints), 0,
                       c->stream, c->grid, c->pm, c->pbuf, c->pxint, c->pres, c->tile_start, c->tile_count,
                       c->tile_list, c->tiles[0], c->tiles[1], c->tiles[2], (T *)o0, (T *)o1, (T *)o2, (T *)o3,
                       (const T *)c->u[0], (const T *)c->u[1], (const T *)c->u[2], mode);
  }
  const double es = c->prm.precision == SPM_FP32 ? 4.0 : 8.0;
  // mode 0 writes 4 fields, mode 1 one, mode 2 reads 3 and writes 3
  tend(c, SPM_K_MAKE_PHI, (mode == 0 ? 4.0 : mode == 1 ? 1.0 : 6.0) * es * (double)ref_count(c->grid));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int p_make_phi(spm_ctx *c, int sw_up) {
  if (c->np == 0) {
    int rc = k_zero(c, c->phi, c->nfield);
    if (rc) return rc;
    if (sw_up)
      for (int d = 0; d < 3; ++d)
        if ((rc = k_zero(c, c->up[d], c->nfield))) return rc;
    return 0;
  }
  int rc = p_build_tiles(c);
  if (rc) return rc;
  return launch_make_phi(c, sw_up ? 0 : 1, c->phi, c->up[0], c->up[1], c->up[2]);
}

// f = up - phi u with phi, up evaluated on the fly (tile lists must be
// current); with no particles f = -0 * u.
int p_make_f(spm_ctx *c) { return launch_make_phi(c, 2, nullptr, c->f[0], c->f[1], c->f[2]); }

int p_calc_f_hydro(spm_ctx *c) { return p_calc_f_hydro_field(c, c->u); }

template <typename T, int W>
void launch_f_hydro_w(spm_ctx *c, const T *u0, const T *u1, const T *u2, double dmy) {
  constexpr int BS = f_hydro_block<W>(), PPB = BS / (64 * W);
  hipLaunchKernelGGL((f_hydro_kernel<T, W>), dim3((unsigned)((c->np + PPB - 1) / PPB)), dim3(BS), 0, c->stream,
                     c->grid, c->pm, c->pbuf, c->pxint, c->pres, c->stencil, c->n_stencil, u0, u1, u2, dmy);
}

template <typename T>
void launch_f_hydro(spm_ctx *c, int w, const T *u0, const T *u1, const T *u2, double dmy) {
  switch (w) {
    case 1: launch_f_hydro_w<T, 1>(c, u0, u1, u2, dmy); break;
    case 2: launch_f_hydro_w<T, 2>(c, u0, u1, u2, dmy); break;
    case 4: launch_f_hydro_w<T, 4>(c, u0, u1, u2, dmy); break;
    case 8: launch_f_hydro_w<T, 8>(c, u0, u1, u2, dmy); break;
    default: launch_f_hydro_w<T, 16>(c, u0, u1, u2, dmy); break;
  }
}

// Calc_f_hydro_correct_precision(p, fld, jikan) on any real-space field
// (the electrolyte step passes the f field, sp_3d_ns.cxx:97-106).
static int f_hydro_launch(spm_ctx *c, double *const fld[3]);
int p_calc_f_hydro_field(spm_ctx *c, double *const fld[3]) {
  if (c->np == 0) return 0;
  // anchors (x_int, residue) for the current positions
  int rc = p_build_tiles(c);
  if (rc) return rc;
  return f_hydro_launch(c, fld);
}
int p_calc_f_hydro_listed(spm_ctx *c) {
  if (c->np == 0) return 0;
  return f_hydro_launch(c, c->u);
}
static int f_hydro_launch(spm_ctx *c, double *const fld[3]) {
  const double dmy = (-c->pm.dx3 * c->pm.rho) / c->prm.dt;
  tbeg(c);
  // waves per particle: about 8 waves per compute unit in all, 1 to 16 each
  const long want = 8L * c->n_cu;
  int w = 1;
  while (w < 16 && (long)c->np * (2 * w) <= want) w *= 2;
  if (c->prm.precision == SPM_FP32)
    launch_f_hydro<float>(c, w, (const float *)fld[0], (const float *)fld[1], (const float *)fld[2], dmy);
  else
    launch_f_hydro<double>(c, w, (const double *)fld[0], (const double *)fld[1], (const double *)fld[2], dmy);
  // gathers u (3 doubles) at every stencil point of every particle
  tend(c, SPM_K_F_HYDRO, 3.0 * (c->prm.precision == SPM_FP32 ? 4.0 : 8.0) * (double)c->n_stencil * c->np);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

// f_hydro1 = f_hydro, torque_hydro1 = torque_hydro (sp_3d_ns.cxx:101-106)
int p_copy_hydro1(spm_ctx *c) {
  if (c->np == 0) return 0;
  tbeg(c);
  hipLaunchKernelGGL(slot_copy_kernel, dim3(blocks(3L * c->np, 256)), dim3(256), 0, c->stream, c->pbuf, c->np,
                     (int)P_F_HYDRO1, (int)P_F_HYDRO);
  hipLaunchKernelGGL(slot_copy_kernel, dim3(blocks(3L * c->np, 256)), dim3(256), 0, c->stream, c->pbuf, c->np,
                     (int)P_TORQUE_HYDRO1, (int)P_TORQUE_HYDRO);
  tend(c, SPM_K_PARTICLE, 96.0 * c->np);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

// phi (no clamp) and the unnormalised surface charge profile; builds the
// tile lists at the current positions first.
// phi and surface are FP64 fields whatever the context's field type.
int p_make_phi_qq(spm_ctx *c, double *phi, double *surface) {
  if (c->np == 0) {
    RCP(k_zero_f64(c, phi, c->nfield));
    return k_zero_f64(c, surface, c->nfield);
  }
  RCP(p_build_tiles(c));
  tbeg(c);
  hipLaunchKernelGGL(make_phi_qq_kernel, dim3(c->ntiles), dim3(kPhiTilePoints), 0, c->stream, c->grid, c->pm, c->pbuf, c->spec,
                     c->pxint, c->tile_start, c->tile_count, c->tile_list, c->tiles[0], c->tiles[1], c->tiles[2],
                     c->qtab, phi, surface);
  tend(c, SPM_K_MAKE_PHI, 16.0 * ref_count(c->grid));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

// Make_surface_normal into c->surface_normal (tile lists must be current).
int p_surface_normal(spm_ctx *c) {  // FP64 normals whatever the field type
  if (c->np == 0) {
    for (int d = 0; d < 3; ++d) RCP(k_zero_f64(c, c->surface_normal[d], c->nfield));
    return 0;
  }
  tbeg(c);
  hipLaunchKernelGGL(surface_normal_kernel, dim3(c->ntiles), dim3(kPhiTilePoints), 0, c->stream, c->grid, c->pm, c->pxint,
                     c->pres, c->tile_start, c->tile_count, c->tile_list, c->tiles[0], c->tiles[1], c->tiles[2],
                     c->surface_normal[0], c->surface_normal[1], c->surface_normal[2]);
  tend(c, SPM_K_MAKE_PHI, 24.0 * ref_count(c->grid));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

// particles up to which the LJ forces take all pairs rather than cell lists
constexpr int kPairsMax = 256;
int p_force(spm_ctx *c) {
  if (c->np == 0) return 0;
  tbeg(c);
  const int all_pairs = c->np <= kPairsMax;
  if (c->pm.lj_on && !all_pairs) {
    hipLaunchKernelGGL(izero_kernel, dim3(blocks(c->ncells, 256)), dim3(256), 0, c->stream, c->cell_count, c->ncells);
    hipLaunchKernelGGL(cell_count_kernel, dim3(blocks(c->np, 128)), dim3(128), 0, c->stream, c->grid, c->pm, c->pbuf,
                       c->cells[0], c->cells[1], c->cells[2], c->cell_count, c->pcell);
    hipLaunchKernelGGL(scan_kernel, dim3(scan_blocks(c->ncells)), dim3(1024), 0, c->stream, c->cell_count, c->cell_start, c->cell_cursor,
                       c->ncells);
    hipLaunchKernelGGL(cell_fill_kernel, dim3(blocks(c->np, 128)), dim3(128), 0, c->stream, c->pm, c->pcell,
                       c->cell_cursor, c->cell_list);
    hipLaunchKernelGGL(sort_segments_kernel, dim3(blocks(c->ncells, 64)), dim3(64), 0, c->stream, c->cell_start,
                       c->cell_count, c->cell_list, c->ncells);
  }
  hipLaunchKernelGGL(force_kernel, dim3(blocks(c->np, 4)), dim3(256), 0, c->stream, c->grid, c->pm, c->pbuf, c->spec,
                     c->pcell, c->cell_start, c->cell_count, c->cell_list, c->cells[0], c->cells[1], c->cells[2],
                     all_pairs);
  tend(c, SPM_K_PARTICLE, 8.0 * 6 * c->np);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int p_velocity(spm_ctx *c, int euler) {
  if (c->np == 0) return 0;
  int rc = p_force(c);
  if (rc) return rc;
  return p_velocity_forced(c, euler);
}
int p_velocity_forced(spm_ctx *c, int euler) {
  if (c->np == 0) return 0;
  tbeg(c);
  hipLaunchKernelGGL(velocity_kernel, dim3(blocks(c->np, 128)), dim3(128), 0, c->stream, c->pm, c->pbuf, c->spec,
                     c->prm.dt, 0.5 * c->prm.dt, euler);
  tend(c, SPM_K_PARTICLE, 8.0 * 60 * c->np);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

// ---- Lees-Edwards shear (device Shear_rate_eff at sre) ----
int p_position_obl(spm_ctx *c, int euler, const double *sre) {
  if (c->np == 0) return 0;
  tbeg(c);
  hipLaunchKernelGGL(position_obl_kernel, dim3(blocks(c->np, 128)), dim3(128), 0, c->stream, c->grid, c->pm, c->pbuf,
                     c->prm.dt, 0.5 * c->prm.dt, euler, c->le_deg, sre);
  tend(c, SPM_K_PARTICLE, 8.0 * 40 * c->np);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int p_calc_f_hydro_obl(spm_ctx *c, double *const ucp[3], const double *sre) {
  if (c->np == 0) return 0;
  RCP(p_build_tiles(c));  // anchors at the current positions
  const double dmy = (-c->pm.dx3 * c->pm.rho) / c->prm.dt;
  tbeg(c);
  hipLaunchKernelGGL(f_hydro_obl_kernel, dim3(blocks(c->np, 4)), dim3(256), 0, c->stream, c->grid, c->pm, c->pbuf,
                     c->pxint, c->pres, c->stencil, c->n_stencil, ucp[0], ucp[1], ucp[2], c->le_deg, sre, dmy,
                     c->le_stress);
  tend(c, SPM_K_F_HYDRO, 24.0 * (double)c->n_stencil * c->np);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int p_velocity_obl(spm_ctx *c, int euler) {
  if (c->np == 0) return 0;
  tbeg(c);
  hipLaunchKernelGGL(force_obl_kernel, dim3(blocks(c->np, 4)), dim3(256), 0, c->stream, c->grid, c->pm, c->pbuf,
                     c->spec, c->le_deg, c->le_stress ? c->le_stress + (long)kStressSlots * c->np : nullptr);
  tend(c, SPM_K_PARTICLE, 24.0 * (double)c->np * c->np);
  SPM_LAUNCH_CHECK(c);
  tbeg(c);
  hipLaunchKernelGGL(velocity_kernel, dim3(blocks(c->np, 128)), dim3(128), 0, c->stream, c->pm, c->pbuf, c->spec,
                     c->prm.dt, 0.5 * c->prm.dt, euler, 1);
  tend(c, SPM_K_PARTICLE, 8.0 * 60 * c->np);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int obl_tile_span(const spm_ctx *c, int axis) {
  const int width = 2 * c->pm.nh + 1 + (axis == 0 ? 2 : 0);
  return std::min(c->tiles[axis], (width + kTile - 1) / kTile + 1);
}

// phi and up at the anchors p_calc_f_hydro_obl built (positions unchanged
// since), through tile lists that include the sheared images
int p_make_phi_obl(spm_ctx *c, const double *sre) {
  if (c->np == 0) {
    RCP(k_zero(c, c->phi, c->nfield));
    for (int d = 0; d < 3; ++d) RCP(k_zero(c, c->up[d], c->nfield));
    return 0;
  }
  const int k[3] = {obl_tile_span(c, 0), obl_tile_span(c, 1), obl_tile_span(c, 2)};
  const long slots = 3L * c->np * k[0] * k[1] * k[2];
  if (slots > c->tile_list_obl_cap) {
    c->err = "sheared tile lists exceed their allocation";
    return SPM_ERR_INVALID;
  }
  tbeg(c);
  hipLaunchKernelGGL(izero_kernel, dim3(blocks(c->ntiles, 256)), dim3(256), 0, c->stream, c->tile_count_obl,
                     c->ntiles);
  hipLaunchKernelGGL(obl_tiles_kernel, dim3(blocks(slots, 256)), dim3(256), 0, c->stream, c->grid, c->pm, c->pbuf,
                     c->pxint, c->le_deg, k[0], k[1], k[2], c->tile_count_obl, c->tile_cursor_obl, c->tile_list_obl, 0);
  hipLaunchKernelGGL(scan_kernel, dim3(scan_blocks(c->ntiles)), dim3(1024), 0, c->stream, c->tile_count_obl,
                     c->tile_start_obl, c->tile_cursor_obl, c->ntiles);
  hipLaunchKernelGGL(obl_tiles_kernel, dim3(blocks(slots, 256)), dim3(256), 0, c->stream, c->grid, c->pm, c->pbuf,
                     c->pxint, c->le_deg, k[0], k[1], k[2], c->tile_count_obl, c->tile_cursor_obl, c->tile_list_obl, 1);
  hipLaunchKernelGGL(sort_segments_kernel, dim3(blocks(c->ntiles, 64)), dim3(64), 0, c->stream, c->tile_start_obl,
                     c->tile_count_obl, c->tile_list_obl, c->ntiles);
  tend(c, SPM_K_PARTICLE, 16.0 * c->ntiles);
  SPM_LAUNCH_CHECK(c);
  tbeg(c);
  hipLaunchKernelGGL(make_phi_obl_kernel, dim3(c->ntiles), dim3(kPhiTilePoints), 0, c->stream, c->grid, c->pm, c->pbuf,
                     c->pxint, c->pres, c->tile_start_obl, c->tile_count_obl, c->tile_list_obl, c->le_deg, sre, c->phi,
                     c->up[0], c->up[1], c->up[2]);
  tend(c, SPM_K_MAKE_PHI, 4.0 * 8.0 * (double)ref_count(c->grid));
  SPM_LAUNCH_CHECK(c);
  return 0;
}

// u += U2u_oblique(up - phi ucp)
int p_f_add_obl(spm_ctx *c, double *const ucp[3]) {
  const long npts = (long)c->grid.nx * c->grid.ny * c->grid.nz;
  tbeg(c);
  hipLaunchKernelGGL(f_add_obl_kernel, dim3(blocks(npts, 256)), dim3(256), 0, c->stream, c->grid, c->up[0], c->up[1],
                     c->up[2], c->phi, ucp[0], ucp[1], ucp[2], c->u[0], c->u[1], c->u[2], c->le_deg);
  tend(c, SPM_K_REALSPACE, 8.0 * npts);  // phi only: the rest depends on where particles reach
  SPM_LAUNCH_CHECK(c);
  return 0;
}

// One block: read the counter, copy the 6 np normals of slot
// counter mod nslots, advance the counter.  The pool is a ring the host
// refills half by half while the other half is read.
__global__ void noise_pool_kernel(const double *__restrict__ pool, int *counter, int nslots, int n6,
                                  double *__restrict__ dst) {
  __shared__ int slot;
  if (threadIdx.x == 0) slot = *counter;
  __syncthreads();
  const long base = (long)(slot % nslots) * n6;
  for (int i = threadIdx.x; i < n6; i += blockDim.x) dst[i] = pool[base + i];
  __syncthreads();
  if (threadIdx.x == 0) *counter = slot + 1;
}

int p_noise_from_pool(spm_ctx *c) {
  if (c->np == 0) return 0;
  hipLaunchKernelGGL(noise_pool_kernel, dim3(1), dim3(256), 0, c->stream, c->pool_dev, c->pool_counter, c->pool_steps,
                     6 * c->np, c->noise_dev);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

int p_thermostat_apply(spm_ctx *c, const double *noise_dev) {
  if (c->np == 0) return 0;
  tbeg(c);
  hipLaunchKernelGGL(thermostat_kernel, dim3(blocks(c->np, 128)), dim3(128), 0, c->stream, c->pm, c->pbuf, c->spec,
                     noise_dev, c->noise_v, c->noise_o);
  tend(c, SPM_K_PARTICLE, 8.0 * 18 * c->np);
  SPM_LAUNCH_CHECK(c);
  return 0;
}

}  // namespace spm
