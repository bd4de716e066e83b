// phi_core.h -- the per-point sums of Make_phi_u_primitive (make_phi.cxx:
// 221-298) over a tile's particle list, shared by the make_phi kernels
// (particles.hip) and the z pass that forms f = up - phi u on the fly
// (fused.hip), so that both produce the same bits.
#pragma once

#include "spm_internal.h"

namespace spm {
namespace phi {

constexpr int kTile = kPhiTile;  // cubic tiles of grid points carry the particle lists

__device__ __forceinline__ int nint(double a) { return a >= 0.0 ? (int)(a + .5) : (int)(a - .5); }

__device__ __forceinline__ double P(const double *pb, int np, int slot, int n) { return pb[(long)slot * np + n]; }

// Phi (profile.h:29-49) with one exponential: H1 / (H1 + H2) = 1 / (1 + H2 / H1),
// H2 / H1 = exp(dx^2/x1^2 - dx^2/x2^2) with x1 = a + xi/2 - d,
// x2 = d - a + xi/2; H1 = 0 (x1 <= 0) gives 0 and H2 = 0 gives 1 as in the
// two-exponential form.  Equal to it up to rounding; A is the arithmetic
// type (FP64, or FP32 for the FP32 field path).
template <typename A>
__device__ __forceinline__ A phi_fast(const PModel &pm, A d) {
  const A x1 = (A)(pm.radius + pm.hxi) - d, x2 = d - (A)(pm.radius - pm.hxi);
  if (!(x1 > (A)0)) return (A)0;
  if (!(x2 > (A)0)) return (A)1;
  const A dx2 = (A)(pm.dx * pm.dx);
  // dx^2/x1^2 - dx^2/x2^2 over one division
  const A x12 = x1 * x2;
  return (A)1 / ((A)1 + exp(dx2 * ((x2 - x1) * (x2 + x1)) / (x12 * x12)));
}

// accumulation type of make_phi: FP64, or FP32 for FP32 fields
template <typename T>
struct PhiAcc {
  using type = double;
};
template <>
struct PhiAcc<float> {
  using type = float;
};

// Distance0 (interaction.h:110-123) from grid point x1 to particle x2.
// x12/L is taken as x12 * (1/L): the nearest-image count differs only when
// x12/L is within one rounding of a half-integer (a pair half a box apart).
__device__ __forceinline__ double distance0(const Grid &g, const double *x1, const double *x2, double *x12) {
  const double L[3] = {g.lx, g.ly, g.lz};
  const double iL[3] = {1.0 / g.lx, 1.0 / g.ly, 1.0 / g.lz};
  double dmy = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    x12[d] = x2[d] - x1[d];
    x12[d] -= (double)nint(x12[d] * iL[d]) * L[d];
    dmy += x12[d] * x12[d];
  }
  return sqrt(dmy);
}

// phi and the unclamped up sums of grid point gi over the particles
// list[0..cnt) in ascending index, as the reference's particle loop adds
// into that point.  The list is the same for every thread of a wavefront,
// so the compiler reads the particle members with scalar loads.  FP32
// fields take phi and its sums in FP32 (the distance itself in FP64).
template <typename T>
__device__ __forceinline__ void phi_point(const Grid &g, const PModel &pm, const double *__restrict__ pb,
                                          const int *__restrict__ pxint, const double *__restrict__ pres,
                                          const int *__restrict__ list, int cnt, const int gi[3], bool sw_up,
                                          typename PhiAcc<T>::type &ph, typename PhiAcc<T>::type &a0,
                                          typename PhiAcc<T>::type &a1, typename PhiAcc<T>::type &a2) {
  using A = typename PhiAcc<T>::type;
  const int N[3] = {g.nx, g.ny, g.nz};
  const double x[3] = {gi[0] * pm.dx, gi[1] * pm.dx, gi[2] * pm.dx};
  const double rcut = pm.radius + pm.hxi;
  // single-precision screen: a point whose distance is below rcut has a
  // float distance below rcut + 1e-3 dx (the rounding of |r| < 20 dx is
  // ~1e-6 dx), so the double-precision distance runs only near the sphere
  const float rscreen2 = (float)((rcut + 1e-3 * pm.dx) * (rcut + 1e-3 * pm.dx));
  ph = a0 = a1 = a2 = 0;
  for (int q = 0; q < cnt; ++q) {
    const int n = list[q];
    int cell[3];
    bool in = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      int o = gi[d] - pxint[3 * n + d];  // |o| < N
      if (o >= N[d] / 2) o -= N[d];
      if (o < -(N[d] / 2)) o += N[d];
      cell[d] = o;
      in = in && (o <= pm.nh && o >= -pm.nh);
    }
    if (!in) continue;
    {
      const float f0 = (float)cell[0] * (float)pm.dx - (float)pres[3 * n];
      const float f1 = (float)cell[1] * (float)pm.dx - (float)pres[3 * n + 1];
      const float f2 = (float)cell[2] * (float)pm.dx - (float)pres[3 * n + 2];
      if (!(f0 * f0 + f1 * f1 + f2 * f2 < rscreen2)) continue;
    }
    double r[3];  // Relative_coord: offset * dx - residue
#pragma unroll
    for (int d = 0; d < 3; ++d) r[d] = (double)cell[d] * pm.dx - pres[3 * n + d];
    double dist;
    if constexpr (sizeof(T) == 8) {  // Distance0 as the reference takes it
      const double xp[3] = {P(pb, pm.np, P_X, n), P(pb, pm.np, P_X + 1, n), P(pb, pm.np, P_X + 2, n)};
      double x12[3];
      dist = distance0(g, x, xp, x12);
    } else {
      dist = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    }
    if (!(dist < rcut)) continue;
    const A dphi = phi_fast<A>(pm, (A)dist);
    ph += dphi;
    if (sw_up) {
      double vr[3] = {0.0, 0.0, 0.0};
      if (pm.rotation) {
        const double w0 = P(pb, pm.np, P_OMEGA, n), w1 = P(pb, pm.np, P_OMEGA + 1, n), w2 = P(pb, pm.np, P_OMEGA + 2, n);
        vr[0] = w1 * r[2] - w2 * r[1];
        vr[1] = w2 * r[0] - w0 * r[2];
        vr[2] = w0 * r[1] - w1 * r[0];
      }
      a0 += (A)(P(pb, pm.np, P_V, n) + vr[0]) * dphi;
      a1 += (A)(P(pb, pm.np, P_V + 1, n) + vr[1]) * dphi;
      a2 += (A)(P(pb, pm.np, P_V + 2, n) + vr[2]) * dphi;
    }
  }
}

// the clamp of make_phi.cxx:280-297: where phi > 1, up is divided by phi
// and phi set to 1
template <typename A>
__device__ __forceinline__ void phi_clamp(A &ph, A &a0, A &a1, A &a2) {
  if (ph > (A)1) {
    const A inv = (A)1 / ph;
    a0 *= inv;
    a1 *= inv;
    a2 *= inv;
    ph = (A)1;
  }
}

}  // namespace phi
}  // namespace spm
