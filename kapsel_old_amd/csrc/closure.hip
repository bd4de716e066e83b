// closure.hip -- the learned force/torque closure of the coarse-grid step
// (include/spm_closure.h): the MLP's forward and backward kernels on the
// matrix cores (v_mfma_f32_16x16x16_bf16: BF16 operands, FP32
// accumulation), Adam, the per-particle inputs, and the hook that adds the
// correction to Calc_f_hydro_correct_precision's force and torque
// (md_force.cxx:263-389).
//
// Layout.  The forward pass takes tiles of 16 samples (particles) per
// block, its four wavefronts splitting the output columns; the backward pass
// tiles of 64, one 16-row slice per wavefront.  Activations of a tile sit in
// LDS as [sample][128] BF16; a layer is a set of 16 x 16 output tiles over
// 8 k-steps of 16.  Weights are kept in FP32 (master) and
// twice in BF16: [out][in] for the forward product (four consecutive k of
// one output column per lane) and [in][out] for the backward product
// dH = dZ W^T.  Weight gradients dW = H^T dZ are products over the tile's
// 64 samples, written per tile and summed over tiles in a fixed order by
// the Adam kernel, so training is run-to-run deterministic.
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/spm_closure.h"
#include "phi_core.h"
#include "spm_internal.h"

namespace spm {
namespace mlp {

constexpr int IN = SPM_CLOSURE_IN, OUT = SPM_CLOSURE_OUT;
constexpr int W = 128;     // padded input and hidden width
constexpr int OUTP = 16;   // padded output width
constexpr int TILE = 64;   // samples per block
constexpr int THREADS = 256;
constexpr int OFF_W[4] = {0, W * W + W, 2 * (W * W + W), 3 * (W * W + W)};
constexpr int OFF_B[4] = {W * W, 2 * W * W + W, 3 * W * W + 2 * W, 3 * (W * W + W) + W * OUTP};
constexpr int NOUT[4] = {W, W, W, OUTP};  // columns of each weight matrix
constexpr int NPARAM = 3 * (W * W + W) + W * OUTP + OUTP;
static_assert(NPARAM == SPM_CLOSURE_NPARAM, "parameter count");

typedef unsigned short bf16_t;
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even (NaN kept quiet)
__device__ __forceinline__ bf16_t f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }

// GELU with the error function, and its derivative
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float dgelu(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__device__ __forceinline__ bf16x4 ld_bf4(const bf16_t *p) { return *reinterpret_cast<const bf16x4 *>(p); }
__device__ __forceinline__ bf16x4 pack_bf4(float a, float b, float c, float d) {
  bf16x4 v;
  v[0] = (short)f2bf(a);
  v[1] = (short)f2bf(b);
  v[2] = (short)f2bf(c);
  v[3] = (short)f2bf(d);
  return v;
}

struct Net {
  const float *p;       // FP32 master parameters
  const bf16_t *wt[4];  // [out][in] per layer
  const bf16_t *wb[4];  // [in][out] per layer
  const float *shift, *scale, *oscale;
};

struct Work {
  bf16_t *h[4];  // h[0]: normalised input, h[1..3]: activations, [n][128]
  float *z[3];   // pre-activations of layers 1..3, [n][128]
  float *dy;     // output gradient [n][16]
  float *loss;   // per-block sums of squared errors
};

// acc[nt] (output columns nt*16 .. +16, rows row0 .. +16 of the tile) =
// act[rows][0..128) x B, with B[k][n] = wt[n][k] (row stride 128)
template <int NT>
__device__ __forceinline__ void mfma_rows(const bf16_t *act, int row0, const bf16_t *__restrict__ wt, f32x4 *acc,
                                          int lane) {
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < W / 16; ++kk) {
    const int k = kk * 16 + 4 * (lane >> 4);
    const bf16x4 a = ld_bf4(act + (row0 + (lane & 15)) * W + k);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const bf16x4 b = ld_bf4(wt + (nt * 16 + (lane & 15)) * W + k);
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
    }
  }
}

// Forward pass of a tile of 16 samples: the four wavefronts split each
// layer's 128 output columns (two 16 x 16 tiles each; the last layer's
// single tile by wavefront 0), so a small batch still spreads over many
// blocks and each wavefront's chain of matrix instructions stays short.
// With `target`, also the output gradient of the mean squared error
// (normalised units) and the tile's sum of squared errors; with wk.z, the
// activations a backward pass needs.
constexpr int FTILE = 16;
__global__ __launch_bounds__(THREADS) void mlp_fwd_kernel(Net net, const float *__restrict__ x, int n, float *y,
                                                          const float *__restrict__ target, Work wk) {
  __shared__ bf16_t act[2][FTILE * W];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const long s0 = (long)blockIdx.x * FTILE;
  for (int e = tid; e < FTILE * W; e += THREADS) {
    const int r = e / W, col = e % W;
    const long s = s0 + r;
    float v = 0.f;
    if (s < n && col < IN) v = (x[s * IN + col] - net.shift[col]) * net.scale[col];
    const bf16_t h = f2bf(v);
    act[0][e] = h;
    if (wk.h[0] && s < n) wk.h[0][s * W + col] = h;
  }
  __syncthreads();
  int cur = 0;
  for (int l = 0; l < 3; ++l) {
    f32x4 acc[2];
    mfma_rows<2>(act[cur], 0, net.wt[l] + 32 * wave * W, acc, lane);
    const float *b = net.p + OFF_B[l];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      const int o = 32 * wave + nt * 16 + (lane & 15);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = 4 * (lane >> 4) + j;
        const long s = s0 + r;
        const float z = acc[nt][j] + b[o];
        const bf16_t h = f2bf(gelu(z));
        act[cur ^ 1][r * W + o] = h;
        if (wk.z[l] && s < n) {
          wk.z[l][s * W + o] = z;
          wk.h[l + 1][s * W + o] = h;
        }
      }
    }
    __syncthreads();
    cur ^= 1;
  }
  float sse = 0.f;
  if (wave == 0) {
    f32x4 acc4[1];
    mfma_rows<1>(act[cur], 0, net.wt[3], acc4, lane);
    const int o = lane & 15;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long s = s0 + 4 * (lane >> 4) + j;
      if (s >= n) continue;
      const float yn = acc4[0][j] + net.p[OFF_B[3] + o];
      if (o < OUT) {
        if (y) y[s * OUT + o] = yn * net.oscale[o];
        if (target) {
          const float d = yn - target[s * OUT + o] / net.oscale[o];
          wk.dy[s * OUTP + o] = 2.f * d / ((float)n * OUT);
          sse += d * d;
        }
      } else if (target) {
        wk.dy[s * OUTP + o] = 0.f;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sse += __shfl_xor(sse, off, 64);
    if (target && lane == 0) wk.loss[blockIdx.x] = sse;
  }
}

// Inference forward of 64 samples per block (no training outputs): each
// layer's BF16 weights are staged once in LDS and shared by the four
// wavefronts, each wavefront takes 16 samples and all output columns, so a
// block reads the weights once per 64 samples instead of once per 16.
// Rows of the LDS arrays are padded to 132 entries (264 bytes), so the 16
// rows a matrix operand reads fall in different banks.  Every output is the
// same chain of matrix instructions in the same k order as mlp_fwd_kernel's,
// so the results are identical.
constexpr int BTILE = 64, LDW = W + 4;
static_assert(IN % 2 == 0, "inputs staged as pairs");
__global__ __launch_bounds__(THREADS) void mlp_infer_kernel(Net net, const float *__restrict__ x, int n,
                                                            float *__restrict__ y) {
  // one activation buffer: a wavefront reads its 16 rows for a layer's
  // products before it writes that layer's outputs over them, and no other
  // wavefront touches them
  __shared__ bf16_t act[BTILE * LDW];
  __shared__ bf16_t wsh[W * LDW];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const long s0 = (long)blockIdx.x * BTILE;
  {
    // inputs as column pairs: thread t always takes pair t % 64, rows
    // t / 64 + 4 i, so its shift and scale are loaded once
    const int cp = tid % 64, c0 = 2 * cp;
    const bool in = c0 < IN;
    const float sh0 = in ? net.shift[c0] : 0.f, sh1 = in ? net.shift[c0 + 1] : 0.f;
    const float sc0 = in ? net.scale[c0] : 0.f, sc1 = in ? net.scale[c0 + 1] : 0.f;
    float2 v[BTILE / 4];
#pragma unroll
    for (int i = 0; i < BTILE / 4; ++i) {
      const long s = s0 + tid / 64 + 4 * i;
      v[i] = (in && s < n) ? *reinterpret_cast<const float2 *>(x + s * IN + c0) : float2{0.f, 0.f};
    }
#pragma unroll
    for (int i = 0; i < BTILE / 4; ++i) {
      const int r = tid / 64 + 4 * i;
      const long s = s0 + r;
      const bool ok = in && s < n;
      act[r * LDW + c0] = f2bf(ok ? (v[i].x - sh0) * sc0 : 0.f);
      act[r * LDW + c0 + 1] = f2bf(ok ? (v[i].y - sh1) * sc1 : 0.f);
    }
  }
  const int row0 = 16 * wave;
  for (int l = 0; l < 4; ++l) {
    const int no = NOUT[l];
    // stage W_l ([out][in], 8 bf16 per 16-byte load) after every wavefront
    // has finished reading the previous layer's weights
    __syncthreads();
#pragma unroll
    for (int i = 0; i < (W * W / 8) / THREADS; ++i) {
      const int e = tid + i * THREADS;
      if (e < no * (W / 8)) {
        const int o = e / (W / 8), k8 = (e % (W / 8)) * 8;
        const uint4 v = *reinterpret_cast<const uint4 *>(net.wt[l] + o * W + k8);
        uint2 *d = reinterpret_cast<uint2 *>(wsh + o * LDW + k8);
        d[0] = uint2{v.x, v.y};
        d[1] = uint2{v.z, v.w};
      }
    }
    __syncthreads();
    const float *b = net.p + OFF_B[l];
    if (l < 3) {
      f32x4 acc[8];
#pragma unroll
      for (int nt = 0; nt < 8; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kk = 0; kk < W / 16; ++kk) {
        const int k = kk * 16 + 4 * (lane >> 4);
        const bf16x4 a = ld_bf4(act + (row0 + (lane & 15)) * LDW + k);
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
          const bf16x4 bw = ld_bf4(wsh + (nt * 16 + (lane & 15)) * LDW + k);
          acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bw, acc[nt], 0, 0, 0);
        }
      }
#pragma unroll
      for (int nt = 0; nt < 8; ++nt) {
        const int o = nt * 16 + (lane & 15);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int r = row0 + 4 * (lane >> 4) + j;
          act[r * LDW + o] = f2bf(gelu(acc[nt][j] + b[o]));
        }
      }
    } else {
      f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kk = 0; kk < W / 16; ++kk) {
        const int k = kk * 16 + 4 * (lane >> 4);
        const bf16x4 a = ld_bf4(act + (row0 + (lane & 15)) * LDW + k);
        const bf16x4 bw = ld_bf4(wsh + (lane & 15) * LDW + k);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bw, acc, 0, 0, 0);
      }
      const int o = lane & 15;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const long s = s0 + row0 + 4 * (lane >> 4) + j;
        if (s < n && o < OUT) y[s * OUT + o] = (acc[j] + b[o]) * net.oscale[o];
      }
    }
  }
}

// Backward pass of a tile: per-tile weight and bias gradients into
// partial[blockIdx.x][NPARAM].
__global__ __launch_bounds__(THREADS) void mlp_bwd_kernel(Net net, Work wk, int n, float *partial) {
  __shared__ float dz[2][TILE * W];  // dZ of the current layer, then of the one below
  __shared__ bf16_t hl[TILE * W];    // activation feeding the current layer
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const long s0 = (long)blockIdx.x * TILE;
  float *part = partial + (long)blockIdx.x * NPARAM;
  for (int e = tid; e < TILE * W; e += THREADS) {
    const int r = e / W, o = e % W;
    const long s = s0 + r;
    dz[0][e] = (s < n && o < OUTP) ? wk.dy[s * OUTP + o] : 0.f;
  }
  int cur = 0;
  for (int l = 3; l >= 0; --l) {
    const int no = NOUT[l];
    for (int e = tid; e < TILE * W; e += THREADS) {
      const long s = s0 + e / W;
      hl[e] = s < n ? wk.h[l][s * W + e % W] : (bf16_t)0;
    }
    __syncthreads();
    // bias gradient: column sums of dZ (fixed order)
    if (tid < no) {
      float sb = 0.f;
      for (int r = 0; r < TILE; ++r) sb += dz[cur][r * W + tid];
      part[OFF_B[l] + tid] = sb;
    }
    // dW[i][o] = sum_s H[s][i] dZ[s][o]: wave w takes inputs i in [32w, 32w+32)
    {
      const int ntn = no / 16;
      f32x4 acc[2][8];
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kk = 0; kk < TILE / 16; ++kk) {
        const int sk = kk * 16 + 4 * (lane >> 4);
        bf16x4 a[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
          const int i = 32 * wave + mt * 16 + (lane & 15);
          bf16x4 v;
#pragma unroll
          for (int j = 0; j < 4; ++j) v[j] = (short)hl[(sk + j) * W + i];
          a[mt] = v;
        }
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
          if (nt >= ntn) break;
          const int o = nt * 16 + (lane & 15);
          const bf16x4 b = pack_bf4(dz[cur][(sk + 0) * W + o], dz[cur][(sk + 1) * W + o], dz[cur][(sk + 2) * W + o],
                                    dz[cur][(sk + 3) * W + o]);
#pragma unroll
          for (int mt = 0; mt < 2; ++mt) acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[mt], b, acc[mt][nt], 0, 0, 0);
        }
      }
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
          if (nt >= ntn) break;
          const int o = nt * 16 + (lane & 15);
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int i = 32 * wave + mt * 16 + 4 * (lane >> 4) + j;
            part[OFF_W[l] + i * no + o] = acc[mt][nt][j];
          }
        }
    }
    if (l == 0) break;
    // dH[s][i] = sum_o dZ[s][o] W[i][o]; dZ below = dH gelu'(Z below)
    {
      f32x4 acc[8];
#pragma unroll
      for (int nt = 0; nt < 8; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
      const int r = 16 * wave + (lane & 15);
      for (int kk = 0; kk < no / 16; ++kk) {
        const int o = kk * 16 + 4 * (lane >> 4);
        const float *d = &dz[cur][r * W + o];
        const bf16x4 a = pack_bf4(d[0], d[1], d[2], d[3]);
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
          const bf16x4 b = ld_bf4(net.wb[l] + (nt * 16 + (lane & 15)) * no + o);
          acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
        }
      }
#pragma unroll
      for (int nt = 0; nt < 8; ++nt) {
        const int i = nt * 16 + (lane & 15);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int rr = 16 * wave + 4 * (lane >> 4) + j;
          const long s = s0 + rr;
          dz[cur ^ 1][rr * W + i] = s < n ? acc[nt][j] * dgelu(wk.z[l - 1][s * W + i]) : 0.f;
        }
      }
    }
    __syncthreads();
    cur ^= 1;
  }
}

// Sum of the per-tile gradients in tile order.
__device__ __forceinline__ float grad_sum(const float *partial, int nblocks, int idx) {
  float g = 0.f;
  for (int b = 0; b < nblocks; ++b) g += partial[(long)b * NPARAM + idx];
  return g;
}

__device__ __forceinline__ void refresh_bf16(int idx, float w, bf16_t *const *wt, bf16_t *const *wb) {
  for (int l = 0; l < 4; ++l) {
    const int no = NOUT[l];
    if (idx >= OFF_W[l] && idx < OFF_W[l] + W * no) {
      const int i = (idx - OFF_W[l]) / no, o = (idx - OFF_W[l]) % no;
      wt[l][o * W + i] = f2bf(w);
      wb[l][i * no + o] = f2bf(w);
    }
  }
}

struct Bf16Ptrs {
  bf16_t *wt[4], *wb[4];
};

__global__ void grad_kernel(const float *partial, int nblocks, float *grad) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < NPARAM) grad[idx] = grad_sum(partial, nblocks, idx);
}

__global__ void adam_kernel(float *p, float *m, float *v, const float *partial, int nblocks, float lr, float b1,
                            float b2, float eps, float c1, float c2, Bf16Ptrs bp) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= NPARAM) return;
  const float g = grad_sum(partial, nblocks, idx);
  const float mm = b1 * m[idx] + (1.f - b1) * g;
  const float vv = b2 * v[idx] + (1.f - b2) * g * g;
  m[idx] = mm;
  v[idx] = vv;
  const float w = p[idx] - lr * (mm * c1) / (sqrtf(vv * c2) + eps);
  p[idx] = w;
  refresh_bf16(idx, w, bp.wt, bp.wb);
}

__global__ void bf16_refresh_kernel(const float *p, Bf16Ptrs bp) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < NPARAM) refresh_bf16(idx, p[idx], bp.wt, bp.wb);
}

__global__ void loss_sum_kernel(const float *part, int nblocks, float *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    float s = 0.f;
    for (int b = 0; b < nblocks; ++b) s += part[b];
    *out = s;
  }
}

}  // namespace mlp

// ---------------------------------------------------------------------
// per-particle inputs and the correction hook
// ---------------------------------------------------------------------
namespace {

// phi at the point x (after the clamp): the sum of the particles' profiles
// Phi(|x - x_p|) (profile.h:29-49) over the list of the tile holding the
// grid point below x, which has every particle whose profile reaches x
// (a particle reaching x has that grid point, within sqrt(3) dx of x,
// inside its stencil).
__device__ __forceinline__ double phi_at(const Grid &g, const PModel &pm, const double *__restrict__ pb,
                                         const int *__restrict__ tile_start, const int *__restrict__ tile_count,
                                         const int *__restrict__ tile_list, const int gi[3], const double x[3]) {
  const int tile = ((gi[0] / phi::kTile) * (g.ny / phi::kTile) + gi[1] / phi::kTile) * (g.nz / phi::kTile) +
                   gi[2] / phi::kTile;
  const int *lst = tile_list + tile_start[tile];
  const int cnt = tile_count[tile];
  const double rcut = pm.radius + pm.hxi;
  double ph = 0.0;
  for (int q = 0; q < cnt; ++q) {
    const int n = lst[q];
    const double xp[3] = {phi::P(pb, pm.np, P_X, n), phi::P(pb, pm.np, P_X + 1, n), phi::P(pb, pm.np, P_X + 2, n)};
    double x12[3];
    const double d = phi::distance0(g, x, xp, x12);
    if (d < rcut) ph += phi::phi_fast<double>(pm, d);
  }
  return ph > 1.0 ? 1.0 : ph;
}

// One thread per (particle, offset): u - v_p (trilinear in the 8 grid
// points around the sample point, periodic) and phi at x_p + h o, so the
// inputs move smoothly with the particle and turn with the lattice's
// symmetries.  The centre thread also writes the relative (slip) velocity:
// the phi-weighted mean of v + w x r - u over the particle, which
// Calc_f_hydro_correct_precision has just summed into f_hydro as
// F = -(rho dx^3 / dt) sum phi (v + w x r - u) (md_force.cxx:263-389),
// taken here as -F dt / (rho V_p), V_p = 4 pi a^3 / 3.
template <typename T>
__global__ void closure_samples_kernel(Grid g, PModel pm, const double *__restrict__ pb,
                                       const int *__restrict__ tile_start, const int *__restrict__ tile_count,
                                       const int *__restrict__ tile_list, const T *__restrict__ u0,
                                       const T *__restrict__ u1, const T *__restrict__ u2, double h,
                                       double slip_scale, float *x) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 27L * pm.np) return;
  const int n = (int)(t / 27), o = (int)(t % 27);
  const int od[3] = {o / 9 - 1, (o / 3) % 3 - 1, o % 3 - 1};
  const int N[3] = {g.nx, g.ny, g.nz};
  const double L[3] = {g.lx, g.ly, g.lz};
  int i0[3];
  double fr[3], xs[3];
  for (int d = 0; d < 3; ++d) {
    xs[d] = phi::P(pb, pm.np, P_X + d, n) + od[d] * h;
    xs[d] -= L[d] * floor(xs[d] / L[d]);
    const double gx = xs[d] / pm.dx;
    const double fl = floor(gx);
    i0[d] = ((int)fl % N[d] + N[d]) % N[d];
    fr[d] = gx - fl;
  }
  double us[3] = {0.0, 0.0, 0.0};
  for (int c = 0; c < 8; ++c) {
    const int cb[3] = {(c >> 2) & 1, (c >> 1) & 1, c & 1};
    long im = 0;
    double wgt = 1.0;
    for (int d = 0; d < 3; ++d) wgt *= cb[d] ? fr[d] : 1.0 - fr[d];
    im = ((long)((i0[0] + cb[0]) % N[0]) * g.ny + (i0[1] + cb[1]) % N[1]) * g.nz_ + (i0[2] + cb[2]) % N[2];
    us[0] += wgt * (double)u0[im];
    us[1] += wgt * (double)u1[im];
    us[2] += wgt * (double)u2[im];
  }
  float *xr = x + (long)n * mlp::IN;
  for (int d = 0; d < 3; ++d) xr[3 + 3 * o + d] = (float)(us[d] - phi::P(pb, pm.np, P_V + d, n));
  xr[84 + o] = (float)phi_at(g, pm, pb, tile_start, tile_count, tile_list, i0, xs);
  if (o == 13)
    for (int d = 0; d < 3; ++d) xr[d] = (float)(slip_scale * phi::P(pb, pm.np, P_F_HYDRO + d, n));
}

// The two nearest other particles (minimum image): one wavefront per
// particle, lanes striding over the others keeping their two nearest, then a
// shuffle reduction merging the pairs (smaller distance first, the lower
// index among equal ones).  Writes, for the nearest, the gap vector (the
// centre separation r times 1 - 2a/|r|), the relative velocity v_j - v_n and
// angular velocity w_j - w_n; for the second nearest, the gap vector and
// relative velocity; zeros where there is no such particle.
__device__ __forceinline__ bool nearer(double d, int i, double e, int j) { return d < e || (d == e && i < j); }

__global__ __launch_bounds__(256) void closure_gap_kernel(Grid g, PModel pm, const double *__restrict__ pb,
                                                          float *x) {
  const int lane = threadIdx.x & 63;
  const int n = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (n >= pm.np) return;
  const double L[3] = {g.lx, g.ly, g.lz};
  const double iL[3] = {1.0 / g.lx, 1.0 / g.ly, 1.0 / g.lz};
  const double xn[3] = {phi::P(pb, pm.np, P_X, n), phi::P(pb, pm.np, P_X + 1, n), phi::P(pb, pm.np, P_X + 2, n)};
  double b1 = 1e300, b2 = 1e300;
  int i1 = pm.np, i2 = pm.np;
  for (int m = lane; m < pm.np; m += 64) {
    if (m == n) continue;
    double r2 = 0.0;
    for (int d = 0; d < 3; ++d) {
      double r = phi::P(pb, pm.np, P_X + d, m) - xn[d];
      r -= L[d] * rint(r * iL[d]);
      r2 += r * r;
    }
    if (nearer(r2, m, b1, i1)) {
      b2 = b1;
      i2 = i1;
      b1 = r2;
      i1 = m;
    } else if (nearer(r2, m, b2, i2)) {
      b2 = r2;
      i2 = m;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o1 = __shfl_xor(b1, off, 64), o2 = __shfl_xor(b2, off, 64);
    const int j1 = __shfl_xor(i1, off, 64), j2 = __shfl_xor(i2, off, 64);
    // the two nearest of {(b1, i1), (b2, i2), (o1, j1), (o2, j2)}, each pair sorted
    if (nearer(o1, j1, b1, i1)) {
      const bool keep_b1 = nearer(b1, i1, o2, j2);
      b2 = keep_b1 ? b1 : o2;
      i2 = keep_b1 ? i1 : j2;
      b1 = o1;
      i1 = j1;
    } else if (nearer(o1, j1, b2, i2)) {
      b2 = o1;
      i2 = j1;
    }
  }
  if (lane != 0) return;
  float *xr = x + (long)n * mlp::IN;
  auto put = [&](int j, double dist2, int at, bool omega) {
    if (j >= pm.np) {
      for (int d = 0; d < (omega ? 9 : 6); ++d) xr[at + d] = 0.f;
      return;
    }
    const double f = 1.0 - 2.0 * pm.radius / sqrt(dist2);
    for (int d = 0; d < 3; ++d) {
      double r = phi::P(pb, pm.np, P_X + d, j) - xn[d];
      r -= L[d] * rint(r * iL[d]);
      xr[at + d] = (float)(r * f);
      xr[at + 3 + d] = (float)(phi::P(pb, pm.np, P_V + d, j) - phi::P(pb, pm.np, P_V + d, n));
      if (omega) xr[at + 6 + d] = (float)(phi::P(pb, pm.np, P_OMEGA + d, j) - phi::P(pb, pm.np, P_OMEGA + d, n));
    }
  };
  put(i1, b1, 111, true);
  put(i2, b2, 120, false);
}

// f_hydro += y[0..3), torque_hydro += y[3..6) (torque only with rotation,
// as the reference stores the torque only then)
__global__ void closure_apply_kernel(PModel pm, double *pb, const float *__restrict__ y) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= pm.np) return;
  for (int d = 0; d < 3; ++d) {
    pb[(long)(P_F_HYDRO + d) * pm.np + n] += (double)y[(long)n * mlp::OUT + d];
    if (pm.rotation) pb[(long)(P_TORQUE_HYDRO + d) * pm.np + n] += (double)y[(long)n * mlp::OUT + 3 + d];
  }
}

}  // namespace
}  // namespace spm

// ---------------------------------------------------------------------
// the network object
// ---------------------------------------------------------------------
struct spm_closure {
  int device = 0;
  int max_batch = 0;
  hipStream_t stream = nullptr;
  std::string err;
  float *p = nullptr, *m = nullptr, *v = nullptr;  // master weights, Adam moments
  spm::mlp::bf16_t *wt[4] = {}, *wb[4] = {};
  float *scal = nullptr;  // shift[128], scale[128], oscale[16]
  int adam_t = 0;
  // work for batches of up to max_batch
  spm::mlp::bf16_t *h[4] = {};
  float *z[3] = {};
  float *dy = nullptr, *loss = nullptr, *loss_sum = nullptr, *partial = nullptr, *grad = nullptr;
  float *x = nullptr, *y = nullptr, *t = nullptr;  // staging of one batch
  // training set
  float *data_x = nullptr, *data_t = nullptr;
  int data_n = 0, data_cap = 0;
};

namespace {

using spm::mlp::NPARAM;

#define CL_HIP(cl, call)                                                 \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) {                                              \
      (cl)->err = std::string(#call) + ": " + hipGetErrorString(e_);     \
      spm::global_error() = (cl)->err;                                   \
      return SPM_ERR_HIP;                                                \
    }                                                                    \
  } while (0)

#define CL_RC(x)          \
  do {                    \
    const int rc_ = (x);  \
    if (rc_) return rc_;  \
  } while (0)

template <typename P>
int cl_alloc(spm_closure *cl, P **p, size_t count) {
  CL_HIP(cl, hipMalloc((void **)p, sizeof(P) * (count > 0 ? count : 1)));
  CL_HIP(cl, hipMemsetAsync(*p, 0, sizeof(P) * (count > 0 ? count : 1), cl->stream));
  return 0;
}

void cl_free(spm_closure *cl) {
  if (!cl) return;
  (void)hipSetDevice(cl->device);
  if (cl->stream) (void)hipStreamSynchronize(cl->stream);
  for (float *q : {cl->p, cl->m, cl->v, cl->scal, cl->z[0], cl->z[1], cl->z[2], cl->dy, cl->loss, cl->loss_sum,
                   cl->partial, cl->grad, cl->x, cl->y, cl->t, cl->data_x, cl->data_t})
    if (q) (void)hipFree(q);
  for (int l = 0; l < 4; ++l) {
    if (cl->wt[l]) (void)hipFree(cl->wt[l]);
    if (cl->wb[l]) (void)hipFree(cl->wb[l]);
    if (cl->h[l]) (void)hipFree(cl->h[l]);
  }
  if (cl->stream) (void)hipStreamDestroy(cl->stream);
  delete cl;
}

spm::mlp::Net net_of(const spm_closure *cl) {
  spm::mlp::Net n{};
  n.p = cl->p;
  for (int l = 0; l < 4; ++l) {
    n.wt[l] = cl->wt[l];
    n.wb[l] = cl->wb[l];
  }
  n.shift = cl->scal;
  n.scale = cl->scal + 128;
  n.oscale = cl->scal + 256;
  return n;
}

spm::mlp::Work work_of(const spm_closure *cl, bool save) {
  spm::mlp::Work w{};
  if (save) {
    for (int l = 0; l < 4; ++l) w.h[l] = cl->h[l];
    for (int l = 0; l < 3; ++l) w.z[l] = cl->z[l];
  }
  w.dy = cl->dy;
  w.loss = cl->loss;
  return w;
}

spm::mlp::Bf16Ptrs bf_of(spm_closure *cl) {
  spm::mlp::Bf16Ptrs b{};
  for (int l = 0; l < 4; ++l) {
    b.wt[l] = cl->wt[l];
    b.wb[l] = cl->wb[l];
  }
  return b;
}

inline unsigned tiles_of(int n) { return (unsigned)((n + spm::mlp::TILE - 1) / spm::mlp::TILE); }
inline unsigned ftiles_of(int n) { return (unsigned)((n + spm::mlp::FTILE - 1) / spm::mlp::FTILE); }
// batches from which inference takes mlp_infer_kernel (below it, the
// 16-sample blocks of mlp_fwd_kernel spread a small batch over more CUs)
constexpr int kInferMin = 2048;

int refresh(spm_closure *cl) {
  hipLaunchKernelGGL(spm::mlp::bf16_refresh_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, cl->stream, cl->p,
                     bf_of(cl));
  CL_HIP(cl, hipGetLastError());
  return 0;
}

// forward of n samples at device x (training saves with target != null)
int fwd(spm_closure *cl, const float *x, float *y, const float *target, int n) {
  if (n <= 0) return 0;
  if (!target && n >= kInferMin) {  // inference of a large batch: weights shared by 64 samples
    hipLaunchKernelGGL(spm::mlp::mlp_infer_kernel, dim3((unsigned)((n + spm::mlp::BTILE - 1) / spm::mlp::BTILE)),
                       dim3(spm::mlp::THREADS), 0, cl->stream, net_of(cl), x, n, y);
    CL_HIP(cl, hipGetLastError());
    return 0;
  }
  hipLaunchKernelGGL(spm::mlp::mlp_fwd_kernel, dim3(ftiles_of(n)), dim3(spm::mlp::THREADS), 0, cl->stream, net_of(cl),
                     x, n, y, target, work_of(cl, target != nullptr));
  CL_HIP(cl, hipGetLastError());
  return 0;
}

// forward + backward of n samples into cl->partial (per tile) and the loss
int fwd_bwd(spm_closure *cl, const float *x, const float *target, int n) {
  CL_RC(fwd(cl, x, nullptr, target, n));
  hipLaunchKernelGGL(spm::mlp::mlp_bwd_kernel, dim3(tiles_of(n)), dim3(spm::mlp::THREADS), 0, cl->stream, net_of(cl),
                     work_of(cl, true), n, cl->partial);
  hipLaunchKernelGGL(spm::mlp::loss_sum_kernel, dim3(1), dim3(64), 0, cl->stream, cl->loss, (int)ftiles_of(n),
                     cl->loss_sum);
  CL_HIP(cl, hipGetLastError());
  return 0;
}

int adam_step(spm_closure *cl, int n, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  ++cl->adam_t;
  const float c1 = 1.f / (1.f - std::pow(b1, (float)cl->adam_t));
  const float c2 = 1.f / (1.f - std::pow(b2, (float)cl->adam_t));
  hipLaunchKernelGGL(spm::mlp::adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, cl->stream, cl->p, cl->m, cl->v,
                     cl->partial, (int)tiles_of(n), lr, b1, b2, eps, c1, c2, bf_of(cl));
  CL_HIP(cl, hipGetLastError());
  return 0;
}

int check_n(spm_closure *cl, int n) {
  if (n < 0 || n > cl->max_batch) {
    cl->err = "batch larger than the network's max_batch";
    spm::global_error() = cl->err;
    return SPM_ERR_INVALID;
  }
  return 0;
}

}  // namespace

namespace spm {
// -dt / (rho V_p): f_hydro to the phi-weighted mean slip velocity
double slip_scale(const spm_ctx *c) {
  const double r = c->prm.a * c->prm.dx;
  return -c->prm.dt / (c->prm.rho * 4.0 / 3.0 * M_PI * r * r * r);
}

// the step's hook: inputs of the context's particles, forward, correction
int closure_apply(spm_ctx *c) {
  spm_closure *cl = c->closure;
  if (!cl || c->np == 0) return 0;
  if (c->np > cl->max_batch) {
    c->err = "more particles than the attached closure's max_batch";
    return SPM_ERR_INVALID;
  }
  const long nthr = 27L * c->np;
  tbeg(c);
  if (c->prm.precision == SPM_FP32)
    hipLaunchKernelGGL(closure_samples_kernel<float>, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, c->stream,
                       c->grid, c->pm, c->pbuf, c->tile_start, c->tile_count, c->tile_list,
                       (const float *)c->u[0], (const float *)c->u[1], (const float *)c->u[2],
                       c->prm.a * c->prm.dx, slip_scale(c), cl->x);
  else
    hipLaunchKernelGGL(closure_samples_kernel<double>, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, c->stream,
                       c->grid, c->pm, c->pbuf, c->tile_start, c->tile_count, c->tile_list,
                       (const double *)c->u[0], (const double *)c->u[1], (const double *)c->u[2],
                       c->prm.a * c->prm.dx, slip_scale(c), cl->x);
  hipLaunchKernelGGL(closure_gap_kernel, dim3((c->np + 3) / 4), dim3(256), 0, c->stream, c->grid, c->pm, c->pbuf,
                     cl->x);
  tend(c, SPM_K_PARTICLE, 4.0 * mlp::IN * c->np);
  SPM_LAUNCH_CHECK(c);
  // the network runs on the context's stream
  hipStream_t keep = cl->stream;
  cl->stream = c->stream;
  tbeg(c);
  const int rc = fwd(cl, cl->x, cl->y, nullptr, c->np);
  tend(c, SPM_K_PARTICLE, 2.0 * 48128.0 * c->np);
  cl->stream = keep;
  if (rc) {
    c->err = cl->err;
    return rc;
  }
  hipLaunchKernelGGL(closure_apply_kernel, dim3((c->np + 127) / 128), dim3(128), 0, c->stream, c->pm, c->pbuf, cl->y);
  SPM_LAUNCH_CHECK(c);
  return 0;
}
}  // namespace spm

// =====================================================================
// C ABI
// =====================================================================
extern "C" {

int spm_closure_create(int device, int max_batch, spm_closure **out) {
  if (!out || max_batch <= 0 || max_batch > (1 << 22)) return SPM_ERR_INVALID;
  *out = nullptr;
  spm::global_error().clear();
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    spm::global_error() = "no HIP device";
    return SPM_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= ndev) {
    spm::global_error() = "device ordinal out of range";
    return SPM_ERR_INVALID;
  }
  spm_closure *cl = new spm_closure();
  cl->device = device;
  cl->max_batch = max_batch;
  int rc = 0;
  auto fail = [&](int r) {
    spm::global_error() = cl->err;
    cl_free(cl);
    return r;
  };
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&cl->stream) != hipSuccess) {
    cl->err = "cannot set the device or create a stream";
    return fail(SPM_ERR_HIP);
  }
  const size_t nb = (size_t)max_batch, tiles = tiles_of(max_batch);
  if ((rc = cl_alloc(cl, &cl->p, NPARAM)) || (rc = cl_alloc(cl, &cl->m, NPARAM)) || (rc = cl_alloc(cl, &cl->v, NPARAM)) ||
      (rc = cl_alloc(cl, &cl->scal, 272)) || (rc = cl_alloc(cl, &cl->dy, nb * 16)) ||
      (rc = cl_alloc(cl, &cl->loss, ftiles_of(max_batch))) || (rc = cl_alloc(cl, &cl->loss_sum, 1)) ||
      (rc = cl_alloc(cl, &cl->partial, tiles * NPARAM)) || (rc = cl_alloc(cl, &cl->grad, NPARAM)) ||
      (rc = cl_alloc(cl, &cl->x, nb * spm::mlp::IN)) || (rc = cl_alloc(cl, &cl->y, nb * spm::mlp::OUT)) ||
      (rc = cl_alloc(cl, &cl->t, nb * spm::mlp::OUT)))
    return fail(rc);
  for (int l = 0; l < 4; ++l)
    if ((rc = cl_alloc(cl, &cl->wt[l], (size_t)128 * spm::mlp::NOUT[l])) ||
        (rc = cl_alloc(cl, &cl->wb[l], (size_t)128 * spm::mlp::NOUT[l])) || (rc = cl_alloc(cl, &cl->h[l], nb * 128)))
      return fail(rc);
  for (int l = 0; l < 3; ++l)
    if ((rc = cl_alloc(cl, &cl->z[l], nb * 128))) return fail(rc);
  // identity scaling
  std::vector<float> sc(272, 0.f);
  for (int i = 128; i < 272; ++i) sc[i] = 1.f;
  if (hipMemcpy(cl->scal, sc.data(), sizeof(float) * 272, hipMemcpyHostToDevice) != hipSuccess) {
    cl->err = "cannot upload the scaling";
    return fail(SPM_ERR_HIP);
  }
  if ((rc = spm_closure_init(cl, 0))) return fail(rc);
  *out = cl;
  return 0;
}

int spm_closure_destroy(spm_closure *cl) {
  cl_free(cl);
  return 0;
}

int spm_closure_init(spm_closure *cl, uint64_t seed) {
  if (!cl) return SPM_ERR_INVALID;
  using namespace spm::mlp;
  std::vector<float> h(NPARAM, 0.f);
  std::mt19937_64 gen(seed);
  const int nin[4] = {IN, W, W, W}, nout[4] = {W, W, W, OUT};
  for (int l = 0; l < 4; ++l) {
    const double lim = std::sqrt(6.0 / (nin[l] + nout[l]));
    std::uniform_real_distribution<double> U(-lim, lim);
    for (int i = 0; i < nin[l]; ++i)
      for (int o = 0; o < nout[l]; ++o) h[OFF_W[l] + i * NOUT[l] + o] = (float)U(gen);
  }
  CL_HIP(cl, hipSetDevice(cl->device));
  CL_HIP(cl, hipMemcpy(cl->p, h.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  CL_HIP(cl, hipMemset(cl->m, 0, sizeof(float) * NPARAM));
  CL_HIP(cl, hipMemset(cl->v, 0, sizeof(float) * NPARAM));
  cl->adam_t = 0;
  CL_RC(refresh(cl));
  CL_HIP(cl, hipStreamSynchronize(cl->stream));
  return 0;
}

int spm_closure_get_params(spm_closure *cl, float *host) {
  if (!cl || !host) return SPM_ERR_INVALID;
  CL_HIP(cl, hipSetDevice(cl->device));
  CL_HIP(cl, hipStreamSynchronize(cl->stream));
  CL_HIP(cl, hipMemcpy(host, cl->p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  return 0;
}

int spm_closure_put_params(spm_closure *cl, const float *host) {
  if (!cl || !host) return SPM_ERR_INVALID;
  using namespace spm::mlp;
  // padding must stay zero: inputs beyond SPM_CLOSURE_IN, outputs beyond 6
  for (int i = IN; i < W; ++i)
    for (int o = 0; o < W; ++o)
      if (host[OFF_W[0] + i * W + o] != 0.f) {
        cl->err = "W1 rows beyond the inputs must be zero";
        return SPM_ERR_INVALID;
      }
  for (int i = 0; i < W; ++i)
    for (int o = OUT; o < OUTP; ++o)
      if (host[OFF_W[3] + i * OUTP + o] != 0.f) {
        cl->err = "W4 columns beyond output 6 must be zero";
        return SPM_ERR_INVALID;
      }
  for (int o = OUT; o < OUTP; ++o)
    if (host[OFF_B[3] + o] != 0.f) {
      cl->err = "b4 beyond output 6 must be zero";
      return SPM_ERR_INVALID;
    }
  CL_HIP(cl, hipSetDevice(cl->device));
  CL_HIP(cl, hipStreamSynchronize(cl->stream));
  CL_HIP(cl, hipMemcpy(cl->p, host, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  CL_RC(refresh(cl));
  CL_HIP(cl, hipStreamSynchronize(cl->stream));
  return 0;
}

int spm_closure_set_scaling(spm_closure *cl, const float *in_shift, const float *in_scale, const float *out_scale) {
  if (!cl || !in_shift || !in_scale || !out_scale) return SPM_ERR_INVALID;
  std::vector<float> sc(272, 0.f);
  for (int i = 128; i < 272; ++i) sc[i] = 1.f;
  for (int i = 0; i < spm::mlp::IN; ++i) {
    sc[i] = in_shift[i];
    sc[128 + i] = in_scale[i];
  }
  for (int o = 0; o < spm::mlp::OUT; ++o) {
    if (!(out_scale[o] != 0.f) || !std::isfinite(out_scale[o])) {
      cl->err = "output scales must be finite and nonzero";
      return SPM_ERR_INVALID;
    }
    sc[256 + o] = out_scale[o];
  }
  CL_HIP(cl, hipSetDevice(cl->device));
  CL_HIP(cl, hipStreamSynchronize(cl->stream));
  CL_HIP(cl, hipMemcpy(cl->scal, sc.data(), sizeof(float) * 272, hipMemcpyHostToDevice));
  return 0;
}

int spm_closure_get_scaling(spm_closure *cl, float *in_shift, float *in_scale, float *out_scale) {
  if (!cl || !in_shift || !in_scale || !out_scale) return SPM_ERR_INVALID;
  std::vector<float> sc(272);
  CL_HIP(cl, hipSetDevice(cl->device));
  CL_HIP(cl, hipStreamSynchronize(cl->stream));
  CL_HIP(cl, hipMemcpy(sc.data(), cl->scal, sizeof(float) * 272, hipMemcpyDeviceToHost));
  std::memcpy(in_shift, sc.data(), sizeof(float) * spm::mlp::IN);
  std::memcpy(in_scale, sc.data() + 128, sizeof(float) * spm::mlp::IN);
  std::memcpy(out_scale, sc.data() + 256, sizeof(float) * spm::mlp::OUT);
  return 0;
}

int spm_closure_forward(spm_closure *cl, const float *x, float *y, int n) {
  if (!cl || (n > 0 && (!x || !y))) return SPM_ERR_INVALID;
  CL_RC(check_n(cl, n));
  if (n == 0) return 0;
  CL_HIP(cl, hipSetDevice(cl->device));
  CL_HIP(cl, hipMemcpyAsync(cl->x, x, sizeof(float) * n * spm::mlp::IN, hipMemcpyHostToDevice, cl->stream));
  CL_RC(fwd(cl, cl->x, cl->y, nullptr, n));
  CL_HIP(cl, hipMemcpyAsync(y, cl->y, sizeof(float) * n * spm::mlp::OUT, hipMemcpyDeviceToHost, cl->stream));
  CL_HIP(cl, hipStreamSynchronize(cl->stream));
  return 0;
}

int spm_closure_gradient(spm_closure *cl, const float *x, const float *target, int n, float *grad, float *loss) {
  if (!cl || !x || !target || !grad || !loss || n <= 0) return SPM_ERR_INVALID;
  CL_RC(check_n(cl, n));
  CL_HIP(cl, hipSetDevice(cl->device));
  CL_HIP(cl, hipMemcpyAsync(cl->x, x, sizeof(float) * n * spm::mlp::IN, hipMemcpyHostToDevice, cl->stream));
  CL_HIP(cl, hipMemcpyAsync(cl->t, target, sizeof(float) * n * spm::mlp::OUT, hipMemcpyHostToDevice, cl->stream));
  CL_RC(fwd_bwd(cl, cl->x, cl->t, n));
  hipLaunchKernelGGL(spm::mlp::grad_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, cl->stream, cl->partial,
                     (int)tiles_of(n), cl->grad);
  CL_HIP(cl, hipGetLastError());
  CL_HIP(cl, hipMemcpyAsync(grad, cl->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, cl->stream));
  CL_HIP(cl, hipMemcpyAsync(loss, cl->loss_sum, sizeof(float), hipMemcpyDeviceToHost, cl->stream));
  CL_HIP(cl, hipStreamSynchronize(cl->stream));
  *loss /= (float)n * spm::mlp::OUT;
  return 0;
}

int spm_closure_set_data(spm_closure *cl, const float *x, const float *target, int n) {
  if (!cl || n <= 0 || !x || !target) return SPM_ERR_INVALID;
  CL_HIP(cl, hipSetDevice(cl->device));
  CL_HIP(cl, hipStreamSynchronize(cl->stream));
  if (n > cl->data_cap) {
    if (cl->data_x) (void)hipFree(cl->data_x);
    if (cl->data_t) (void)hipFree(cl->data_t);
    cl->data_x = cl->data_t = nullptr;
    cl->data_cap = 0;
    CL_HIP(cl, hipMalloc((void **)&cl->data_x, sizeof(float) * (size_t)n * spm::mlp::IN));
    CL_HIP(cl, hipMalloc((void **)&cl->data_t, sizeof(float) * (size_t)n * spm::mlp::OUT));
    cl->data_cap = n;
  }
  CL_HIP(cl, hipMemcpy(cl->data_x, x, sizeof(float) * (size_t)n * spm::mlp::IN, hipMemcpyHostToDevice));
  CL_HIP(cl, hipMemcpy(cl->data_t, target, sizeof(float) * (size_t)n * spm::mlp::OUT, hipMemcpyHostToDevice));
  cl->data_n = n;
  return 0;
}

int spm_closure_train(spm_closure *cl, int steps, int batch, float lr, float *loss) {
  if (!cl || steps < 0 || batch <= 0 || !(lr > 0.f)) return SPM_ERR_INVALID;
  if (cl->data_n <= 0) {
    cl->err = "no training set (spm_closure_set_data)";
    return SPM_ERR_INVALID;
  }
  if (batch > cl->data_n) batch = cl->data_n;
  CL_RC(check_n(cl, batch));
  CL_HIP(cl, hipSetDevice(cl->device));
  int start = 0;
  for (int s = 0; s < steps; ++s) {
    const int nb = std::min(batch, cl->data_n - start);
    CL_RC(fwd_bwd(cl, cl->data_x + (size_t)start * spm::mlp::IN, cl->data_t + (size_t)start * spm::mlp::OUT, nb));
    CL_RC(adam_step(cl, nb, lr));
    if (loss) {
      float l = 0.f;
      CL_HIP(cl, hipMemcpyAsync(&l, cl->loss_sum, sizeof(float), hipMemcpyDeviceToHost, cl->stream));
      CL_HIP(cl, hipStreamSynchronize(cl->stream));
      loss[s] = l / ((float)nb * spm::mlp::OUT);
    }
    start += nb;
    if (start >= cl->data_n) start = 0;
  }
  CL_HIP(cl, hipStreamSynchronize(cl->stream));
  return 0;
}

int spm_set_closure(spm_ctx *c, spm_closure *cl) {
  if (!c) return SPM_ERR_INVALID;
  if (cl && cl->device != c->device) {
    c->err = "the closure lives on another device";
    return SPM_ERR_INVALID;
  }
  if (cl && c->np > cl->max_batch) {
    c->err = "more particles than the closure's max_batch";
    return SPM_ERR_INVALID;
  }
  if (cl && (c->le || c->electro)) {
    c->err = "the closure is for the Navier-Stokes step";
    return SPM_ERR_UNSUPPORTED;
  }
  c->closure = cl;
  return 0;
}

int spm_closure_features(spm_ctx *c, float *x) {
  if (!c || (!x && c->np > 0)) return SPM_ERR_INVALID;
  if (c->np == 0) return 0;
  (void)hipSetDevice(c->device);
  if (int rs = spm::settle(c)) return rs;  // the inputs read u
  float *d = nullptr;
  SPM_HIP(c, hipMalloc((void **)&d, sizeof(float) * (size_t)c->np * spm::mlp::IN));
  int rc = spm::p_build_tiles(c);
  if (!rc) {
    const long nthr = 27L * c->np;
    if (c->prm.precision == SPM_FP32)
      hipLaunchKernelGGL(spm::closure_samples_kernel<float>, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0,
                         c->stream, c->grid, c->pm, c->pbuf, c->tile_start, c->tile_count,
                         c->tile_list, (const float *)c->u[0], (const float *)c->u[1], (const float *)c->u[2],
                         c->prm.a * c->prm.dx, spm::slip_scale(c), d);
    else
      hipLaunchKernelGGL(spm::closure_samples_kernel<double>, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0,
                         c->stream, c->grid, c->pm, c->pbuf, c->tile_start, c->tile_count,
                         c->tile_list, (const double *)c->u[0], (const double *)c->u[1], (const double *)c->u[2],
                         c->prm.a * c->prm.dx, spm::slip_scale(c), d);
    hipLaunchKernelGGL(spm::closure_gap_kernel, dim3((c->np + 3) / 4), dim3(256), 0, c->stream, c->grid, c->pm,
                       c->pbuf, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(x, d, sizeof(float) * (size_t)c->np * spm::mlp::IN, hipMemcpyDeviceToHost,
                                            c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      c->err = std::string("closure features: ") + hipGetErrorString(e);
      rc = SPM_ERR_HIP;
    }
  }
  (void)hipFree(d);
  return rc;
}

int spm_set_particle_motion(spm_ctx *c, int motion) {
  if (!c || (motion != SPM_MOTION_FREE && motion != SPM_MOTION_PRESCRIBED)) return SPM_ERR_INVALID;
  if (motion == SPM_MOTION_PRESCRIBED && (c->le || c->electro)) {
    c->err = "prescribed motion is for the Navier-Stokes step";
    return SPM_ERR_UNSUPPORTED;
  }
  c->motion = motion;
  return 0;
}

}  // extern "C"
