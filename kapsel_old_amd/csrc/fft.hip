// fft.hip -- 3-D real FFTs in KAPSEL's layout, hand-written for gfx950.
//
// Replaces A2a_k / A_k2a (fft_wrapper.h:432-631), which call Ooura's
// rdft3d + rdft3dsort (fftsg3d.c:560-723).  Same transform:
//   forward  X(k) = sum_x a(x) exp(+2 pi i k.x / N)   (fftsg3d.c:109-120)
//   inverse  a(x) = (1/N) sum_k X(k) exp(-2 pi i k.x / N)
// on a [NX][NY][NZ_] array, real data in k < NZ, spectrum as NZ/2+1
// interleaved complex values per row.
//
// A 3-D transform is three passes, each one read and one write of the
// field, staged through LDS:
//   z pass  : rows are contiguous; NZ reals are packed as NZ/2 complex
//             values, transformed, and split into the half spectrum.
//   y, x    : complex lines with stride NZ_/2 (y) or NY*NZ_/2 (x); a block
//             loads C adjacent columns of N points, so global accesses are
//             C-element contiguous segments.  The x pass numbers its
//             NY*(NZ/2+1) columns as one run, so a tile's segment starts on
//             a 128-byte boundary; the y pass keeps the kz = NZ/2 column of
//             each x plane for tiles of its own (LineTiles).
// Inside a block the transform is a radix-8 Stockham FFT: each thread
// keeps 8 points in registers per stage, stages exchange through LDS laid
// out [point][column] so a wavefront touches consecutive LDS words.
#include <climits>
#include <cstdio>
#include <functional>
#include <map>
#include <mutex>
#include <tuple>

#include "fft_core.h"

namespace spm {

// ---------------------------------------------------------------------
// y / x passes: complex lines of N points at element stride `estride`
// (complex units).  A block transforms C columns; LineTiles says which.
// blockIdx.x = tile (XCD-remapped), blockIdx.z = field.
//
// The fused schedule only carries modes the 2/3 rule keeps (|kx| < TRN_X,
// |ky| < TRN_Y, kz < TRN_Z).  For its y passes: columns whose index within
// the x plane (their kz) is >= col_cut are left alone; of each line, the
// points j with min(j, N - j) >= row_keep (|ky| >= TRN_Y) are taken as zero
// on the way in (inverse) or not stored (forward).  col_cut = INT_MAX and
// row_keep = N give the plain transform.
// ---------------------------------------------------------------------
template <typename T, int N, int R, int C, int SIGN>
__global__ __launch_bounds__(C *N / R) void fft_lines_kernel(FieldPtrs<T> fp, const double2 *__restrict__ tw_g,
                                                             LineTiles lt, long estride, int col_cut, int row_keep) {
  using C2 = cplx<T>;
  __shared__ C2 lds[N * C];
  const int tid = threadIdx.x;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  if (line_first_col(lt, tile, C) >= col_cut) return;  // uniform over the block
  // twiddle tables in LDS: the transform's barriers then wait on LDS
  // only, and global loads issued ahead stay in flight across them
  __shared__ double2 tw[N];
  for (int t = tid; t < N; t += blockDim.x) tw[t] = tw_g[t];
  const long col = line_column(lt, tile, tid % C, C);
  const bool col_ok = col >= 0 && line_first_col(lt, tile, C) + (tile < lt.main ? tid % C : 0) < col_cut;
  C2 *g = reinterpret_cast<C2 *>(fp.p[blockIdx.z]) + (col_ok ? col : 0);
  auto row_kept = [&](int row) { return (row < N - row ? row : N - row) < row_keep; };
  auto ld = [&](int row, int, int) -> C2 {
    return (col_ok && (SIGN > 0 || row_kept(row))) ? g[row * estride] : C2{0, 0};
  };
  auto st = [&](int row, int, int, C2 v) {
    if (col_ok && (SIGN < 0 || row_kept(row))) g[row * estride] = v;
  };
  fft_stages_io<T, N, R, C, C, SIGN, 1, true, true>(lds, tw, tid, ld, st);
}

// ---------------------------------------------------------------------
// z pass, forward: ROWS rows per block, each NZ = 2M reals -> M+1 complex.
// ---------------------------------------------------------------------
template <typename T, int M, int R, int ROWS>
__global__ __launch_bounds__(ROWS *M / R) void fft_z_r2c_kernel(FieldPtrs<T> fp, const double2 *__restrict__ tw_g,
                                                               const double2 *__restrict__ twr_g, int nrows,
                                                               int nz_, int kz_cut) {
  using C2 = cplx<T>;
  constexpr int NT = ROWS * M / R;
  constexpr int CP = ROWS + 1;
  constexpr int H = M + 1;
  __shared__ C2 lds[M * CP];
  T *a = fp.p[blockIdx.z];
  const int tid = threadIdx.x;
  // twiddle tables in LDS: the transform's barriers then wait on LDS
  // only, and global loads issued ahead stay in flight across them
  __shared__ double2 tw[M];
  for (int t = tid; t < M; t += blockDim.x) tw[t] = tw_g[t];
  __shared__ double2 twr[M + 1];
  for (int t = tid; t < M + 1; t += blockDim.x) twr[t] = twr_g[t];
  const long row0 = (long)blockIdx.x * ROWS;
#pragma unroll 4
  for (int f = tid; f < ROWS * M; f += NT) {
    const int m = f % M, r = f / M;
    const long row = row0 + r;
    lds[m * CP + r] = (row < nrows) ? *reinterpret_cast<const C2 *>(a + row * nz_ + 2 * m) : C2{0, 0};
  }
  __syncthreads();
  fft_stages<T, M, R, ROWS, CP, +1, 1>(lds, tw, tid);
  for (int f = tid; f < ROWS * H; f += NT) {
    const int k = f % H, r = f / H;
    const long row = row0 + r;
    const C2 z = lds[(k % M) * CP + r];
    C2 zc = lds[((M - k) % M) * CP + r];
    zc.y = -zc.y;
    const C2 e{(T)0.5 * (z.x + zc.x), (T)0.5 * (z.y + zc.y)};
    const C2 o{(T)0.5 * (z.y - zc.y), (T)-0.5 * (z.x - zc.x)};  // (z - zc) / (2i)
    const double2 w = twr[k];
    const C2 x = cadd(e, cmul(o, C2{(T)w.x, (T)w.y}));
    if (row < nrows && k < kz_cut) *reinterpret_cast<C2 *>(a + row * nz_ + 2 * k) = x;
  }
}

// ---------------------------------------------------------------------
// z pass, inverse: M+1 complex -> NZ = 2M reals per row, times `scale`.
// ---------------------------------------------------------------------
template <typename T, int M, int R, int ROWS>
__global__ __launch_bounds__(ROWS *M / R) void fft_z_c2r_kernel(FieldPtrs<T> fp, const double2 *__restrict__ tw_g,
                                                               const double2 *__restrict__ twr_g, int nrows, int nz_,
                                                               T scale, int kz_cut) {
  using C2 = cplx<T>;
  constexpr int NT = ROWS * M / R;
  constexpr int CP = ROWS + 1;
  constexpr int H = M + 1;
  __shared__ C2 lds[H * CP];
  T *a = fp.p[blockIdx.z];
  const int tid = threadIdx.x;
  // twiddle tables in LDS: the transform's barriers then wait on LDS
  // only, and global loads issued ahead stay in flight across them
  __shared__ double2 tw[M];
  for (int t = tid; t < M; t += blockDim.x) tw[t] = tw_g[t];
  __shared__ double2 twr[M + 1];
  for (int t = tid; t < M + 1; t += blockDim.x) twr[t] = twr_g[t];
  const long row0 = (long)blockIdx.x * ROWS;
  for (int f = tid; f < ROWS * H; f += NT) {
    const int k = f % H, r = f / H;
    const long row = row0 + r;
    lds[k * CP + r] = (row < nrows && k < kz_cut) ? *reinterpret_cast<const C2 *>(a + row * nz_ + 2 * k) : C2{0, 0};
  }
  __syncthreads();
  // Z'[k] = (X[k] + conj X[M-k]) + i (X[k] - conj X[M-k]) exp(-2 pi i k/NZ),
  // formed by the first Stockham stage from the pair in LDS
  auto ld = [&](int k, int r, int) -> C2 {
    const C2 xk = lds[k * CP + r], xkk = lds[(M - k) * CP + r];
    const C2 cj{xkk.x, -xkk.y};
    const C2 e = cadd(xk, cj), d = csub(xk, cj);
    const double2 w = twr[k];
    const C2 o = cmul(d, C2{(T)w.x, (T)-w.y});
    return C2{e.x - o.y, e.y + o.x};
  };
  auto none = [](int, int, int, C2) {};
  fft_stages_io<T, M, R, ROWS, CP, -1, 1, true, false>(lds, tw, tid, ld, none);
  for (int f = tid; f < ROWS * H; f += NT) {
    const int m = f % H, r = f / H;
    const long row = row0 + r;
    if (row < nrows) {
      C2 v{0, 0};
      if (m < M) {
        v = lds[m * CP + r];
        v.x *= scale;
        v.y *= scale;
      }
      *reinterpret_cast<C2 *>(a + row * nz_ + 2 * m) = v;  // m == M writes the zero padding
    }
  }
}

// ---------------------------------------------------------------------
// The inverse reads the kz = 0 and kz = NZ/2 planes as Ooura's rdft3d does
// (rdft3dsort(-1), fftsg3d.c:645-668): of each pair (k, -k) in a plane one
// member is kept and the other taken as its conjugate, and the four
// self-conjugate points keep their real part.  Kept on kz = 0: 1 <= j < NY/2,
// or j in {0, NY/2} with i <= NX/2; on kz = NZ/2: j > NY/2, or j in
// {0, NY/2} with i > NX/2.  For a Hermitian spectrum this changes nothing;
// for one that is not (the Nyquist modes of a derivative) it makes the
// result the reference's.  A kept point is never written, so in place is
// safe.  One thread per (i, j) and plane.
// ---------------------------------------------------------------------
template <typename T>
__global__ void ooura_planes_kernel(FieldPtrs<T> fp, int nx, int ny, int nz_, int m) {
  using C2 = cplx<T>;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2L * nx * ny) return;
  const int plane = (int)(t / ((long)nx * ny));
  const int ij = (int)(t % ((long)nx * ny));
  const int i = ij / ny, j = ij % ny, n1h = nx / 2, n2h = ny / 2;
  const int kz = plane == 0 ? 0 : m;
  const bool edge = (j == 0 || j == n2h);
  const bool selfc = edge && (i == 0 || i == n1h);
  const bool keep = plane == 0 ? ((j >= 1 && j < n2h) || (edge && i <= n1h)) : ((j > n2h) || (edge && i > n1h) || selfc);
  T *a = fp.p[blockIdx.y];
  C2 *x = reinterpret_cast<C2 *>(a + ((long)i * ny + j) * nz_ + 2 * kz);
  if (selfc) {
    x->y = (T)0;
  } else if (!keep) {
    const C2 mv = *reinterpret_cast<const C2 *>(a + ((long)((nx - i) % nx) * ny + (ny - j) % ny) * nz_ + 2 * kz);
    *x = C2{mv.x, -mv.y};
  }
}

// ---------------------------------------------------------------------
// Lengths with factors 3 and 5 (grids of 48, 80, 96, 160 or 192 points per
// axis; the reference allows powers of two only, input.cxx:1397-1402): the
// same passes with the mixed-radix stages of fft_core.h on two LDS buffers.
// A block of TH threads takes C columns (y, x) or ROWS rows (z); which
// (C, TH) or ROWS a launch uses is chosen per length and pass by timing
// (tune_mixed_transforms), the defaults below otherwise.
// ---------------------------------------------------------------------
constexpr int kMixedThreads = 256;
constexpr size_t kMixedMaxLds = 65536;
template <typename T>
constexpr int kMixedC() { return sizeof(T) == 8 ? 8 : 16; }  // one 128-byte run per row
constexpr int kMixedRows(int m) { return m >= 1024 ? 1 : (1024 / m > 32 ? 32 : 1024 / m); }
template <typename T>
constexpr size_t lines_mixed_lds(int n, int c) { return 2 * (size_t)n * c * sizeof(cplx<T>) + 16 * (size_t)n; }
template <typename T>
constexpr size_t z_mixed_lds(int m, int rows) { return (2 * (size_t)m + 1) * rows * sizeof(cplx<T>) + 16 * (2 * (size_t)m + 1); }

template <typename T, int N, int C, int TH, int SIGN>
__global__ __launch_bounds__(TH) void fft_lines_mixed_kernel(FieldPtrs<T> fp, const double2 *__restrict__ tw_g,
                                                             LineTiles lt, long estride) {
  using C2 = cplx<T>;
  __shared__ C2 b0[N * C], b1[N * C];
  __shared__ double2 tw[N];
  const int tid = threadIdx.x;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  for (int t = tid; t < N; t += TH) tw[t] = tw_g[t];
  C2 *g = reinterpret_cast<C2 *>(fp.p[blockIdx.z]);
  for (int e = tid; e < N * C; e += TH) {
    const long col = line_column(lt, tile, e % C, C);
    b0[e] = col >= 0 ? g[col + (e / C) * estride] : C2{0, 0};
  }
  __syncthreads();
  mixed_stages<T, N, 1, C, SIGN>(b0, b1, tw, tid, TH);
  const C2 *res = (mixed_nstages(N) % 2 == 0) ? b0 : b1;
  for (int e = tid; e < N * C; e += TH) {
    const long col = line_column(lt, tile, e % C, C);
    if (col >= 0) g[col + (e / C) * estride] = res[e];
  }
}

// z pass forward: ROWS rows of NZ = 2M reals -> M + 1 complex each
template <typename T, int M, int ROWS>
__global__ __launch_bounds__(kMixedThreads) void fft_z_r2c_mixed_kernel(FieldPtrs<T> fp,
                                                                       const double2 *__restrict__ tw_g,
                                                                       const double2 *__restrict__ twr_g, int nrows,
                                                                       int nz_) {
  using C2 = cplx<T>;
  constexpr int H = M + 1;
  __shared__ C2 b0[M * ROWS], b1[M * ROWS];
  __shared__ double2 tw[M], twr[M + 1];
  T *a = fp.p[blockIdx.z];
  const int tid = threadIdx.x;
  for (int t = tid; t < M; t += kMixedThreads) tw[t] = tw_g[t];
  for (int t = tid; t < M + 1; t += kMixedThreads) twr[t] = twr_g[t];
  const long row0 = (long)blockIdx.x * ROWS;
  for (int f = tid; f < ROWS * M; f += kMixedThreads) {
    const int m = f % M, r = f / M;
    const long row = row0 + r;
    b0[m * ROWS + r] = (row < nrows) ? *reinterpret_cast<const C2 *>(a + row * nz_ + 2 * m) : C2{0, 0};
  }
  __syncthreads();
  mixed_stages<T, M, 1, ROWS, +1>(b0, b1, tw, tid, kMixedThreads);
  const C2 *res = (mixed_nstages(M) % 2 == 0) ? b0 : b1;
  for (int f = tid; f < ROWS * H; f += kMixedThreads) {
    const int k = f % H, r = f / H;
    const long row = row0 + r;
    const C2 z = res[(k % M) * ROWS + r];
    C2 zc = res[((M - k) % M) * ROWS + r];
    zc.y = -zc.y;
    const C2 e{(T)0.5 * (z.x + zc.x), (T)0.5 * (z.y + zc.y)};
    const C2 o{(T)0.5 * (z.y - zc.y), (T)-0.5 * (z.x - zc.x)};  // (z - zc) / (2i)
    const double2 w = twr[k];
    if (row < nrows) *reinterpret_cast<C2 *>(a + row * nz_ + 2 * k) = cadd(e, cmul(o, C2{(T)w.x, (T)w.y}));
  }
}

// z pass inverse: M + 1 complex -> NZ = 2M reals per row, times `scale`
template <typename T, int M, int ROWS>
__global__ __launch_bounds__(kMixedThreads) void fft_z_c2r_mixed_kernel(FieldPtrs<T> fp,
                                                                       const double2 *__restrict__ tw_g,
                                                                       const double2 *__restrict__ twr_g, int nrows,
                                                                       int nz_, T scale) {
  using C2 = cplx<T>;
  constexpr int H = M + 1, HP = M / 2 + 1;
  __shared__ C2 b0[H * ROWS], b1[M * ROWS];
  __shared__ double2 tw[M], twr[M + 1];
  T *a = fp.p[blockIdx.z];
  const int tid = threadIdx.x;
  for (int t = tid; t < M; t += kMixedThreads) tw[t] = tw_g[t];
  for (int t = tid; t < M + 1; t += kMixedThreads) twr[t] = twr_g[t];
  const long row0 = (long)blockIdx.x * ROWS;
  for (int f = tid; f < ROWS * H; f += kMixedThreads) {
    const int k = f % H, r = f / H;
    const long row = row0 + r;
    b0[k * ROWS + r] = (row < nrows) ? *reinterpret_cast<const C2 *>(a + row * nz_ + 2 * k) : C2{0, 0};
  }
  __syncthreads();
  // Z'[k] = (X[k] + conj X[M-k]) + i (X[k] - conj X[M-k]) exp(-2 pi i k/NZ), in place per pair
  for (int f = tid; f < ROWS * HP; f += kMixedThreads) {
    const int p = f % HP, r = f / HP;
    const int k = p, kk = M - p;
    const C2 xk = b0[k * ROWS + r], xkk = b0[kk * ROWS + r];
    {
      const C2 cj{xkk.x, -xkk.y};
      const C2 e = cadd(xk, cj), d = csub(xk, cj);
      const double2 w = twr[k];
      const C2 o = cmul(d, C2{(T)w.x, (T)-w.y});
      b0[k * ROWS + r] = C2{e.x - o.y, e.y + o.x};
    }
    if (p > 0 && kk != k) {
      const C2 cj{xk.x, -xk.y};
      const C2 e = cadd(xkk, cj), d = csub(xkk, cj);
      const double2 w = twr[kk];
      const C2 o = cmul(d, C2{(T)w.x, (T)-w.y});
      b0[kk * ROWS + r] = C2{e.x - o.y, e.y + o.x};
    }
  }
  __syncthreads();
  mixed_stages<T, M, 1, ROWS, -1>(b0, b1, tw, tid, kMixedThreads);
  const C2 *res = (mixed_nstages(M) % 2 == 0) ? b0 : b1;
  for (int f = tid; f < ROWS * H; f += kMixedThreads) {
    const int m = f % H, r = f / H;
    const long row = row0 + r;
    if (row < nrows) {
      C2 v{0, 0};
      if (m < M) {
        v = res[m * ROWS + r];
        v.x *= scale;
        v.y *= scale;
      }
      *reinterpret_cast<C2 *>(a + row * nz_ + 2 * m) = v;  // m == M writes the zero padding
    }
  }
}

// ---------------------------------------------------------------------
// host side: configuration per size and launch
// ---------------------------------------------------------------------
namespace {


template <typename T, int N, int SIGN>
void launch_lines_n(const FieldPtrs<T> &fp, int nf, const double2 *tw, int ncols, long estride, long ostride,
                    int nouter, hipStream_t st, int col_cut, int row_keep) {
  constexpr int R = kRadix(N), C = kLinesC<T>(N);
  const LineTiles lt = make_line_tiles(ncols, nouter, ostride, C);
  hipLaunchKernelGGL((fft_lines_kernel<T, N, R, C, SIGN>), dim3(lt.tiles, 1, nf), dim3(C * N / R), 0, st, fp, tw, lt,
                     estride, col_cut, row_keep);
}

// Launch-shape variants of the mixed-radix passes.  Lines: (C, TH) =
// (C0/2, 128), (C0, 256), (2 C0, 256), (C0, 128) with C0 = kMixedC<T>();
// z: ROWS = R0/2, R0, 2 R0 with R0 = kMixedRows(M).  A variant whose LDS
// would pass 64 KiB does not exist (its launch returns -1).
constexpr int kLinesVariants = 4, kZVariants = 3, kLinesDefault = 1, kZDefault = 1;

template <typename T, int N, int C, int TH, int SIGN>
int lines_mixed_go(const FieldPtrs<T> &fp, int nf, const double2 *tw, int ncols, long estride, long ostride,
                   int nouter, hipStream_t st) {
  if constexpr (lines_mixed_lds<T>(N, C) > kMixedMaxLds || C < 1) {
    return -1;
  } else {
    const LineTiles lt = make_line_tiles(ncols, nouter, ostride, C);
    hipLaunchKernelGGL((fft_lines_mixed_kernel<T, N, C, TH, SIGN>), dim3(lt.tiles, 1, nf), dim3(TH), 0, st, fp, tw,
                       lt, estride);
    return 0;
  }
}

template <typename T, int N, int SIGN>
int launch_lines_mixed(int v, const FieldPtrs<T> &fp, int nf, const double2 *tw, int ncols, long estride,
                       long ostride, int nouter, hipStream_t st) {
  constexpr int C0 = kMixedC<T>();
  switch (v) {
    case 0: return lines_mixed_go<T, N, C0 / 2, 128, SIGN>(fp, nf, tw, ncols, estride, ostride, nouter, st);
    case 1: return lines_mixed_go<T, N, C0, 256, SIGN>(fp, nf, tw, ncols, estride, ostride, nouter, st);
    case 2: return lines_mixed_go<T, N, 2 * C0, 256, SIGN>(fp, nf, tw, ncols, estride, ostride, nouter, st);
    case 3: return lines_mixed_go<T, N, C0, 128, SIGN>(fp, nf, tw, ncols, estride, ostride, nouter, st);
  }
  return -1;
}

template <typename T, int M, int ROWS>
int z_mixed_go(bool fwd, const FieldPtrs<T> &fp, int nf, const double2 *tw, const double2 *twr, int nrows, int nz_,
               T scale, hipStream_t st) {
  if constexpr (z_mixed_lds<T>(M, ROWS) > kMixedMaxLds || ROWS < 1) {
    return -1;
  } else {
    dim3 grid((nrows + ROWS - 1) / ROWS, 1, nf);
    if (fwd)
      hipLaunchKernelGGL((fft_z_r2c_mixed_kernel<T, M, ROWS>), grid, dim3(kMixedThreads), 0, st, fp, tw, twr, nrows,
                         nz_);
    else
      hipLaunchKernelGGL((fft_z_c2r_mixed_kernel<T, M, ROWS>), grid, dim3(kMixedThreads), 0, st, fp, tw, twr, nrows,
                         nz_, scale);
    return 0;
  }
}

template <typename T, int M>
int launch_z_mixed(int v, bool fwd, const FieldPtrs<T> &fp, int nf, const double2 *tw, const double2 *twr, int nrows,
                   int nz_, T scale, hipStream_t st) {
  constexpr int R0 = kMixedRows(M);
  switch (v) {
    case 0: return z_mixed_go<T, M, R0 / 2>(fwd, fp, nf, tw, twr, nrows, nz_, scale, st);
    case 1: return z_mixed_go<T, M, R0>(fwd, fp, nf, tw, twr, nrows, nz_, scale, st);
    case 2: return z_mixed_go<T, M, 2 * R0>(fwd, fp, nf, tw, twr, nrows, nz_, scale, st);
  }
  return -1;
}

// runtime length -> template
template <typename T, int SIGN>
int lines_mixed_rt(int n, int v, const FieldPtrs<T> &fp, int nf, const double2 *tw, int ncols, long estride,
                   long ostride, int nouter, hipStream_t st) {
  switch (n) {
    case 48: return launch_lines_mixed<T, 48, SIGN>(v, fp, nf, tw, ncols, estride, ostride, nouter, st);
    case 80: return launch_lines_mixed<T, 80, SIGN>(v, fp, nf, tw, ncols, estride, ostride, nouter, st);
    case 96: return launch_lines_mixed<T, 96, SIGN>(v, fp, nf, tw, ncols, estride, ostride, nouter, st);
    case 160: return launch_lines_mixed<T, 160, SIGN>(v, fp, nf, tw, ncols, estride, ostride, nouter, st);
    case 192: return launch_lines_mixed<T, 192, SIGN>(v, fp, nf, tw, ncols, estride, ostride, nouter, st);
  }
  return -1;
}

template <typename T>
int z_mixed_rt(int m, int v, bool fwd, const FieldPtrs<T> &fp, int nf, const double2 *tw, const double2 *twr,
               int nrows, int nz_, T scale, hipStream_t st) {
  switch (m) {
    case 24: return launch_z_mixed<T, 24>(v, fwd, fp, nf, tw, twr, nrows, nz_, scale, st);
    case 40: return launch_z_mixed<T, 40>(v, fwd, fp, nf, tw, twr, nrows, nz_, scale, st);
    case 48: return launch_z_mixed<T, 48>(v, fwd, fp, nf, tw, twr, nrows, nz_, scale, st);
    case 80: return launch_z_mixed<T, 80>(v, fwd, fp, nf, tw, twr, nrows, nz_, scale, st);
    case 96: return launch_z_mixed<T, 96>(v, fwd, fp, nf, tw, twr, nrows, nz_, scale, st);
  }
  return -1;
}

// The autotuner's choices, per pass kind, element size, length and the
// shape of the pass (columns and outer count for lines, rows for z).
enum MixedKind { MIX_LINES = 0, MIX_ZFWD = 1, MIX_ZINV = 2 };
struct MixedKey {
  int kind, esize, n, a, b;
  bool operator<(const MixedKey &o) const {
    return std::tie(kind, esize, n, a, b) < std::tie(o.kind, o.esize, o.n, o.a, o.b);
  }
};
std::mutex g_tune_mu;
std::map<MixedKey, int> g_tune;

int tuned_variant(const MixedKey &k, int dflt) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(k);
  return it == g_tune.end() ? dflt : it->second;
}

// one switch arm per supported line length
template <typename T, int SIGN>
int launch_lines(int n, const FieldPtrs<T> &fp, int nf, const double2 *tw, int ncols, long estride, long ostride,
                 int nouter, hipStream_t st, int col_cut = INT_MAX, int row_keep = -1) {
  if (row_keep < 0) row_keep = n;
  switch (n) {
    case 8: launch_lines_n<T, 8, SIGN>(fp, nf, tw, ncols, estride, ostride, nouter, st, col_cut, row_keep); return 0;
    case 16: launch_lines_n<T, 16, SIGN>(fp, nf, tw, ncols, estride, ostride, nouter, st, col_cut, row_keep); return 0;
    case 32: launch_lines_n<T, 32, SIGN>(fp, nf, tw, ncols, estride, ostride, nouter, st, col_cut, row_keep); return 0;
    case 64: launch_lines_n<T, 64, SIGN>(fp, nf, tw, ncols, estride, ostride, nouter, st, col_cut, row_keep); return 0;
    case 128: launch_lines_n<T, 128, SIGN>(fp, nf, tw, ncols, estride, ostride, nouter, st, col_cut, row_keep); return 0;
    case 256: launch_lines_n<T, 256, SIGN>(fp, nf, tw, ncols, estride, ostride, nouter, st, col_cut, row_keep); return 0;
    case 512: launch_lines_n<T, 512, SIGN>(fp, nf, tw, ncols, estride, ostride, nouter, st, col_cut, row_keep); return 0;
    case 1024: launch_lines_n<T, 1024, SIGN>(fp, nf, tw, ncols, estride, ostride, nouter, st, col_cut, row_keep); return 0;
  }
  // mixed radix: whole transforms only (no band limits)
  if (col_cut != INT_MAX || row_keep != n) return -1;
  const int v = tuned_variant(MixedKey{MIX_LINES, (int)sizeof(T), n, ncols, nouter}, kLinesDefault);
  return lines_mixed_rt<T, SIGN>(n, v, fp, nf, tw, ncols, estride, ostride, nouter, st);
}

template <typename T, int M>
void launch_z_n(bool fwd, const FieldPtrs<T> &fp, int nf, const double2 *tw, const double2 *twr, int nrows, int nz_,
                T scale, hipStream_t st, int kz_cut) {
  constexpr int R = kZRadix<T>(M), ROWS = kRows<T>(M);
  dim3 grid((nrows + ROWS - 1) / ROWS, 1, nf);
  if (fwd)
    hipLaunchKernelGGL((fft_z_r2c_kernel<T, M, R, ROWS>), grid, dim3(ROWS * M / R), 0, st, fp, tw, twr, nrows, nz_,
                       kz_cut);
  else
    hipLaunchKernelGGL((fft_z_c2r_kernel<T, M, R, ROWS>), grid, dim3(ROWS * M / R), 0, st, fp, tw, twr, nrows, nz_,
                       scale, kz_cut);
}

template <typename T>
int launch_z(int m, bool fwd, const FieldPtrs<T> &fp, int nf, const double2 *tw, const double2 *twr, int nrows,
             int nz_, T scale, hipStream_t st, int kz_cut = INT_MAX) {
  switch (m) {
    case 4: launch_z_n<T, 4>(fwd, fp, nf, tw, twr, nrows, nz_, scale, st, kz_cut); return 0;
    case 8: launch_z_n<T, 8>(fwd, fp, nf, tw, twr, nrows, nz_, scale, st, kz_cut); return 0;
    case 16: launch_z_n<T, 16>(fwd, fp, nf, tw, twr, nrows, nz_, scale, st, kz_cut); return 0;
    case 32: launch_z_n<T, 32>(fwd, fp, nf, tw, twr, nrows, nz_, scale, st, kz_cut); return 0;
    case 64: launch_z_n<T, 64>(fwd, fp, nf, tw, twr, nrows, nz_, scale, st, kz_cut); return 0;
    case 128: launch_z_n<T, 128>(fwd, fp, nf, tw, twr, nrows, nz_, scale, st, kz_cut); return 0;
    case 256: launch_z_n<T, 256>(fwd, fp, nf, tw, twr, nrows, nz_, scale, st, kz_cut); return 0;
    case 512: launch_z_n<T, 512>(fwd, fp, nf, tw, twr, nrows, nz_, scale, st, kz_cut); return 0;
  }
  if (kz_cut != INT_MAX) return -1;  // mixed radix: whole transforms only
  const int v = tuned_variant(MixedKey{fwd ? MIX_ZFWD : MIX_ZINV, (int)sizeof(T), m, nrows, 0}, kZDefault);
  return z_mixed_rt<T>(m, v, fwd, fp, nf, tw, twr, nrows, nz_, scale, st);
}

int upload_roots(spm_ctx *c, double2 **dst, int n, int count) {
  std::vector<double2> h(count);
  for (int t = 0; t < count; ++t) {
    long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
    h[t] = double2{(double)cosl(ang), (double)sinl(ang)};
  }
  SPM_HIP(c, hipMalloc(dst, sizeof(double2) * count));
  SPM_HIP(c, hipMemcpy(*dst, h.data(), sizeof(double2) * count, hipMemcpyHostToDevice));
  return 0;
}

}  // namespace

int fft_init_tables(spm_ctx *c) {
  const Grid &g = c->grid;
  int rc;
  if ((rc = upload_roots(c, &c->tw_x, g.nx, g.nx))) return rc;
  if ((rc = upload_roots(c, &c->tw_y, g.ny, g.ny))) return rc;
  if ((rc = upload_roots(c, &c->tw_z, g.nz / 2, g.nz / 2))) return rc;
  if ((rc = upload_roots(c, &c->tw_rz, g.nz, g.nz / 2 + 1))) return rc;
  return 0;
}

// Field buffers are double* handles; an FP32 context stores floats in them.
template <typename T>
static int fft_many_t(spm_ctx *c, double *const *a, int n, bool fwd) {
  const Grid &g = c->grid;
  for (int b = 0; b < n; b += 5) {
    FieldPtrs<T> fp{};
    const int nf = (n - b) < 5 ? (n - b) : 5;
    for (int q = 0; q < nf; ++q) fp.p[q] = reinterpret_cast<T *>(a[b + q]);
    const int nrows = g.nx * g.ny;
    const long h = g.hp;
    // every pass reads and writes each field once: 2 * NX*NY*NZ_ elements
    const double bytes = 2.0 * nf * (double)ref_count(g) * sizeof(T);
    int rc = 0;
    if (fwd) {
      tbeg(c);
      rc |= launch_z<T>(g.nz / 2, true, fp, nf, c->tw_z, c->tw_rz, nrows, g.nz_, (T)1, c->stream);
      tend(c, SPM_K_FFT_Z_R2C, bytes);
      tbeg(c);
      rc |= launch_lines<T, +1>(g.ny, fp, nf, c->tw_y, g.hnz, h, (long)g.ny * h, g.nx, c->stream);
      tend(c, SPM_K_FFT_Y, bytes);
      tbeg(c);
      rc |= launch_lines<T, +1>(g.nx, fp, nf, c->tw_x, g.ny * g.hp, (long)g.ny * h, 0, 1, c->stream);
      tend(c, SPM_K_FFT_X, bytes);
    } else {
      const T scale = (T)(1.0 / ((double)g.nx * g.ny * g.nz));
      tbeg(c);
      hipLaunchKernelGGL((ooura_planes_kernel<T>), dim3((unsigned)((2L * g.nx * g.ny + 255) / 256), nf), dim3(256), 0,
                         c->stream, fp, g.nx, g.ny, g.nz_, g.nz / 2);
      tend(c, SPM_K_SPECTRAL, 2.0 * nf * 2.0 * g.nx * g.ny * 2 * sizeof(T));
      tbeg(c);
      rc |= launch_lines<T, -1>(g.nx, fp, nf, c->tw_x, g.ny * g.hp, (long)g.ny * h, 0, 1, c->stream);
      tend(c, SPM_K_FFT_X, bytes);
      tbeg(c);
      rc |= launch_lines<T, -1>(g.ny, fp, nf, c->tw_y, g.hnz, h, (long)g.ny * h, g.nx, c->stream);
      tend(c, SPM_K_FFT_Y, bytes);
      tbeg(c);
      rc |= launch_z<T>(g.nz / 2, false, fp, nf, c->tw_z, c->tw_rz, nrows, g.nz_, scale, c->stream);
      tend(c, SPM_K_FFT_Z_C2R, bytes);
    }
    if (rc) {
      c->err = "unsupported FFT size";
      return SPM_ERR_UNSUPPORTED;
    }
    SPM_LAUNCH_CHECK(c);
  }
  return 0;
}

static int fft_many(spm_ctx *c, double *const *a, int n, bool fwd) {
  return c->prm.precision == SPM_FP32 ? fft_many_t<float>(c, a, n, fwd) : fft_many_t<double>(c, a, n, fwd);
}

int fft_r2c_many(spm_ctx *c, double *const *a, int n) { return fft_many(c, a, n, true); }
int fft_c2r_many(spm_ctx *c, double *const *a, int n) { return fft_many(c, a, n, false); }
int fft_r2c(spm_ctx *c, double *a) { return fft_many(c, &a, 1, true); }
int fft_c2r(spm_ctx *c, double *a) { return fft_many(c, &a, 1, false); }
int fft_r2c_f64(spm_ctx *c, double *a) { return fft_many_t<double>(c, &a, 1, true); }
int fft_c2r_many_f64(spm_ctx *c, double *const *a, int n) { return fft_many_t<double>(c, a, n, false); }

// ---------------------------------------------------------------------
// Autotuner of the mixed-radix launch shapes: every variant of every
// mixed-length pass of the context's grid is timed on the context's two
// scratch fields (work[0], work[1], zeroed first; they carry nothing
// between steps) and the fastest is kept for that pass shape, for every
// context of the process.  Shapes already in the table are not re-timed.
// ---------------------------------------------------------------------
namespace {
struct TuneLog {
  MixedKey key;
  std::vector<float> ms;  // per variant, -1 where it does not exist
  int pick;
};
std::vector<TuneLog> g_tune_log;

template <typename T>
int tune_mixed_t(spm_ctx *c) {
  const Grid &g = c->grid;
  FieldPtrs<T> fp{};
  fp.p[0] = reinterpret_cast<T *>(c->work[0]);
  fp.p[1] = reinterpret_cast<T *>(c->work[1]);
  const int nf = 2, nrows = g.nx * g.ny;
  const long h = g.hp;
  const T scale = (T)(1.0 / ((double)g.nx * g.ny * g.nz));
  struct Pass {
    MixedKey key;
    int nvar;
    std::function<int(int)> go;
  };
  std::vector<Pass> passes;
  const int es = (int)sizeof(T);
  if (!grid_pow2(g.nz)) {
    const int m = g.nz / 2;
    passes.push_back({MixedKey{MIX_ZFWD, es, m, nrows, 0}, kZVariants, [&, m](int v) {
                        return z_mixed_rt<T>(m, v, true, fp, nf, c->tw_z, c->tw_rz, nrows, g.nz_, (T)1, c->stream);
                      }});
    passes.push_back({MixedKey{MIX_ZINV, es, m, nrows, 0}, kZVariants, [&, m](int v) {
                        return z_mixed_rt<T>(m, v, false, fp, nf, c->tw_z, c->tw_rz, nrows, g.nz_, scale, c->stream);
                      }});
  }
  if (!grid_pow2(g.ny))
    passes.push_back({MixedKey{MIX_LINES, es, g.ny, g.hnz, g.nx}, kLinesVariants, [&](int v) {
                        return lines_mixed_rt<T, +1>(g.ny, v, fp, nf, c->tw_y, g.hnz, h, (long)g.ny * h, g.nx,
                                                     c->stream);
                      }});
  if (!grid_pow2(g.nx))
    passes.push_back({MixedKey{MIX_LINES, es, g.nx, g.ny * g.hp, 1}, kLinesVariants, [&](int v) {
                        return lines_mixed_rt<T, +1>(g.nx, v, fp, nf, c->tw_x, g.ny * g.hp, (long)g.ny * h, 0, 1,
                                                     c->stream);
                      }});
  if (passes.empty()) return 0;
  for (int q = 0; q < 2; ++q)
    if (int r0 = k_zero(c, c->work[q], c->nfield)) return r0;
  hipEvent_t e0, e1;
  SPM_HIP(c, hipEventCreate(&e0));
  SPM_HIP(c, hipEventCreate(&e1));
  constexpr int kReps = 10;
  int rc = 0;
  for (Pass &p : passes) {
    if (tuned_variant(p.key, -1) >= 0) continue;
    TuneLog log{p.key, std::vector<float>(p.nvar, -1.f), -1};
    for (int v = 0; v < p.nvar && rc == 0; ++v) {
      if (p.go(v) != 0) continue;  // no such variant for this length
      if (hipEventRecord(e0, c->stream) != hipSuccess) rc = SPM_ERR_HIP;
      for (int r = 0; r < kReps; ++r) (void)p.go(v);
      if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) rc = SPM_ERR_HIP;
      float ms = 0.f;
      if (rc == 0 && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = SPM_ERR_HIP;
      log.ms[v] = ms / kReps;
      if (log.pick < 0 || log.ms[v] < log.ms[log.pick]) log.pick = v;
    }
    if (hipGetLastError() != hipSuccess) rc = SPM_ERR_HIP;
    if (rc) break;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[p.key] = log.pick;
    g_tune_log.push_back(log);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) c->err = "autotuning the mixed-radix passes failed";
  return rc;
}
}  // namespace

int tune_mixed_transforms(spm_ctx *c) {
  return c->prm.precision == SPM_FP32 ? tune_mixed_t<float>(c) : tune_mixed_t<double>(c);
}

std::string tune_report() {
  static const char *kinds[] = {"lines", "z-r2c", "z-c2r"};
  std::lock_guard<std::mutex> lk(g_tune_mu);
  std::string out;
  char line[256];
  for (const TuneLog &l : g_tune_log) {
    int n = std::snprintf(line, sizeof line, "%s fp%d n=%d shape=%dx%d pick=%d ms=", kinds[l.key.kind],
                          8 * l.key.esize, l.key.n, l.key.a, l.key.b, l.pick);
    for (size_t v = 0; v < l.ms.size() && n < (int)sizeof line - 16; ++v)
      n += std::snprintf(line + n, sizeof line - n, "%s%.4f", v ? "," : "", l.ms[v]);
    out += line;
    out += "\n";
  }
  return out;
}

}  // namespace spm

namespace spm {

#define RC_(x)            \
  do {                    \
    const int rc_ = (x);  \
    if (rc_) return rc_;  \
  } while (0)

// Single passes, used by the fused schedule (fused.hip supplies the x and
// z passes that carry pointwise work).  kind: 0 = y pass, 1 = z forward,
// 2 = z inverse.  They carry only the modes the 2/3 rule keeps: kz <
// TRN_Z columns and entries, and in the y direction |ky| < TRN_Y (taken as
// zero on the way into an inverse, not stored by a forward pass).
template <typename T>
static int fft_pass_t(spm_ctx *c, double *const *a, int n, int kind, int sign) {
  const Grid &g = c->grid;
  for (int b = 0; b < n; b += 5) {
    FieldPtrs<T> fp{};
    const int nf = (n - b) < 5 ? (n - b) : 5;
    for (int q = 0; q < nf; ++q) fp.p[q] = reinterpret_cast<T *>(a[b + q]);
    const long h = g.hp;
    // bytes moved with the band limits: field fractions kept in kz and ky
    const double fz = (double)g.trn_z / g.hnz, fy = (double)(2 * g.trn_y - 1) / g.ny;
    const double fb = nf * (double)ref_count(g) * sizeof(T);
    int rc;
    tbeg(c);
    if (kind == 0) {
      rc = sign > 0 ? launch_lines<T, +1>(g.ny, fp, nf, c->tw_y, g.hnz, h, (long)g.ny * h, g.nx, c->stream, g.trn_z,
                                          g.trn_y)
                    : launch_lines<T, -1>(g.ny, fp, nf, c->tw_y, g.hnz, h, (long)g.ny * h, g.nx, c->stream, g.trn_z,
                                          g.trn_y);
      tend(c, SPM_K_FFT_Y, fb * fz * (1.0 + fy));
    } else if (kind == 1) {
      rc = launch_z<T>(g.nz / 2, true, fp, nf, c->tw_z, c->tw_rz, g.nx * g.ny, g.nz_, (T)1, c->stream, g.trn_z);
      tend(c, SPM_K_FFT_Z_R2C, fb * (1.0 + fz));
    } else {
      const T scale = (T)(1.0 / ((double)g.nx * g.ny * g.nz));
      rc = launch_z<T>(g.nz / 2, false, fp, nf, c->tw_z, c->tw_rz, g.nx * g.ny, g.nz_, scale, c->stream, g.trn_z);
      tend(c, SPM_K_FFT_Z_C2R, fb * (1.0 + fz));
    }
    if (rc) {
      c->err = "unsupported FFT size";
      return SPM_ERR_UNSUPPORTED;
    }
    SPM_LAUNCH_CHECK(c);
  }
  return 0;
}

static int fft_pass(spm_ctx *c, double *const *a, int n, int kind, int sign) {
  return c->prm.precision == SPM_FP32 ? fft_pass_t<float>(c, a, n, kind, sign)
                                      : fft_pass_t<double>(c, a, n, kind, sign);
}

// Whole transforms with band-limited passes: the inverse of spectra that
// vanish outside the 2/3 box, and the forward transform whose modes outside
// the box are not read on (they are left unspecified).  The x pass runs over
// the kz < TRN_Z columns of every ky row (as NY outer blocks of NZ/2+1
// columns), the y and z passes are those of the fused schedule.
template <typename T>
static int fft_band_t(spm_ctx *c, double *const *a, int n, bool fwd, bool with_z = true) {
  const Grid &g = c->grid;
  for (int b = 0; b < n; b += 5) {
    const int nf = (n - b) < 5 ? (n - b) : 5;
    FieldPtrs<T> fp{};
    for (int q = 0; q < nf; ++q) fp.p[q] = reinterpret_cast<T *>(a[b + q]);
    const long h = g.hp;
    const double bytes = 2.0 * nf * (double)ref_count(g) * sizeof(T);
    const double fz = (double)g.trn_z / g.hnz;
    int rc = 0;
    if (fwd) {
      if (with_z) RC_(fft_pass_t<T>(c, a + b, nf, 1, +1));
      RC_(fft_pass_t<T>(c, a + b, nf, 0, +1));
      tbeg(c);
      rc |= launch_lines<T, +1>(g.nx, fp, nf, c->tw_x, g.hnz, (long)g.ny * h, h, g.ny, c->stream, g.trn_z);
      tend(c, SPM_K_FFT_X, bytes * fz);
    } else {
      tbeg(c);
      hipLaunchKernelGGL((ooura_planes_kernel<T>), dim3((unsigned)((2L * g.nx * g.ny + 255) / 256), nf), dim3(256), 0,
                         c->stream, fp, g.nx, g.ny, g.nz_, g.nz / 2);
      tend(c, SPM_K_SPECTRAL, 2.0 * nf * 2.0 * g.nx * g.ny * 2 * sizeof(T));
      tbeg(c);
      rc |= launch_lines<T, -1>(g.nx, fp, nf, c->tw_x, g.hnz, (long)g.ny * h, h, g.ny, c->stream, g.trn_z);
      tend(c, SPM_K_FFT_X, bytes * fz);
      if (!rc) RC_(fft_pass_t<T>(c, a + b, nf, 0, -1));
      if (!rc && with_z) RC_(fft_pass_t<T>(c, a + b, nf, 2, -1));
    }
    if (rc) {
      c->err = "unsupported FFT size";
      return SPM_ERR_UNSUPPORTED;
    }
    SPM_LAUNCH_CHECK(c);
  }
  return 0;
}

int fft_r2c_many_band(spm_ctx *c, double *const *a, int n) {
  return c->prm.precision == SPM_FP32 ? fft_band_t<float>(c, a, n, true) : fft_band_t<double>(c, a, n, true);
}
int fft_c2r_many_band(spm_ctx *c, double *const *a, int n) {
  return c->prm.precision == SPM_FP32 ? fft_band_t<float>(c, a, n, false) : fft_band_t<double>(c, a, n, false);
}
int fft_c2r_band_xy(spm_ctx *c, double *const *a, int n) {
  return c->prm.precision == SPM_FP32 ? fft_band_t<float>(c, a, n, false, false)
                                      : fft_band_t<double>(c, a, n, false, false);
}
int fft_r2c_band_yx(spm_ctx *c, double *const *a, int n) {
  return c->prm.precision == SPM_FP32 ? fft_band_t<float>(c, a, n, true, false)
                                      : fft_band_t<double>(c, a, n, true, false);
}

int fft_pass_y(spm_ctx *c, double *const *a, int n, int sign) { return fft_pass(c, a, n, 0, sign); }
int fft_pass_z_r2c(spm_ctx *c, double *const *a, int n) { return fft_pass(c, a, n, 1, +1); }
int fft_pass_z_c2r(spm_ctx *c, double *const *a, int n) { return fft_pass(c, a, n, 2, -1); }

// ---------------------------------------------------------------------
// Transforms of one caller-owned FP64 field with no context (the
// reference's A2a_k / A_k2a are free functions on a caller's array,
// fft_wrapper.h:432-631).  The roots of unity live in a workspace that
// every call rewrites: four tables, each starting on a 256-byte boundary,
//   x: exp(+2 pi i t/NX), t < NX      y: exp(+2 pi i t/NY), t < NY
//   z: exp(+2 pi i t/(NZ/2)), t < NZ/2     rz: exp(+2 pi i k/NZ), k <= NZ/2
// computed on the device (sincospi of 2t/n: exact for powers of two, one
// rounding of the argument for the mixed lengths), so the
// call queues kernels only.
// ---------------------------------------------------------------------
namespace {
inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct RootTables {
  size_t off[4];  // byte offsets of x, y, z, rz
  int len[4];
  size_t bytes;
};

RootTables root_tables(int nx, int ny, int nz) {
  RootTables t{};
  t.len[0] = nx;
  t.len[1] = ny;
  t.len[2] = nz / 2;
  t.len[3] = nz / 2 + 1;
  size_t o = 0;
  for (int q = 0; q < 4; ++q) {
    t.off[q] = o;
    o += align256(sizeof(double2) * t.len[q]);
  }
  t.bytes = o;
  return t;
}

bool fft_size_ok(int n) { return grid_size_ok(n); }
}  // namespace

// one thread per table entry over the four tables
__global__ void roots_kernel(char *ws, RootTables t, int nx, int ny, int nz) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  int q = 0, e = g;
  while (q < 4 && e >= t.len[q]) e -= t.len[q++];
  if (q >= 4) return;
  const int n = q == 0 ? nx : (q == 1 ? ny : (q == 2 ? nz / 2 : nz));
  double sn, cs;
  sincospi(2.0 * (double)e / (double)n, &sn, &cs);
  reinterpret_cast<double2 *>(ws + t.off[q])[e] = double2{cs, sn};
}

size_t fft3d_workspace_bytes(int nx, int ny, int nz) {
  if (!fft_size_ok(nx) || !fft_size_ok(ny) || !fft_size_ok(nz)) return 0;
  return root_tables(nx, ny, nz).bytes;
}

// Queues one transform of `a` on `st` using the roots in `ws` (at least
// fft3d_workspace_bytes bytes of device memory).  direction +1: A2a_k,
// -1: A_k2a.
int fft3d_raw(int nx, int ny, int nz, int direction, double *a, void *ws, hipStream_t st) {
  if (!fft_size_ok(nx) || !fft_size_ok(ny) || !fft_size_ok(nz)) {
    global_error() = "transform sizes must be powers of two in [8, 1024] or 48, 80, 96, 160, 192";
    return SPM_ERR_INVALID;
  }
  if (!a || !ws || (direction != 1 && direction != -1)) {
    global_error() = "null field or workspace, or direction not +1 / -1";
    return SPM_ERR_INVALID;
  }
  const RootTables t = root_tables(nx, ny, nz);
  char *w = static_cast<char *>(ws);
  const int nent = t.len[0] + t.len[1] + t.len[2] + t.len[3];
  hipLaunchKernelGGL(roots_kernel, dim3((nent + 255) / 256), dim3(256), 0, st, w, t, nx, ny, nz);
  const double2 *tw_x = reinterpret_cast<const double2 *>(w + t.off[0]);
  const double2 *tw_y = reinterpret_cast<const double2 *>(w + t.off[1]);
  const double2 *tw_z = reinterpret_cast<const double2 *>(w + t.off[2]);
  const double2 *tw_rz = reinterpret_cast<const double2 *>(w + t.off[3]);
  FieldPtrs<double> fp{};
  fp.p[0] = a;
  const int hnz = nz / 2 + 1, nz_ = 2 * hnz, nrows = nx * ny;
  int rc = 0;
  if (direction > 0) {
    rc |= launch_z<double>(nz / 2, true, fp, 1, tw_z, tw_rz, nrows, nz_, 1.0, st);
    rc |= launch_lines<double, +1>(ny, fp, 1, tw_y, hnz, hnz, (long)ny * hnz, nx, st);
    rc |= launch_lines<double, +1>(nx, fp, 1, tw_x, ny * hnz, (long)ny * hnz, 0, 1, st);
  } else {
    hipLaunchKernelGGL((ooura_planes_kernel<double>), dim3((unsigned)((2L * nx * ny + 255) / 256), 1), dim3(256), 0,
                       st, fp, nx, ny, nz_, nz / 2);
    rc |= launch_lines<double, -1>(nx, fp, 1, tw_x, ny * hnz, (long)ny * hnz, 0, 1, st);
    rc |= launch_lines<double, -1>(ny, fp, 1, tw_y, hnz, hnz, (long)ny * hnz, nx, st);
    rc |= launch_z<double>(nz / 2, false, fp, 1, tw_z, tw_rz, nrows, nz_, 1.0 / ((double)nx * ny * nz), st);
  }
  if (rc) {
    global_error() = "unsupported transform size";
    return SPM_ERR_UNSUPPORTED;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    global_error() = std::string("kernel launch: ") + hipGetErrorString(e);
    return SPM_ERR_HIP;
  }
  return 0;
}

}  // namespace spm

extern "C" {

size_t spm_fft_workspace_size(int nx, int ny, int nz) { return spm::fft3d_workspace_bytes(nx, ny, nz); }

int spm_fft3d(int nx, int ny, int nz, int direction, double *field, void *workspace, size_t workspace_bytes,
              void *stream) {
  spm::global_error().clear();
  const size_t need = spm::fft3d_workspace_bytes(nx, ny, nz);
  if (need == 0) {
    spm::global_error() = "transform sizes must be powers of two in [8, 1024] or 48, 80, 96, 160, 192";
    return SPM_ERR_INVALID;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (workspace) {
    if (workspace_bytes < need) {
      spm::global_error() = "workspace smaller than spm_fft_workspace_size()";
      return SPM_ERR_INVALID;
    }
    return spm::fft3d_raw(nx, ny, nz, direction, field, workspace, st);
  }
  // no workspace given: allocate one for this call and free it once the
  // stream has finished with it
  void *ws = nullptr;
  hipError_t e = hipMalloc(&ws, need);
  if (e != hipSuccess) {
    spm::global_error() = std::string("hipMalloc: ") + hipGetErrorString(e);
    return SPM_ERR_OOM;
  }
  int rc = spm::fft3d_raw(nx, ny, nz, direction, field, ws, st);
  e = hipStreamSynchronize(st);
  if (!rc && e != hipSuccess) {
    spm::global_error() = std::string("hipStreamSynchronize: ") + hipGetErrorString(e);
    rc = SPM_ERR_HIP;
  }
  (void)hipFree(ws);
  return rc;
}

}  // extern "C"
