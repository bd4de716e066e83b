// fft_core.h -- in-LDS Stockham FFT building blocks shared by the plain
// FFT passes (fft.hip) and the fused passes of the step (fused.hip).
#pragma once

#include "spm_internal.h"

namespace spm {

// Launch shapes.  Each may be overridden at compile time (-D...) to
// measure another shape; the defaults are the measured choices for gfx950
// at 256^3 (FP64 fields move twice the bytes per point of FP32 ones, so
// some choices differ by precision).
#ifndef SPM_LINE_POINTS_F64
#define SPM_LINE_POINTS_F64 2048  // points (line length x columns) per line-pass block, FP64 fields
#endif
#ifndef SPM_LINE_POINTS_F32
#define SPM_LINE_POINTS_F32 4096
#endif
#ifndef SPM_LINE_RADIX
#define SPM_LINE_RADIX 8  // Stockham radix of the plain y/x passes
#endif
#ifndef SPM_Z_RADIX_F64
#define SPM_Z_RADIX_F64 4  // Stockham radix of the plain z passes, FP64 fields
#endif
#ifndef SPM_Z_RADIX_F32
#define SPM_Z_RADIX_F32 8
#endif
#ifndef SPM_Z_POINTS_F64
#define SPM_Z_POINTS_F64 2048  // complex points (NZ/2 x rows) per z-pass block
#endif
#ifndef SPM_Z_POINTS_F32
#define SPM_Z_POINTS_F32 2048
#endif
#ifndef SPM_FUSED_RADIX
#define SPM_FUSED_RADIX 4  // radix of the fused x passes (more live registers)
#endif
#ifndef SPM_ZADV_RADIX
#define SPM_ZADV_RADIX 4  // radix of the fused z pass (advection products)
#endif
#ifndef SPM_ZADV_POINTS_F64
#define SPM_ZADV_POINTS_F64 1024  // complex points per fused z-pass block, FP64 fields
#endif
#ifndef SPM_ZADV_POINTS_F32
#define SPM_ZADV_POINTS_F32 2048
#endif
#ifndef SPM_FUSED_X_POINTS
#define SPM_FUSED_X_POINTS 2048  // points (line length x columns) per fused x-pass block
#endif
#ifndef SPM_XREMAP_F64
#define SPM_XREMAP_F64 0  // XCD-grouped tiles in the fused x passes, FP64 (tiles are whole lines)
#endif
#ifndef SPM_XREMAP_F32
#define SPM_XREMAP_F32 1  // FP32 (two tiles share each 128-byte line)
#endif

constexpr int clamp_rows(int r) { return r > 32 ? 32 : (r < 1 ? 1 : r); }
// columns per line-pass block (at most 32)
template <typename T>
constexpr int kLinesC(int n) {
  constexpr int pts = sizeof(T) == 8 ? SPM_LINE_POINTS_F64 : SPM_LINE_POINTS_F32;
  return n >= pts ? 1 : (pts / n > 32 ? 32 : pts / n);
}
constexpr int kRadix(int n) { return n >= SPM_LINE_RADIX ? SPM_LINE_RADIX : n; }
// rows and radix of a plain z-pass block of m complex points
template <typename T>
constexpr int kRows(int m) {
  return clamp_rows((sizeof(T) == 8 ? SPM_Z_POINTS_F64 : SPM_Z_POINTS_F32) / m);
}
template <typename T>
constexpr int kZRadix(int m) {
  return m >= (sizeof(T) == 8 ? SPM_Z_RADIX_F64 : SPM_Z_RADIX_F32) ? (sizeof(T) == 8 ? SPM_Z_RADIX_F64 : SPM_Z_RADIX_F32)
                                                                   : m;
}
constexpr int kZadvR(int n) { return n >= SPM_ZADV_RADIX ? SPM_ZADV_RADIX : n; }
template <typename T>
constexpr int kZadvRows(int m) {
  return clamp_rows((sizeof(T) == 8 ? SPM_ZADV_POINTS_F64 : SPM_ZADV_POINTS_F32) / m);
}
constexpr int kFusedR(int n) { return n >= SPM_FUSED_RADIX ? SPM_FUSED_RADIX : n; }
constexpr int kFusedC(int n) {
  return n >= SPM_FUSED_X_POINTS ? 1 : (SPM_FUSED_X_POINTS / n > 32 ? 32 : SPM_FUSED_X_POINTS / n);
}
template <typename T>
__device__ __forceinline__ int fused_x_tile() {
  return (sizeof(T) == 8 ? SPM_XREMAP_F64 : SPM_XREMAP_F32) ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
}

template <typename C2, int SIGN>
__device__ __forceinline__ C2 mul_si(C2 a) {  // a * (SIGN i)
  return SIGN > 0 ? C2{-a.y, a.x} : C2{a.y, -a.x};
}

template <typename C2, int SIGN>
__device__ __forceinline__ void dft2(C2 *a) {
  C2 t = a[0];
  a[0] = cadd(t, a[1]);
  a[1] = csub(t, a[1]);
}

template <typename C2, int SIGN>
__device__ __forceinline__ void dft4v(C2 &a0, C2 &a1, C2 &a2, C2 &a3) {
  C2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3);
  C2 t3 = mul_si<C2, SIGN>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a1 = cadd(t1, t3);
  a2 = csub(t0, t2);
  a3 = csub(t1, t3);
}

template <typename C2, int SIGN>
__device__ __forceinline__ void dft4(C2 *a) {
  dft4v<C2, SIGN>(a[0], a[1], a[2], a[3]);
}

template <typename C2, int SIGN>
__device__ __forceinline__ void dft8(C2 *a) {
  using R = decltype(a[0].x);
  C2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  C2 o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4v<C2, SIGN>(e0, e1, e2, e3);
  dft4v<C2, SIGN>(o0, o1, o2, o3);
  const R r = (R)0.70710678118654752440084436210485;
  const R s = (R)SIGN;
  o1 = C2{r * (o1.x - s * o1.y), r * (o1.y + s * o1.x)};  // * exp(SIGN i pi/4)
  o2 = mul_si<C2, SIGN>(o2);                               // * exp(SIGN i pi/2)
  o3 = C2{-r * (o3.x + s * o3.y), r * (s * o3.x - o3.y)};  // * exp(SIGN 3 i pi/4)
  a[0] = cadd(e0, o0);
  a[4] = csub(e0, o0);
  a[1] = cadd(e1, o1);
  a[5] = csub(e1, o1);
  a[2] = cadd(e2, o2);
  a[6] = csub(e2, o2);
  a[3] = cadd(e3, o3);
  a[7] = csub(e3, o3);
}

// radix 3: y_m = sum_r a_r w^(rm), w = exp(SIGN 2 pi i / 3)
template <typename C2, int SIGN>
__device__ __forceinline__ void dft3(C2 *a) {
  using R = decltype(a[0].x);
  const R h = (R)(SIGN * 0.86602540378443864676372317075294);  // SIGN sin(2 pi / 3)
  const C2 t1 = cadd(a[1], a[2]);
  const C2 t2{a[0].x - (R)0.5 * t1.x, a[0].y - (R)0.5 * t1.y};
  const C2 d = csub(a[1], a[2]);
  const C2 id{-h * d.y, h * d.x};  // i SIGN sin(2 pi/3) (a1 - a2)
  a[0] = cadd(a[0], t1);
  a[1] = cadd(t2, id);
  a[2] = csub(t2, id);
}

// radix 5: y_m = sum_r a_r w^(rm), w = exp(SIGN 2 pi i / 5), from the
// symmetric sums a_1 + a_4, a_2 + a_3 and differences a_1 - a_4, a_2 - a_3
template <typename C2, int SIGN>
__device__ __forceinline__ void dft5(C2 *a) {
  using R = decltype(a[0].x);
  const R c1 = (R)0.30901699437494742410229341718282, c2 = (R)-0.80901699437494742410229341718282;
  const R s1 = (R)(SIGN * 0.95105651629515357211643933337938), s2 = (R)(SIGN * 0.58778525229247312916870595463907);
  const C2 b1 = cadd(a[1], a[4]), b2 = cadd(a[2], a[3]);
  const C2 d1 = csub(a[1], a[4]), d2 = csub(a[2], a[3]);
  const C2 e1{a[0].x + c1 * b1.x + c2 * b2.x, a[0].y + c1 * b1.y + c2 * b2.y};
  const C2 e2{a[0].x + c2 * b1.x + c1 * b2.x, a[0].y + c2 * b1.y + c1 * b2.y};
  const C2 f1{s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y};
  const C2 f2{s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y};
  a[0] = cadd(a[0], cadd(b1, b2));
  a[1] = C2{e1.x - f1.y, e1.y + f1.x};  // e1 + i f1
  a[4] = C2{e1.x + f1.y, e1.y - f1.x};
  a[2] = C2{e2.x - f2.y, e2.y + f2.x};
  a[3] = C2{e2.x + f2.y, e2.y - f2.x};
}

template <typename C2, int RS, int SIGN>
__device__ __forceinline__ void dft(C2 *a) {
  if constexpr (RS == 2) dft2<C2, SIGN>(a);
  else if constexpr (RS == 3) dft3<C2, SIGN>(a);
  else if constexpr (RS == 4) dft4<C2, SIGN>(a);
  else if constexpr (RS == 5) dft5<C2, SIGN>(a);
  else dft8<C2, SIGN>(a);
}

// ---------------------------------------------------------------------
// Mixed-radix lengths (factors 2, 3, 5): out-of-place Stockham stages on two
// LDS buffers of N points x C columns ([point][column]), radices 4, 2, 3,
// 5 in that order of preference; each thread loops over butterflies.
// ---------------------------------------------------------------------
constexpr int mixed_radix(int rem) {
  return rem % 4 == 0 ? 4 : (rem % 2 == 0 ? 2 : (rem % 3 == 0 ? 3 : (rem % 5 == 0 ? 5 : 0)));
}
constexpr bool radix235(int n) {
  if (n < 1) return false;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  return n == 1;
}
constexpr int mixed_nstages(int n) {
  int s = 0;
  for (int rem = n; rem > 1; rem /= mixed_radix(rem)) ++s;
  return s;
}

// stages from span NS on; src holds the data, the result ends in src when
// the number of remaining stages is even, else in dst
template <typename T, int N, int NS, int C, int SIGN>
__device__ __forceinline__ void mixed_stages(cplx<T> *src, cplx<T> *dst, const double2 *__restrict__ tw, int tid,
                                             int nt) {
  using C2 = cplx<T>;
  if constexpr (NS < N) {
    constexpr int RS = mixed_radix(N / NS);
    static_assert(RS > 0, "lengths with factors 2, 3 and 5 only");
    constexpr int NB = N / RS;
    for (int w = tid; w < NB * C; w += nt) {
      const int c = w % C, j = w / C, k = j % NS;
      C2 a[RS];
#pragma unroll
      for (int r = 0; r < RS; ++r) a[r] = src[(j + r * NB) * C + c];
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < RS; ++r) {
          const double2 t = tw[(r * k * (N / (NS * RS))) % N];
          a[r] = cmul(a[r], C2{(T)t.x, (T)(SIGN * t.y)});
        }
      }
      dft<C2, RS, SIGN>(a);
      const int base = (j / NS) * NS * RS + k;
#pragma unroll
      for (int r = 0; r < RS; ++r) dst[(base + r * NS) * C + c] = a[r];
    }
    __syncthreads();
    mixed_stages<T, N, NS * mixed_radix(N / NS), C, SIGN>(dst, src, tw, tid, nt);
  }
}

// One radix-RS Stockham stage of span NS on C columns of N points held in
// LDS at lds[point * CP + column].  The block has C * N / R threads; each
// performs R / RS butterflies.  Twiddles come from tw[t] = exp(+2 pi i t/N).
template <typename T, int N, int R, int RS, int NS, int C, int CP, int SIGN>
__device__ __forceinline__ void stockham_stage(cplx<T> *lds, const double2 *__restrict__ tw, int tid) {
  using C2 = cplx<T>;
  constexpr int NB = N / RS;  // butterflies per column
  constexpr int PER = R / RS;
  const int c = tid % C;
  const int j0 = tid / C;
  C2 a[R];
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    const int j = j0 + s * (N / R);
#pragma unroll
    for (int r = 0; r < RS; ++r) a[s * RS + r] = lds[(j + r * NB) * CP + c];
    if constexpr (NS > 1) {
      // w^r for r = 1..RS-1 from one table load: w^2 and w^4 by squaring,
      // odd powers by one more product (at most three roundings deep)
      const int k = j % NS;
      const double2 t1 = tw[k * (N / (NS * RS))];
      C2 w[RS];
      w[1] = C2{(T)t1.x, (T)(SIGN * t1.y)};
#pragma unroll
      for (int r = 2; r < RS; ++r) w[r] = (r % 2 == 0) ? cmul(w[r / 2], w[r / 2]) : cmul(w[r - 1], w[1]);
#pragma unroll
      for (int r = 1; r < RS; ++r) a[s * RS + r] = cmul(a[s * RS + r], w[r]);
    }
    dft<C2, RS, SIGN>(&a[s * RS]);
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    const int j = j0 + s * (N / R);
    const int k = j % NS;
    const int base = (j / NS) * NS * RS + k;
#pragma unroll
    for (int r = 0; r < RS; ++r) lds[(base + r * NS) * CP + c] = a[s * RS + r];
  }
  __syncthreads();
}

template <typename T, int N, int R, int C, int CP, int SIGN, int NS>
__device__ __forceinline__ void fft_stages(cplx<T> *lds, const double2 *__restrict__ tw, int tid) {
  if constexpr (NS < N) {
    constexpr int RS = (N / NS >= R) ? R : N / NS;
    stockham_stage<T, N, R, RS, NS, C, CP, SIGN>(lds, tw, tid);
    fft_stages<T, N, R, C, CP, SIGN, NS * RS>(lds, tw, tid);
  }
}

// The same stages with the first one reading its inputs through a load
// functor and the last one handing its outputs to a store functor instead
// of going through LDS: for a line pass that saves the initial copy into LDS
// and the final copy out of it, and each thread issues its first-stage loads
// back to back.  Functors see (point, column, q) where q = s*RS + r is the
// compile-time slot of the value in the thread's registers.
//   ld(point, column, q) -> cplx<T>
//   st(point, column, q, value)
// The first stage synchronises before writing LDS, so a caller may reuse
// the LDS of a previous transform as soon as its own reads are issued.
template <typename T, int N, int R, int RS, int NS, int C, int CP, int SIGN, bool IN_G, bool OUT_G, class LD,
          class ST>
__device__ __forceinline__ void stockham_stage_io(cplx<T> *lds, const double2 *__restrict__ tw, int tid, LD &ld,
                                                  ST &st) {
  using C2 = cplx<T>;
  constexpr int NB = N / RS;
  constexpr int PER = R / RS;
  const int c = tid % C;
  const int j0 = tid / C;
  C2 a[R];
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    const int j = j0 + s * (N / R);
#pragma unroll
    for (int r = 0; r < RS; ++r) {
      if constexpr (IN_G)
        a[s * RS + r] = ld(j + r * NB, c, s * RS + r);
      else
        a[s * RS + r] = lds[(j + r * NB) * CP + c];
    }
    if constexpr (NS > 1) {
      const int k = j % NS;
      const double2 t1 = tw[k * (N / (NS * RS))];
      C2 w[RS];
      w[1] = C2{(T)t1.x, (T)(SIGN * t1.y)};
#pragma unroll
      for (int r = 2; r < RS; ++r) w[r] = (r % 2 == 0) ? cmul(w[r / 2], w[r / 2]) : cmul(w[r - 1], w[1]);
#pragma unroll
      for (int r = 1; r < RS; ++r) a[s * RS + r] = cmul(a[s * RS + r], w[r]);
    }
    dft<C2, RS, SIGN>(&a[s * RS]);
  }
  if constexpr (OUT_G) {
    if constexpr (IN_G) __syncthreads();  // single stage in place: all reads before any write
#pragma unroll
    for (int s = 0; s < PER; ++s) {
      const int j = j0 + s * (N / R);
      const int base = (j / NS) * NS * RS + j % NS;
#pragma unroll
      for (int r = 0; r < RS; ++r) st(base + r * NS, c, s * RS + r, a[s * RS + r]);
    }
  } else {
    __syncthreads();  // every read of this LDS (this stage's or a previous transform's) is done
#pragma unroll
    for (int s = 0; s < PER; ++s) {
      const int j = j0 + s * (N / R);
      const int base = (j / NS) * NS * RS + j % NS;
#pragma unroll
      for (int r = 0; r < RS; ++r) lds[(base + r * NS) * CP + c] = a[s * RS + r];
    }
    __syncthreads();
  }
}

template <typename T, int N, int R, int C, int CP, int SIGN, int NS, bool FIRST_G, bool LAST_G, class LD, class ST>
__device__ __forceinline__ void fft_stages_io(cplx<T> *lds, const double2 *__restrict__ tw, int tid, LD &ld, ST &st) {
  if constexpr (NS < N) {
    constexpr int RS = (N / NS >= R) ? R : N / NS;
    constexpr bool in_g = FIRST_G && NS == 1;
    constexpr bool out_g = LAST_G && NS * RS == N;
    stockham_stage_io<T, N, R, RS, NS, C, CP, SIGN, in_g, out_g>(lds, tw, tid, ld, st);
    fft_stages_io<T, N, R, C, CP, SIGN, NS * RS, FIRST_G, LAST_G>(lds, tw, tid, ld, st);
  }
}

// Tiles of a line pass.  Column (o, col) starts at complex element
// o * ostride + col.  Regular tiles take C adjacent columns of one outer
// index o.  When an outer index has 1 + a multiple of C columns (the
// NZ/2+1 columns of a y pass), its last column goes to edge tiles that take
// that column of C consecutive outer indices, so no regular tile runs with
// one column out of C.
struct LineTiles {
  long ostride;   // complex elements between outer indices
  int ncols;      // columns per outer index in regular tiles
  int per_outer;  // regular tiles per outer index
  int nouter;     // outer indices
  int main;       // regular tiles
  int tiles;      // regular + edge tiles
};

inline LineTiles make_line_tiles(int ncols, int nouter, long ostride, int C) {
  LineTiles t;
  const bool edge = nouter > 1 && ncols > C && (ncols - 1) % C == 0;
  t.ostride = ostride;
  t.ncols = edge ? ncols - 1 : ncols;
  t.per_outer = (t.ncols + C - 1) / C;
  t.nouter = nouter;
  t.main = nouter * t.per_outer;
  t.tiles = t.main + (edge ? (nouter + C - 1) / C : 0);
  return t;
}

// index within its outer index (the kz of a y pass) of a tile's first column
__device__ __forceinline__ int line_first_col(const LineTiles &t, int tile, int C) {
  return tile < t.main ? (tile % t.per_outer) * C : t.ncols;
}

// element offset of column c of a tile, or -1 where the tile has no column c
__device__ __forceinline__ long line_column(const LineTiles &t, int tile, int c, int C) {
  if (tile < t.main) {
    const int o = tile / t.per_outer;
    const int col = (tile - o * t.per_outer) * C + c;
    return col < t.ncols ? o * t.ostride + col : -1;
  }
  const int o = (tile - t.main) * C + c;
  return o < t.nouter ? o * t.ostride + t.ncols : -1;
}


// Loads through a buffer resource over a whole field: an offset past the
// field's end reads as zero (the hardware's range check), so a load that
// must read zero where a column or mode is not kept needs no branch around
// it.  With no branch the compiler counts the loads in flight exactly, and
// loads issued ahead for the next field stay in flight while the current
// field's complete (a load guarded by `if` is drained at the join after it).
// Fields must be under 4 GiB (the resource's byte count and offsets are 32
// bits); the fused schedule checks that.
struct FieldBuf {
  __amdgpu_buffer_rsrc_t r;
};
template <typename T>
__device__ __forceinline__ FieldBuf field_buf(const T *p, size_t n) {
  return FieldBuf{__builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(p), (short)0, (int)(unsigned)(n * sizeof(T)),
                                                    0x00020000)};
}
// the complex value at element offset `off`, or zero where `ok` is false
template <typename T>
__device__ __forceinline__ cplx<T> buf_load(const FieldBuf &b, long off, bool ok) {
  const unsigned o = ok ? (unsigned)(off * (long)sizeof(T)) : 0xfffffff0u;
  if constexpr (sizeof(T) == 8) {
    return __builtin_bit_cast(cplx<T>, __builtin_amdgcn_raw_buffer_load_b128(b.r, (int)o, 0, 0));
  } else {
    return __builtin_bit_cast(cplx<T>, __builtin_amdgcn_raw_buffer_load_b64(b.r, (int)o, 0, 0));
  }
}

template <typename T>
struct FieldPtrs {
  T *p[5];
};

// ---------------------------------------------------------------------
// Direct-to-LDS tile loads (FP64 complex, 16 bytes per lane).  gfx950's
// buffer_load_dwordx4 ... lds writes lane l's 16 bytes to LDS at the wave's
// base address (M0) + 16 l, so a load in flight holds no VGPRs: a block can
// keep the next spectrum's whole tile in flight while it transforms the
// current one in another LDS buffer.
// ---------------------------------------------------------------------
// Hides v's value from the optimizer, so that what is computed from it is
// computed where it is used instead of being kept live from an earlier use.
template <typename V>
__device__ __forceinline__ V opaque(V v) {
  asm volatile("" : "+v"(v));
  return v;
}
// The same, placed after every memory access before it: what is computed
// from v starts after those accesses (a loop whose iterations would
// otherwise be interleaved, each holding many registers, runs them in turn).
template <typename V>
__device__ __forceinline__ V opaque_after_memory(V v) {
  asm volatile("" : "+v"(v) : : "memory");
  return v;
}

using lds_void = __attribute__((address_space(3))) void;

// one 16-byte load per lane into LDS at wave_base + 16 * lane; an offset
// past the field's end reads as zero
__device__ __forceinline__ void dma16(const FieldBuf &b, void *wave_base, unsigned byte_off) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(b.r, (lds_void *)wave_base, 16, byte_off, 0, 0, 0);
}

// Barrier for hand-offs through LDS while direct-to-LDS loads are in
// flight: it waits for this wave's LDS reads and writes, then for every wave
// of the block, but not for the wave's vector-memory loads (__syncthreads()
// would wait for those too, since an LDS load writes LDS).  The compiler
// still waits for an LDS load before the first read of its destination.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Output point of register slot q of the last Stockham stage (radix
// N / NS_last) for a thread whose first-stage column row is j0.
template <int N, int R>
__device__ __forceinline__ int last_stage_point(int j0, int q) {
  constexpr int NS = [] {
    int ns = 1;
    while (ns * R < N) ns *= R;
    return ns;
  }();
  constexpr int RS = N / NS;
  const int s = q / RS, r = q % RS;
  const int j = j0 + s * (N / R);
  return (j / NS) * NS * RS + j % NS + r * NS;
}

// One Stockham stage over a tile held in LDS (lds[point * C + column]); the
// first stage (NS == 1) reads the points a thread's own lanes loaded, so it
// needs no barrier before it.  After the stage's reads, every wave passes an
// LDS barrier and the first stage then calls hook() (which may start loads
// into a buffer the previous transform has finished reading).  The last
// stage hands its outputs to st(point, column, q, value).
template <typename T, int N, int R, int RS, int NS, int C, int SIGN, class HOOK, class ST>
__device__ __forceinline__ void stockham_stage_lds(cplx<T> *lds, const double2 *__restrict__ tw, int tid,
                                                   HOOK &hook, ST &st) {
  using C2 = cplx<T>;
  constexpr int NB = N / RS;
  constexpr int PER = R / RS;
  constexpr bool last = NS * RS == N;
  // addresses recomputed in each stage rather than kept live in registers
  // across the transforms of a fused pass
  tid = opaque(tid);
  const int c = tid % C;
  const int j0 = tid / C;
  C2 a[R];
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    const int j = j0 + s * (N / R);
#pragma unroll
    for (int r = 0; r < RS; ++r) a[s * RS + r] = lds[(j + r * NB) * C + c];
    if constexpr (NS > 1) {
      const int k = j % NS;
      const double2 t1 = tw[k * (N / (NS * RS))];
      C2 w[RS];
      w[1] = C2{(T)t1.x, (T)(SIGN * t1.y)};
#pragma unroll
      for (int r = 2; r < RS; ++r) w[r] = (r % 2 == 0) ? cmul(w[r / 2], w[r / 2]) : cmul(w[r - 1], w[1]);
#pragma unroll
      for (int r = 1; r < RS; ++r) a[s * RS + r] = cmul(a[s * RS + r], w[r]);
    }
    dft<C2, RS, SIGN>(&a[s * RS]);
  }
  if constexpr (NS == 1 || !last) {
    lds_barrier();
    if constexpr (NS == 1) hook();
  }
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    const int j = j0 + s * (N / R);
    const int base = (j / NS) * NS * RS + j % NS;
#pragma unroll
    for (int r = 0; r < RS; ++r) {
      if constexpr (last)
        st(base + r * NS, c, s * RS + r, a[s * RS + r]);
      else
        lds[(base + r * NS) * C + c] = a[s * RS + r];
    }
  }
  if constexpr (!last) lds_barrier();
}

template <typename T, int N, int R, int C, int SIGN, int NS, class HOOK, class ST>
__device__ __forceinline__ void fft_stages_lds(cplx<T> *lds, const double2 *__restrict__ tw, int tid, HOOK &hook,
                                               ST &st) {
  if constexpr (NS < N) {
    constexpr int RS = (N / NS >= R) ? R : N / NS;
    stockham_stage_lds<T, N, R, RS, NS, C, SIGN>(lds, tw, tid, hook, st);
    fft_stages_lds<T, N, R, C, SIGN, NS * RS>(lds, tw, tid, hook, st);
  }
}

// Starts the direct-to-LDS load of a tile's column points: thread tid's
// slot q is point tid / C + q N / R of column tid % C, the point its own
// first Stockham stage reads; off[q] is that point's byte offset in the
// field (0xfffffff0 where the column is not read: zeros).
template <int N, int R, int C>
__device__ __forceinline__ void dma_tile(const FieldBuf &b, cplx<double> *buf, int tid, const unsigned *off) {
  constexpr int NT = C * N / R;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
#pragma unroll
  for (int q = 0; q < R; ++q) dma16(b, &buf[q * NT + wave * 64], off[q]);
}

// Direct-to-LDS loads of a tile of FP32 complex values, 4 bytes per lane
// (gfx950 has no 8-byte LDS load): dword k NT + tid of the tile, NT = C N / R
// threads and k < 2R, is float f % (2C) of row f / (2C), f = k NT + tid, at
// LDS float offset f.  The threads of a point are then not those that load
// it, so a block waits for its loads (s_waitcnt vmcnt(0)) and passes a
// barrier before the first stage reads the tile.
template <int N, int R, int C>
__device__ __forceinline__ void dma_tile4(const FieldBuf &b, cplx<float> *buf, int tid, const unsigned *off) {
  constexpr int NT = C * N / R;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  float *fb = reinterpret_cast<float *>(buf);
#pragma unroll
  for (int k = 0; k < 2 * R; ++k)
    __builtin_amdgcn_raw_ptr_buffer_load_lds(b.r, (lds_void *)&fb[k * NT + wave * 64], 4, off[k], 0, 0, 0);
}


}  // namespace spm
