// Oversampled-grid FFT with the stages around it fused, for grid lengths
// whose prime factors are 2, 3, 5, 7, 11, 13, 17 and 19 (other lengths use
// rocFFT).
//
// The reference runs, per coil, INVERSE fftshift -> cuFFT -> INVERSE fftshift
// -> crop -> 1/sqrt(Ng) -> deapodization -> conj(S) coil sum for the adjoint
// and S -> deapodization -> zero pad -> INVERSE fftshift -> cuFFT ->
// FORWARD fftshift for the forward (CUDA/src/gpuNUFFT_operator.cpp:682-850,
// :1143-1261; CUDA/src/gpu/std_gpuNUFFT_kernels.cu:114-218, :341-372,
// :493-521, :569-654, :699-793), each a separate pass over the whole grid.
//
// Here a d-dimensional transform is d passes, one per axis. A pass loads a
// tile of lines into LDS, transforms them with a Stockham FFT (radices 8, 4,
// 2, 3, 5 in closed form, the primes 7 .. 19 as direct DFTs) and writes
// them back. The passes are pruned to the data that matters:
//   adjoint: the z and y passes write only the cropped image range of their
//     axis; the last (x) pass writes image pixels and applies the shift,
//     crop, 1/sqrt(Ng), deapodization and the conj(S) coil sum on the way out;
//   forward: the first (x) pass reads image lines (S * x * deapo, zero
//     padding and shift fused into the load) and the y and z passes read only
//     the image range of their axis into otherwise zero lines.
// The shifts are index maps: the adjoint's grid arrives INVERSE-shifted from
// the fold; the forward's grid leaves as the plain spectrum and interp
// applies the output shift in its own indices.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "gnufft_internal.hpp"

namespace gnufft {

namespace {

constexpr int kFftThreads = 512;
constexpr int kTileElems = 8192;  // complex elements per LDS tile (64 KB)
constexpr int kPerThread = kTileElems / kFftThreads;  // tile elements per thread
// y / z passes of lines longer than 511 points: 1024 threads and a 128 KB
// tile of 16 lines (one workgroup per CU), the same work per thread
constexpr int kBigThreads = 1024;
constexpr int kBigTileElems = kPerThread * kBigThreads;
constexpr int kMaxStripe = 32;    // lines per tile of a y / z pass (256 B rows)
constexpr int kMaxPix = 4;        // image pixels per thread in the x passes
constexpr int kMaxLen = kMaxPix * kFftThreads;  // longest transform (2048)
constexpr int kCoilLoads = 4;  // per-pixel coil loads in flight in the x passes
constexpr int kTileLoads = 4;  // grid loads per thread in flight in the adjoint x pass

// ---------------------------------------------------------------------------
// Small DFTs with kernel exp(sign * 2 pi i / R)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// sign * i * a
__device__ __forceinline__ float2 mul_si(float2 a, float sign) { return make_float2(-sign * a.y, sign * a.x); }

template <int R>
__device__ __forceinline__ void dft(float2* v, float sign);

template <>
__device__ __forceinline__ void dft<2>(float2* v, float) {
  const float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void dft<4>(float2* v, float sign) {
  const float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  const float2 t2 = cadd(v[1], v[3]), t3 = mul_si(csub(v[1], v[3]), sign);
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <>
__device__ __forceinline__ void dft<8>(float2* v, float sign) {
  // radix-4 DFTs of the even and odd inputs, combined with w8^k
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  dft<4>(e, sign);
  dft<4>(o, sign);
  const float h = 0.70710678118654752f;
  // w8 = (1 + sign i)/sqrt2, w8^2 = sign i, w8^3 = (-1 + sign i)/sqrt2
  const float2 o1 = make_float2(h * (o[1].x - sign * o[1].y), h * (o[1].y + sign * o[1].x));
  const float2 o2 = mul_si(o[2], sign);
  const float2 o3 = make_float2(h * (-o[3].x - sign * o[3].y), h * (-o[3].y + sign * o[3].x));
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

template <>
__device__ __forceinline__ void dft<3>(float2* v, float sign) {
  const float s = 0.86602540378443865f * sign;  // sin(2 pi / 3)
  const float2 a = v[0];
  const float2 t = cadd(v[1], v[2]);
  const float2 u = csub(v[1], v[2]);
  const float2 m = make_float2(a.x - 0.5f * t.x, a.y - 0.5f * t.y);
  v[0] = cadd(a, t);
  v[1] = make_float2(m.x - s * u.y, m.y + s * u.x);
  v[2] = make_float2(m.x + s * u.y, m.y - s * u.x);
}

template <>
__device__ __forceinline__ void dft<5>(float2* v, float sign) {
  const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;  // cos 72, cos 144
  const float s1 = 0.95105651629515357f * sign, s2 = 0.58778525229247313f * sign;
  const float2 a = v[0];
  const float2 t1 = cadd(v[1], v[4]), u1 = csub(v[1], v[4]);
  const float2 t2 = cadd(v[2], v[3]), u2 = csub(v[2], v[3]);
  const float2 m1 = make_float2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
  const float2 m2 = make_float2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
  const float2 n1 = make_float2(-(s1 * u1.y + s2 * u2.y), s1 * u1.x + s2 * u2.x);
  const float2 n2 = make_float2(-(s2 * u1.y - s1 * u2.y), s2 * u1.x - s1 * u2.x);
  v[0] = cadd(a, cadd(t1, t2));
  v[1] = cadd(m1, n1);
  v[4] = csub(m1, n1);
  v[2] = cadd(m2, n2);
  v[3] = csub(m2, n2);
}

// Odd prime R (7, 11, 13, 17, 19): the direct DFT, inputs paired as
// x_n +- x_{R-n} so that X_k and X_{R-k} share their sums:
// X_k = x_0 + sum_n (x_n + x_{R-n}) cos(2 pi n k / R)
//           + i sum_n (x_n - x_{R-n}) sign sin(2 pi n k / R), n = 1 .. (R-1)/2.
// The constants come from the stage's twiddle table: w[m * step] =
// exp(sign 2 pi i m / R) for a line of step * R points. Each output goes to
// out(k, X_k) as soon as it is formed (no output array: fewer registers).
template <int R, typename Out>
__device__ __forceinline__ void dft_prime_out(const float2* v, const float2* w, int step, Out out) {
  constexpr int H = (R - 1) / 2;
  float c[H], sn[H];
#pragma unroll
  for (int m = 1; m <= H; m++) {
    const float2 e = w[m * step];
    c[m - 1] = e.x;
    sn[m - 1] = e.y;
  }
  float2 t[H], u[H];
  const float2 x0 = v[0];
  float2 sum = x0;
#pragma unroll
  for (int n = 1; n <= H; n++) {
    t[n - 1] = cadd(v[n], v[R - n]);
    u[n - 1] = csub(v[n], v[R - n]);
    sum = cadd(sum, t[n - 1]);
  }
  out(0, sum);
#pragma unroll
  for (int k = 1; k <= H; k++) {
    float2 a = x0, b = make_float2(0.f, 0.f);
#pragma unroll
    for (int n = 1; n <= H; n++) {
      const int m = (n * k) % R;  // compile-time
      const float cm = m <= H ? c[m - 1] : c[R - m - 1];
      const float sm = m <= H ? sn[m - 1] : -sn[R - m - 1];
      a = make_float2(a.x + t[n - 1].x * cm, a.y + t[n - 1].y * cm);
      b = make_float2(b.x + u[n - 1].x * sm, b.y + u[n - 1].y * sm);
    }
    out(k, make_float2(a.x - b.y, a.y + b.x));
    out(R - k, make_float2(a.x + b.y, a.y - b.x));
  }
}

// n / d for n * d < 2^32 via a multiply-high with m = ceil(2^32 / d); d == 1
// has no 32-bit m and is passed as m == 0.
__device__ __forceinline__ int qdiv(int n, uint32_t m) {
  return m ? (int)__umulhi((uint32_t)n, m) : n;
}

// p mod L for p in [-L, 2L)
__device__ __forceinline__ int wrap1(int p, int L) {
  p = p < 0 ? p + L : p;
  return p >= L ? p - L : p;
}

__device__ __forceinline__ uint32_t qdiv_const_dev(int d) {
  return d == 1 ? 0u : (uint32_t)((0x100000000ull + (uint32_t)d - 1) / (uint32_t)d);
}

// One Stockham stage of a length-L transform (Govindaraju et al., "High
// performance discrete Fourier transforms on graphics processors", SC08):
// butterfly j of a line reads elements j + s L/R, multiplies them by
// w^(k s L/(Ns R)) with k = j mod Ns, runs a radix-R DFT and writes
// elements (j / Ns) Ns R + k + s Ns.
//
// Butterflies of a tile are numbered b = j * lines + l: the lanes of a wave
// walk the lines at a fixed j, so the LDS addresses l * LP + f(j) of one
// access (odd line pitch LP) fall into distinct banks in every stage.
struct FftStage {
  int R, Ns, per_line, tws;
  uint32_t m_ns;  // qdiv constant for Ns
  uint32_t pad;   // 24 bytes: records pack into float2 slots of the table
};

struct FftRadices {
  int n;
  FftStage st[12];
};

template <int R>
struct Bfly {
  static constexpr int KB = (kPerThread + R - 1) / R;  // butterflies per thread
};

// Twiddle, radix-R DFT and store of butterfly j (values v) at
// (j / Ns) Ns R + k + s Ns of line l, through `put(l, index, value)`.
template <int R, typename Put>
__device__ __forceinline__ void finish_butterfly(float2* v, int j, int l, const FftStage& st,
                                        const float2* tw, float sign, Put put) {
  const int jq = qdiv(j, st.m_ns);
  const int k = j - jq * st.Ns;
  if (k) {
    // k s tws < Ns R tws = L: the index needs no wrap
#pragma unroll
    for (int s = 1; s < R; s++) v[s] = cmul(v[s], tw[k * s * st.tws]);
  }
  const int base = jq * st.Ns * R + k;
  if constexpr (R == 2 || R == 3 || R == 4 || R == 5 || R == 8) {
    dft<R>(v, sign);
#pragma unroll
    for (int s = 0; s < R; s++) put(l, base + s * st.Ns, v[s]);
  } else {
    dft_prime_out<R>(v, tw, st.per_line,
                     [&](int s, float2 x) { put(l, base + s * st.Ns, x); });
  }
}

// Stage with its inputs fetched by get(l, index) and outputs stored by
// put(l, index, value); a barrier separates all fetches from all stores
// when `sync` (in-place LDS stages).
template <int R, int NT = kFftThreads, typename Get, typename Put>
__device__ __forceinline__ void stage(int lines, uint32_t m_lines, const FftStage& st, const float2* tw,
                             float sign, bool sync, Get get, Put put) {
  constexpr int KB = Bfly<R>::KB;
  const int B = lines * st.per_line;
  float2 v[KB][R];
#pragma unroll
  for (int q = 0; q < KB; q++) {
    const int b = threadIdx.x + q * NT;
    if (b < B) {
      const int j = qdiv(b, m_lines);
      const int l = b - j * lines;
#pragma unroll
      for (int s = 0; s < R; s++) v[q][s] = get(l, j + s * st.per_line);
    }
  }
  if (sync) __syncthreads();
#pragma unroll
  for (int q = 0; q < KB; q++) {
    const int b = threadIdx.x + q * NT;
    if (b < B) {
      const int j = qdiv(b, m_lines);
      const int l = b - j * lines;
      finish_butterfly<R>(v[q], j, l, st, tw, sign, put);
    }
  }
}

// PR: also the prime radices 7 .. 19 (their registers would otherwise set
// the occupancy of every table-driven pass).
template <int NT = kFftThreads, bool PR = false, typename Get, typename Put>
__device__ __forceinline__ void stage_any(int lines, uint32_t m_lines, const FftStage& table_entry,
                                 const float2* tw, float sign, bool sync, Get get, Put put) {
  // a register copy: the table lives in global memory, and stores of the
  // stage to global memory would otherwise force reloads of its fields
  const FftStage st = table_entry;
  switch (st.R) {
    case 8: stage<8, NT>(lines, m_lines, st, tw, sign, sync, get, put); break;
    case 4: stage<4, NT>(lines, m_lines, st, tw, sign, sync, get, put); break;
    case 2: stage<2, NT>(lines, m_lines, st, tw, sign, sync, get, put); break;
    case 3: stage<3, NT>(lines, m_lines, st, tw, sign, sync, get, put); break;
    case 5: stage<5, NT>(lines, m_lines, st, tw, sign, sync, get, put); break;
    default:
      if constexpr (PR) {
        switch (st.R) {
          case 7: stage<7, NT>(lines, m_lines, st, tw, sign, sync, get, put); break;
          case 11: stage<11, NT>(lines, m_lines, st, tw, sign, sync, get, put); break;
          case 13: stage<13, NT>(lines, m_lines, st, tw, sign, sync, get, put); break;
          case 17: stage<17, NT>(lines, m_lines, st, tw, sign, sync, get, put); break;
          default: stage<19, NT>(lines, m_lines, st, tw, sign, sync, get, put); break;
        }
      }
      break;
  }
}

// Transform `lines` lines held in LDS (stride LP) in place. Callers
// synchronise after filling the lines; every stage ends with a barrier.
template <bool PR>
__device__ __forceinline__ void fft_lines(float2* buf, int lines, uint32_t m_lines, int LP,
                                 const FftStage* st, int n_st, const float2* tw, float sign) {
  auto get = [&](int l, int i) { return buf[l * LP + i]; };
  auto put = [&](int l, int i, float2 v) { buf[l * LP + i] = v; };
  for (int i = 0; i < n_st; i++) {
    stage_any<kFftThreads, PR>(lines, m_lines, st[i], tw, sign, true, get, put);
    __syncthreads();
  }
}

// Compile-time stage tables for the common line lengths: every stage
// constant (radix, Ns, strides, division constants) folds into the code.
constexpr uint32_t qmagic(int d) {
  return d == 1 ? 0u : (uint32_t)((0x100000000ull + (unsigned)d - 1) / (unsigned)d);
}

template <int L, int... R>
struct FixedStages {
  static constexpr int n = sizeof...(R);
  static constexpr int rad[n] = {R...};
  static_assert((R * ...) == L, "radices multiply to the line length");
  static_assert(n >= 2, "first and last stages are distinct");
  // points of the sub-transforms that stage i combines
  static constexpr int ns(int i) { return i == 0 ? 1 : ns(i - 1) * rad[i - 1]; }
  template <int I>
  static constexpr FftStage s{rad[I], ns(I), L / rad[I], L / (ns(I) * rad[I]), qmagic(ns(I)), 0u};
};

// Stages I... of F over `lines` lines, each followed by a barrier.
template <typename F, int... I, typename Get, typename Put>
__device__ __forceinline__ void fixed_stages(int lines, uint32_t m_lines, const float2* tw, float sign,
                                             bool sync, Get get, Put put,
                                             std::integer_sequence<int, I...>) {
  ((stage<F::rad[I]>(lines, m_lines, F::template s<I>, tw, sign, sync, get, put), __syncthreads()),
   ...);
}

// Stages 1 .. n-2 of F (the in-LDS stages between a global first and last).
template <typename F, int... I, typename Get, typename Put>
__device__ __forceinline__ void fixed_middle(int lines, uint32_t m_lines, const float2* tw, float sign,
                                             Get get, Put put, std::integer_sequence<int, I...>) {
  ((stage<F::rad[I + 1]>(lines, m_lines, F::template s<I + 1>, tw, sign, true, get, put),
    __syncthreads()),
   ...);
}

// Stage runners: all stages of `lines` LDS lines (the x passes).
template <bool PR>
struct RunTable {
  const FftStage* st;
  int n;
  __device__ __forceinline__ void operator()(float2* buf, int lines, uint32_t m_lines, int LP,
                                             const float2* tw, float sign) const {
    fft_lines<PR>(buf, lines, m_lines, LP, st, n, tw, sign);
  }
};

template <int L, int... R>
struct RunFixed {
  __device__ __forceinline__ void operator()(float2* buf, int lines, uint32_t m_lines, int LP,
                                             const float2* tw, float sign) const {
    using F = FixedStages<L, R...>;
    auto get = [&](int l, int i) { return buf[l * LP + i]; };
    auto put = [&](int l, int i, float2 v) { buf[l * LP + i] = v; };
    fixed_stages<F>(lines, m_lines, tw, sign, true, get, put,
                    std::make_integer_sequence<int, F::n>{});
  }
};

// ---------------------------------------------------------------------------
// y / z pass over [coil][d2][d1][d0] arrays. Lines run along the pass axis;
// a tile is `stripe` lines adjacent in x (a power of two, up to 32, that
// fits the LDS tile) at one position of the remaining axis, so every global
// access of a wave covers rows of `stripe` consecutive elements.
// The first stage reads its inputs from global memory (input extent in_len;
// element i of the input sits at FFT index wrap(i + in_off, L) of an
// otherwise zero line when in_scatter) and the last stage writes its
// outputs there (output extent out_len; output o is FFT index
// wrap(o + out_off, L) when out_gather); the stages between run in LDS.
// ---------------------------------------------------------------------------
struct AxisPass {
  const float2* in;
  float2* out;
  long long in_coil, out_coil;  // elements per coil
  long long in_step, out_step;  // stride along the pass axis
  long long in_ob, out_ob;      // stride along the remaining axis
  int nx;                       // extent in x
  int nob;                      // extent of the remaining axis
  int L, LP;
  int in_len, out_len;
  int in_scatter, in_off;
  int out_gather, out_off;
  int stripe, x_tiles, coils;
  int big;  // 1: lines above 511 points, kBigThreads and a kBigTileElems tile
  uint32_t m_stripe;
  float sign;
  const FftStage* st;  // stage table in device memory (beside the twiddles)
  int n_st;
  const float2* tw;
};

template <int NT, bool PR>
__device__ __forceinline__ void fft_axis_body(const AxisPass& p, int block) {
  extern __shared__ float2 lds[];
  float2* tw = lds;
  float2* buf = lds + p.L;
  // block = (coil * nob + ob) * x_tiles + xt
  const int xt = block % p.x_tiles;
  const int r = block / p.x_tiles;
  const int ob = r % p.nob;
  const int coil = r / p.nob;
  const int x0 = xt * p.stripe;
  const int nl = min(p.stripe, p.nx - x0);
  for (int i = threadIdx.x; i < p.L; i += NT) tw[i] = p.tw[i];
  const float2* in = p.in + coil * p.in_coil + ob * p.in_ob + x0;
  float2* out = p.out + coil * p.out_coil + ob * p.out_ob + x0;
  const int L = p.L, LP = p.LP;
  // offsets within a coil's array fit 32 bits (grids below 2^31 points);
  // in_off, out_off lie in (-L, L), so one conditional step wraps an index
  const uint32_t in_step = (uint32_t)p.in_step, out_step = (uint32_t)p.out_step;
  // lines l >= nl of a partial tile are neither read nor written
  auto get_global = [&](int l, int i) {
    int idx = i;
    if (p.in_scatter) {
      idx = wrap1(i - p.in_off, L);
      if (idx >= p.in_len) return make_float2(0.f, 0.f);
    }
    return l < nl ? in[(uint32_t)idx * in_step + (uint32_t)l] : make_float2(0.f, 0.f);
  };
  auto put_global = [&](int l, int o, float2 v) {
    if (p.out_gather) o = wrap1(o - p.out_off, L);
    if (l < nl && o < p.out_len) out[(uint32_t)o * out_step + (uint32_t)l] = v;
  };
  auto get_lds = [&](int l, int i) { return buf[l * LP + i]; };
  auto put_lds = [&](int l, int i, float2 v) { buf[l * LP + i] = v; };
  const int S = p.stripe;
  const int n = p.n_st;
  __syncthreads();  // twiddles
  if (n == 1) {
    stage_any<NT, PR>(S, p.m_stripe, p.st[0], tw, p.sign, false, get_global, put_global);
    return;
  }
  stage_any<NT, PR>(S, p.m_stripe, p.st[0], tw, p.sign, false, get_global, put_lds);
  __syncthreads();
  for (int i = 1; i < n - 1; i++) {
    stage_any<NT, PR>(S, p.m_stripe, p.st[i], tw, p.sign, true, get_lds, put_lds);
    __syncthreads();
  }
  stage_any<NT, PR>(S, p.m_stripe, p.st[n - 1], tw, p.sign, false, get_lds, put_global);
}

template <bool PR>
__global__ void __launch_bounds__(kFftThreads, 4) fft_axis_kernel(AxisPass p) {
  fft_axis_body<kFftThreads, PR>(p, blockIdx.x);
}

template <bool PR>
__global__ void __launch_bounds__(kBigThreads, 1) fft_axis_big_kernel(AxisPass p) {
  fft_axis_body<kBigThreads, PR>(p, blockIdx.x);
}

// The same pass for a line length L = R0 R1 ... (two or more stages) and a
// tile of S lines known at compile time (S a power of two).
template <int L, int S, int... R>
__device__ __forceinline__ void fft_axis_fixed_body(const AxisPass& p, int block) {
  using F = FixedStages<L, R...>;
  constexpr int LP = (L + 1) | 1;
  constexpr uint32_t mS = qmagic(S);
  extern __shared__ float2 lds[];
  float2* tw = lds;
  float2* buf = lds + L;
  const int xt = block % p.x_tiles;
  const int r = block / p.x_tiles;
  const int ob = r % p.nob;
  const int coil = r / p.nob;
  const int x0 = xt * S;
  const int nl = min(S, p.nx - x0);
  for (int i = threadIdx.x; i < L; i += kFftThreads) tw[i] = p.tw[i];
  const float2* in = p.in + coil * p.in_coil + ob * p.in_ob + x0;
  float2* out = p.out + coil * p.out_coil + ob * p.out_ob + x0;
  const uint32_t in_step = (uint32_t)p.in_step, out_step = (uint32_t)p.out_step;
  auto get_global = [&](int l, int i) {
    int idx = i;
    if (p.in_scatter) {
      idx = wrap1(i - p.in_off, L);
      if (idx >= p.in_len) return make_float2(0.f, 0.f);
    }
    return l < nl ? in[(uint32_t)idx * in_step + (uint32_t)l] : make_float2(0.f, 0.f);
  };
  auto put_global = [&](int l, int o, float2 v) {
    if (p.out_gather) o = wrap1(o - p.out_off, L);
    if (l < nl && o < p.out_len) out[(uint32_t)o * out_step + (uint32_t)l] = v;
  };
  auto get_lds = [&](int l, int i) { return buf[l * LP + i]; };
  auto put_lds = [&](int l, int i, float2 v) { buf[l * LP + i] = v; };
  __syncthreads();  // twiddles
  constexpr int n = F::n;
  stage<F::rad[0]>(S, mS, F::template s<0>, tw, p.sign, false, get_global, put_lds);
  __syncthreads();
  fixed_middle<F>(S, mS, tw, p.sign, get_lds, put_lds, std::make_integer_sequence<int, n - 2>{});
  stage<F::rad[n - 1]>(S, mS, F::template s<n - 1>, tw, p.sign, false, get_lds, put_global);
}

template <int L, int S, int... R>
__global__ void __launch_bounds__(kFftThreads, 4) fft_axis_fixed_kernel(AxisPass p) {
  fft_axis_fixed_body<L, S, R...>(p, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Adjoint last pass (x): lines [coil][N2][N1][G0]; pixel x of an image line
// reads FFT index wrap(x + crop + shift_inv, G0), scaled by 1/sqrt(Ng)
// [* deapo], then either summed over coils with conj(S) or written per coil
// (CUDA/src/gpu/std_gpuNUFFT_kernels.cu:114-218, :341-372, :493-521).
// A workgroup takes Q image lines and walks the coils in groups of Cg.
// ---------------------------------------------------------------------------
struct LastPass {
  const float2* in;
  float2* out;  // [img] with sens, else [coil][img]
  const float* deapo;
  const float2* sens;  // null: per-coil output
  int G0, N0, LP;
  int lines;  // N1 * N2
  int crop0, shift0;
  int coils, Cg, Q;
  uint32_t m_g0, m_n0;  // qdiv constants for G0, N0
  float scale;
  int apply_deapo, accumulate;
  float sign;
  const FftStage* st;  // stage table in device memory (beside the twiddles)
  int n_st;
  const float2* tw;
};

template <typename Run>
__device__ __forceinline__ void fft_adjoint_last_body(const LastPass& p, int block, const Run& run) {
  extern __shared__ float2 lds[];
  float2* tw = lds;
  float2* buf = lds + p.G0;
  const int l0 = block * p.Q;
  const int nq = min(p.Q, p.lines - l0);
  for (int i = threadIdx.x; i < p.G0; i += kFftThreads) tw[i] = p.tw[i];
  const long long Ni = (long long)p.N0 * p.lines;
  const long long cin = (long long)p.G0 * p.lines;
  const long long pix0 = (long long)l0 * p.N0;
  const int npix = nq * p.N0;
  const int per_coil = nq * p.G0;  // the Q lines of one coil are contiguous
  for (int c0 = 0; c0 < p.coils; c0 += p.Cg) {
    const int ng = min(p.Cg, p.coils - c0);
    __syncthreads();  // the previous group's reads of buf are done
    // the group's lines, all coils in one flat loop: kTileLoads loads per
    // thread in flight before their LDS writes (a loop per coil leaves one
    // or two loads per thread, and a memory round trip per coil)
    {
      const int total = ng * per_coil;
      const uint32_t m_pc = qdiv_const_dev(per_coil);
      const float2* src = p.in + (long long)c0 * cin + (long long)l0 * p.G0;
      for (int e0 = threadIdx.x; e0 < total; e0 += kTileLoads * kFftThreads) {
        float2 v[kTileLoads];
#pragma unroll
        for (int u = 0; u < kTileLoads; u++) {
          const int e = e0 + u * kFftThreads;
          if (e < total) {
            const int c = qdiv(e, m_pc);
            v[u] = src[(long long)c * cin + (e - c * per_coil)];
          }
        }
#pragma unroll
        for (int u = 0; u < kTileLoads; u++) {
          const int e = e0 + u * kFftThreads;
          if (e < total) {
            const int c = qdiv(e, m_pc);
            const int r = e - c * per_coil;
            const int q = qdiv(r, p.m_g0);
            buf[(c * nq + q) * p.LP + (r - q * p.G0)] = v[u];
          }
        }
      }
    }
    __syncthreads();
    run(buf, ng * nq, qdiv_const_dev(ng * nq), p.LP, tw, p.sign);
    // the coil sum of later coil groups (and of later calls when
    // accumulating) adds to what this thread wrote for the same pixels
    const bool add = c0 > 0 || p.accumulate;
    for (int e = threadIdx.x; e < npix; e += kFftThreads) {
      const int q = qdiv(e, p.m_n0);
      const int x = e - q * p.N0;
      const long long pix = pix0 + e;
      const int m = wrap1(x + p.crop0 + p.shift0, p.G0);
      const float d = p.apply_deapo ? p.deapo[pix] : 1.0f;
      if (p.sens) {
        float2 acc = add ? p.out[pix] : make_float2(0.f, 0.f);
        // coils in groups of kCoilLoads: their sensitivity loads are in
        // flight together (the sum keeps the coil order)
        for (int cb = 0; cb < ng; cb += kCoilLoads) {
          float2 sc[kCoilLoads];
#pragma unroll
          for (int u = 0; u < kCoilLoads; u++)
            sc[u] = cb + u < ng ? p.sens[(long long)(c0 + cb + u) * Ni + pix] : make_float2(0.f, 0.f);
#pragma unroll
          for (int u = 0; u < kCoilLoads; u++) {
            if (cb + u >= ng) break;
            float2 v = buf[((cb + u) * nq + q) * p.LP + m];
            v.x = (v.x * p.scale) * d;
            v.y = (v.y * p.scale) * d;
            const float2 s = sc[u];
            acc.x += v.x * s.x + v.y * s.y;
            acc.y += v.y * s.x - v.x * s.y;
          }
        }
        p.out[pix] = acc;
      } else {
        for (int c = 0; c < ng; c++) {
          float2 v = buf[(c * nq + q) * p.LP + m];
          v.x = (v.x * p.scale) * d;
          v.y = (v.y * p.scale) * d;
          p.out[(long long)(c0 + c) * Ni + pix] = v;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Forward first pass (x): image lines -> [coil][N2][N1][G0]. Pixel x goes to
// position wrap(x + crop - shift_inv, G0) of a zero line (zero padding and
// INVERSE fftshift) with value (S_c * x) * deapo
// (CUDA/src/gpu/std_gpuNUFFT_kernels.cu:142-218, :493-521, :699-793).
// ---------------------------------------------------------------------------
struct FirstPass {
  const float2* img;  // [img] (sens) or [coil][img]
  float2* out;
  const float* deapo;
  const float2* sens;
  int img_per_coil;
  int G0, N0, LP;
  int lines;
  int crop0, shift0;
  int coils, Cg, Q;
  uint32_t m_g0, m_n0, m_pad;  // qdiv constants for G0, N0, G0 - N0
  float sign;
  const FftStage* st;  // stage table in device memory (beside the twiddles)
  int n_st;
  const float2* tw;
};

template <typename Run>
__device__ __forceinline__ void fft_forward_first_body(const FirstPass& p, int block, const Run& run) {
  extern __shared__ float2 lds[];
  float2* tw = lds;
  float2* buf = lds + p.G0;
  const int l0 = block * p.Q;
  const int nq = min(p.Q, p.lines - l0);
  for (int i = threadIdx.x; i < p.G0; i += kFftThreads) tw[i] = p.tw[i];
  const long long Ni = (long long)p.N0 * p.lines;
  const long long cout = (long long)p.G0 * p.lines;
  const long long pix0 = (long long)l0 * p.N0;
  const int npix = nq * p.N0;
  const int per_coil = nq * p.G0;
  for (int c0 = 0; c0 < p.coils; c0 += p.Cg) {
    const int ng = min(p.Cg, p.coils - c0);
    __syncthreads();  // the previous group's reads of buf are done
    // zero padding: positions of "pixels" N0 .. G0-1 of every line
    const int pad = p.G0 - p.N0;
    for (int e = threadIdx.x; e < ng * nq * pad; e += kFftThreads) {
      const int l = qdiv(e, p.m_pad);
      const int x = p.N0 + e - l * pad;
      buf[l * p.LP + wrap1(x + p.crop0 - p.shift0, p.G0)] = make_float2(0.f, 0.f);
    }
    for (int e = threadIdx.x; e < npix; e += kFftThreads) {
      const int q = qdiv(e, p.m_n0);
      const int x = e - q * p.N0;
      const long long pix = pix0 + e;
      const float d = p.deapo[pix];
      const int pos = wrap1(x + p.crop0 - p.shift0, p.G0);
      const float2 shared_img = p.img_per_coil ? make_float2(0.f, 0.f) : p.img[pix];
      for (int cb = 0; cb < ng; cb += kCoilLoads) {
        // kCoilLoads coils' loads in flight before their LDS writes
        float2 v[kCoilLoads];
#pragma unroll
        for (int u = 0; u < kCoilLoads; u++) {
          const long long o = (long long)(c0 + cb + u) * Ni + pix;
          v[u] = cb + u < ng && p.img_per_coil ? p.img[o] : shared_img;
          if (cb + u < ng && p.sens) v[u] = cmul(v[u], p.sens[o]);
        }
#pragma unroll
        for (int u = 0; u < kCoilLoads; u++)
          if (cb + u < ng) buf[((cb + u) * nq + q) * p.LP + pos] = make_float2(v[u].x * d, v[u].y * d);
      }
    }
    __syncthreads();
    run(buf, ng * nq, qdiv_const_dev(ng * nq), p.LP, tw, p.sign);
    for (int c = 0; c < ng; c++) {
      float2* dst = p.out + (long long)(c0 + c) * cout + (long long)l0 * p.G0;
#pragma unroll 4
      for (int e = threadIdx.x; e < per_coil; e += kFftThreads) {
        const int q = qdiv(e, p.m_g0);
        const int x = e - q * p.G0;
        dst[e] = buf[(c * nq + q) * p.LP + x];
      }
    }
  }
}

template <bool PR>
__global__ void __launch_bounds__(kFftThreads, 4) fft_adjoint_last_kernel(LastPass p) {
  fft_adjoint_last_body(p, blockIdx.x, RunTable<PR>{p.st, p.n_st});
}

template <int L, int... R>
__global__ void __launch_bounds__(kFftThreads, 4) fft_adjoint_last_fixed_kernel(LastPass p) {
  fft_adjoint_last_body(p, blockIdx.x, RunFixed<L, R...>{});
}

template <bool PR>
__global__ void __launch_bounds__(kFftThreads, 4) fft_forward_first_kernel(FirstPass p) {
  fft_forward_first_body(p, blockIdx.x, RunTable<PR>{p.st, p.n_st});
}

template <int L, int... R>
__global__ void __launch_bounds__(kFftThreads, 4) fft_forward_first_fixed_kernel(FirstPass p) {
  fft_forward_first_body(p, blockIdx.x, RunFixed<L, R...>{});
}

// ---------------------------------------------------------------------------
// Batches: the passes of many operators, one launch per pass. An item
// (operator, block) names a workgroup's tile; FftOpDesc holds the
// operator's pass descriptors in device memory.
// ---------------------------------------------------------------------------
struct FftOpDesc {
  AxisPass adj_ax[2];  // adjoint z (3-D), y
  LastPass adj_x;
  FirstPass fwd_x;
  AxisPass fwd_ax[2];  // forward y, z (3-D)
  int n_ax;            // d - 1
};

template <bool PR>
__global__ void __launch_bounds__(kFftThreads, 4) fft_axis_batch_kernel(const FftOpDesc* __restrict__ d,
                                                                        const uint2* __restrict__ items,
                                                                        int adjoint, int which) {
  const uint2 it = items[blockIdx.x];
  const FftOpDesc& o = d[it.x];
  // register copies: the descriptors live in global memory, which the
  // passes also store to
  const AxisPass p = adjoint ? o.adj_ax[which] : o.fwd_ax[which];
  fft_axis_body<kFftThreads, PR>(p, (int)it.y);
}

template <bool PR>
__global__ void __launch_bounds__(kBigThreads, 1) fft_axis_batch_big_kernel(
    const FftOpDesc* __restrict__ d, const uint2* __restrict__ items, int adjoint, int which) {
  const uint2 it = items[blockIdx.x];
  const FftOpDesc& o = d[it.x];
  const AxisPass p = adjoint ? o.adj_ax[which] : o.fwd_ax[which];
  fft_axis_body<kBigThreads, PR>(p, (int)it.y);
}

template <bool PR>
__global__ void __launch_bounds__(kFftThreads, 4) fft_adjoint_last_batch_kernel(
    const FftOpDesc* __restrict__ d, const uint2* __restrict__ items) {
  const uint2 it = items[blockIdx.x];
  const LastPass p = d[it.x].adj_x;
  fft_adjoint_last_body(p, (int)it.y, RunTable<PR>{p.st, p.n_st});
}

template <bool PR>
__global__ void __launch_bounds__(kFftThreads, 4) fft_forward_first_batch_kernel(
    const FftOpDesc* __restrict__ d, const uint2* __restrict__ items) {
  const uint2 it = items[blockIdx.x];
  const FirstPass p = d[it.x].fwd_x;
  fft_forward_first_body(p, (int)it.y, RunTable<PR>{p.st, p.n_st});
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
uint32_t qdiv_const(int d) { return d == 1 ? 0u : udiv_magic((uint32_t)d); }

bool make_radices(int L, FftRadices* rad) {
  rad->n = 0;
  int n = L;
  const int order[10] = {8, 4, 2, 3, 5, 7, 11, 13, 17, 19};
  for (int r : order)
    while (n > 1 && n % r == 0) {
      if (rad->n == 12) return false;
      rad->st[rad->n++].R = r;
      n /= r;
    }
  if (n != 1 || rad->n == 0) return false;
  int Ns = 1;
  for (int i = 0; i < rad->n; i++) {
    FftStage& s = rad->st[i];
    s.Ns = Ns;
    s.per_line = L / s.R;
    s.tws = L / (Ns * s.R);
    s.m_ns = qdiv_const(Ns);
    Ns *= s.R;
  }
  return true;
}

// Lines whose transform has a stage of radix 7, 11, 13, 17 or 19 (run by
// the kernel variants compiled with those stages).
bool prime_radix(int L) { return L % 7 == 0 || L % 11 == 0 || L % 13 == 0 || L % 17 == 0 || L % 19 == 0; }

// odd line pitch: the lanes of a wave, one line each, hit distinct banks
int pad_of(int L) { return (L + 1) | 1; }

size_t tile_bytes(int L, int lines) {
  return ((size_t)L + (size_t)lines * pad_of(L)) * sizeof(float2);
}

template <typename K>
hipError_t allow_lds(K kernel, size_t lds) {
  if (lds <= 64 * 1024) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

// x-pass tile: Cg coils x Q lines with Cg * Q * LP <= kTileElems and
// Q * N0 <= kMaxPix * kFftThreads (N0 <= G0 <= kMaxLen keeps Q >= 1 inside
// both bounds); a multiple of 32 lines where that fits (one line per lane).
// Coils that need several groups are split into equal groups (16 coils at
// 15 lines per tile: 8 + 8, not 15 + 1).
void x_tiling(int G0, int N0, int coils, int* Cg, int* Q) {
  int per_tile = std::max(1, kTileElems / pad_of(G0));
  if (per_tile >= 32) per_tile -= per_tile % 32;
  const int groups = (std::max(1, coils) + per_tile - 1) / per_tile;
  *Cg = std::max(1, (coils + groups - 1) / groups);
  *Q = std::max(1, std::min(per_tile / *Cg, (kMaxPix * kFftThreads) / N0));
}

// The stage table of axis a inside the device table buffer.
const FftStage* stage_table(const float2* tw, const size_t tw_off[3][3], int a) {
  return reinterpret_cast<const FftStage*>(tw + tw_off[a][2]);
}

// Tiling of a y / z pass (the caller fills the geometry fields).
bool finish_axis(AxisPass* p, int coils, const FftStage* st) {
  p->coils = coils;
  p->LP = pad_of(p->L);
  FftRadices r;
  if (!make_radices(p->L, &r)) return false;
  p->st = st;
  p->n_st = r.n;
  // 16-line tiles at least (128-byte rows): lines above 511 points take the
  // large tile
  p->big = 16 * p->LP > kTileElems ? 1 : 0;
  const int tile = p->big ? kBigTileElems : kTileElems;
  int lg = 0;
  while ((2 << lg) <= kMaxStripe && (2 << lg) * p->LP <= tile) lg++;
  p->stripe = 1 << lg;
  p->m_stripe = qdiv_const(p->stripe);
  p->x_tiles = (p->nx + p->stripe - 1) / p->stripe;
  return true;
}

uint32_t axis_blocks(const AxisPass& p) { return (uint32_t)p.coils * p.nob * p.x_tiles; }
size_t axis_lds(const AxisPass& p) { return tile_bytes(p.L, p.stripe); }
uint32_t last_blocks(const LastPass& p) { return (uint32_t)((p.lines + p.Q - 1) / p.Q); }
size_t last_lds(const LastPass& p) { return tile_bytes(p.G0, p.Q * p.Cg); }
uint32_t first_blocks(const FirstPass& p) { return (uint32_t)((p.lines + p.Q - 1) / p.Q); }
size_t first_lds(const FirstPass& p) { return tile_bytes(p.G0, p.Q * p.Cg); }

// Line lengths with compile-time kernels: (L, tile lines of a y / z pass,
// radices). The tile must be the one finish_axis picks for L. Besides the
// common smooth lengths (and 320, the 160^3 osf-2 grid), the 3-D grids of
// 96^3 .. 160^3 images at osf 1.5 and 2: those with a prime factor 7 .. 19
// and the smooth ones that need four stages (144, 180, 216, 240, 288).
#define GNUFFT_FIXED_FFT_SIZES(X) \
  X(96, 32, 8, 4, 3)              \
  X(128, 32, 8, 8, 2)             \
  X(160, 32, 8, 4, 5)             \
  X(192, 32, 8, 8, 3)             \
  X(256, 16, 8, 8, 4)             \
  X(156, 32, 4, 3, 13)            \
  X(168, 32, 8, 3, 7)             \
  X(204, 32, 4, 3, 17)            \
  X(208, 32, 8, 2, 13)            \
  X(224, 32, 8, 4, 7)             \
  X(228, 32, 4, 3, 19)            \
  X(272, 16, 8, 2, 17)            \
  X(304, 16, 8, 2, 19)            \
  X(320, 16, 8, 8, 5)             \
  X(144, 32, 4, 4, 3, 3)          \
  X(180, 32, 4, 5, 3, 3)          \
  X(216, 32, 8, 3, 3, 3)          \
  X(240, 32, 4, 4, 5, 3)          \
  X(288, 16, 8, 4, 3, 3)

// Line lengths with compile-time kernels for the x passes only (their y / z
// passes use the large tiles).
#define GNUFFT_FIXED_X_SIZES(X) \
  X(512, 0, 8, 8, 8)

template <typename K>
hipError_t launch_pass(K kernel, uint32_t blocks, size_t lds, hipStream_t s,
                       const AxisPass& p, int threads = kFftThreads) {
  hipError_t e = allow_lds(kernel, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(threads), lds, s, p);
  return hipGetLastError();
}

hipError_t launch_axis(const AxisPass& p, hipStream_t s) {
  const size_t lds = axis_lds(p);
  const uint32_t blocks = axis_blocks(p);
#define GNUFFT_AXIS_CASE(LEN, TILE, ...) \
  if (p.L == LEN && p.stripe == TILE)    \
    return launch_pass(fft_axis_fixed_kernel<LEN, TILE, __VA_ARGS__>, blocks, lds, s, p);
  GNUFFT_FIXED_FFT_SIZES(GNUFFT_AXIS_CASE)
#undef GNUFFT_AXIS_CASE
  if (prime_radix(p.L)) {
    if (p.big) return launch_pass(fft_axis_big_kernel<true>, blocks, lds, s, p, kBigThreads);
    return launch_pass(fft_axis_kernel<true>, blocks, lds, s, p);
  }
  if (p.big) return launch_pass(fft_axis_big_kernel<false>, blocks, lds, s, p, kBigThreads);
  return launch_pass(fft_axis_kernel<false>, blocks, lds, s, p);
}

template <typename K, typename P>
hipError_t launch_x(K kernel, uint32_t blocks, size_t lds, hipStream_t s, const P& p) {
  hipError_t e = allow_lds(kernel, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kFftThreads), lds, s, p);
  return hipGetLastError();
}

hipError_t launch_last(const LastPass& p, hipStream_t s) {
  const size_t lds = last_lds(p);
  const uint32_t blocks = last_blocks(p);
#define GNUFFT_LAST_CASE(LEN, TILE, ...) \
  if (p.G0 == LEN) return launch_x(fft_adjoint_last_fixed_kernel<LEN, __VA_ARGS__>, blocks, lds, s, p);
  GNUFFT_FIXED_FFT_SIZES(GNUFFT_LAST_CASE)
  GNUFFT_FIXED_X_SIZES(GNUFFT_LAST_CASE)
#undef GNUFFT_LAST_CASE
  if (prime_radix(p.G0)) return launch_x(fft_adjoint_last_kernel<true>, blocks, lds, s, p);
  return launch_x(fft_adjoint_last_kernel<false>, blocks, lds, s, p);
}

hipError_t launch_first(const FirstPass& p, hipStream_t s) {
  const size_t lds = first_lds(p);
  const uint32_t blocks = first_blocks(p);
#define GNUFFT_FIRST_CASE(LEN, TILE, ...) \
  if (p.G0 == LEN) return launch_x(fft_forward_first_fixed_kernel<LEN, __VA_ARGS__>, blocks, lds, s, p);
  GNUFFT_FIXED_FFT_SIZES(GNUFFT_FIRST_CASE)
  GNUFFT_FIXED_X_SIZES(GNUFFT_FIRST_CASE)
#undef GNUFFT_FIRST_CASE
  if (prime_radix(p.G0)) return launch_x(fft_forward_first_kernel<true>, blocks, lds, s, p);
  return launch_x(fft_forward_first_kernel<false>, blocks, lds, s, p);
}

// Pass descriptors of the adjoint: grid [coils][G2][G1][G0] (the
// INVERSE-shifted spectrum) -> image; ax[0 .. d-2] then the x pass.
bool build_adjoint(const Geom& g, float2* grid, float2* tmp, int coils, const float* deapo,
                   const float2* sens, float2* out, int stop, bool accumulate, const float2* tw,
                   const size_t tw_off[3][3], AxisPass ax[2], LastPass* last) {
  const float sign = 1.0f;  // unnormalised inverse transform
  const int G0 = g.G[0], G1 = g.G[1];
  const int N1 = g.N[1], N2 = g.dim == 3 ? g.N[2] : 1;
  const float2* cur = grid;
  int n = 0;
  if (g.dim == 3) {
    // z: [c][G2][G1][G0] -> tmp [c][N2][G1][G0]
    AxisPass& z = ax[n++];
    z = AxisPass{};
    z.in = grid;
    z.out = tmp;
    z.in_coil = (long long)G0 * G1 * g.G[2];
    z.out_coil = (long long)G0 * G1 * N2;
    z.in_step = z.out_step = (long long)G0 * G1;
    z.in_ob = z.out_ob = G0;
    z.nx = G0;
    z.nob = G1;
    z.L = g.G[2];
    z.in_len = g.G[2];
    z.out_len = N2;
    z.out_gather = 1;
    z.out_off = g.crop[2] + g.shift_inv[2];
    z.sign = sign;
    z.tw = tw + tw_off[2][1];
    if (!finish_axis(&z, coils, stage_table(tw, tw_off, 2))) return false;
    cur = tmp;
  }
  // y: [c][N2][G1][G0] -> [c][N2][N1][G0] (3-D: tmp -> grid, 2-D: grid -> tmp)
  float2* ydst = g.dim == 3 ? grid : tmp;
  AxisPass& y = ax[n++];
  y = AxisPass{};
  y.in = cur;
  y.out = ydst;
  y.in_coil = (long long)G0 * G1 * N2;
  y.out_coil = (long long)G0 * N1 * N2;
  y.in_step = y.out_step = G0;
  y.in_ob = (long long)G0 * G1;
  y.out_ob = (long long)G0 * N1;
  y.nx = G0;
  y.nob = N2;
  y.L = G1;
  y.in_len = G1;
  y.out_len = N1;
  y.out_gather = 1;
  y.out_off = g.crop[1] + g.shift_inv[1];
  y.sign = sign;
  y.tw = tw + tw_off[1][1];
  if (!finish_axis(&y, coils, stage_table(tw, tw_off, 1))) return false;
  LastPass& p = *last;
  p = LastPass{};
  p.in = ydst;
  p.out = out;
  p.deapo = deapo;
  p.sens = stop == 1 ? nullptr : sens;
  p.G0 = G0;
  p.N0 = g.N[0];
  p.LP = pad_of(G0);
  p.lines = N1 * N2;
  p.crop0 = g.crop[0];
  p.shift0 = g.shift_inv[0];
  p.coils = coils;
  x_tiling(G0, p.N0, coils, &p.Cg, &p.Q);
  p.m_g0 = qdiv_const(G0);
  p.m_n0 = qdiv_const(p.N0);
  p.scale = g.fft_scale;
  p.apply_deapo = stop == 1 ? 0 : 1;
  p.accumulate = accumulate ? 1 : 0;
  p.sign = sign;
  p.tw = tw + tw_off[0][1];
  FftRadices r;
  if (!make_radices(G0, &r)) return false;
  p.st = stage_table(tw, tw_off, 0);
  p.n_st = r.n;
  return true;
}

// Pass descriptors of the forward: image -> grid [coils][G2][G1][G0]; the x
// pass, then ax[0 .. d-2] (y, z).
bool build_forward(const Geom& g, const float2* img, bool img_per_coil, const float2* sens,
                   const float* deapo, float2* grid, float2* tmp, int coils, const float2* tw,
                   const size_t tw_off[3][3], FirstPass* first, AxisPass ax[2]) {
  const float sign = -1.0f;
  const int G0 = g.G[0], G1 = g.G[1];
  const int N1 = g.N[1], N2 = g.dim == 3 ? g.N[2] : 1;
  FirstPass& f = *first;
  f = FirstPass{};
  f.img = img;
  f.deapo = deapo;
  f.sens = sens;
  f.img_per_coil = img_per_coil ? 1 : 0;
  f.G0 = G0;
  f.N0 = g.N[0];
  f.LP = pad_of(G0);
  f.lines = N1 * N2;
  f.crop0 = g.crop[0];
  f.shift0 = g.shift_inv[0];
  f.coils = coils;
  x_tiling(G0, f.N0, coils, &f.Cg, &f.Q);
  f.m_g0 = qdiv_const(G0);
  f.m_n0 = qdiv_const(f.N0);
  f.m_pad = G0 > f.N0 ? qdiv_const(G0 - f.N0) : 0;
  f.sign = sign;
  f.tw = tw + tw_off[0][0];
  FftRadices r;
  if (!make_radices(G0, &r)) return false;
  f.st = stage_table(tw, tw_off, 0);
  f.n_st = r.n;
  // 3-D: x -> grid (as [c][N2][N1][G0]), y -> tmp, z -> grid
  // 2-D: x -> tmp, y -> grid
  f.out = g.dim == 3 ? grid : tmp;
  // y: [c][N2][N1][G0] -> [c][N2][G1][G0]
  AxisPass& y = ax[0];
  y = AxisPass{};
  y.in = f.out;
  y.out = g.dim == 3 ? tmp : grid;
  y.in_coil = (long long)G0 * N1 * N2;
  y.out_coil = (long long)G0 * G1 * N2;
  y.in_step = y.out_step = G0;
  y.in_ob = (long long)G0 * N1;
  y.out_ob = (long long)G0 * G1;
  y.nx = G0;
  y.nob = N2;
  y.L = G1;
  y.in_len = N1;
  y.out_len = G1;
  y.in_scatter = 1;
  y.in_off = g.crop[1] - g.shift_inv[1];
  y.sign = sign;
  y.tw = tw + tw_off[1][0];
  if (!finish_axis(&y, coils, stage_table(tw, tw_off, 1))) return false;
  if (g.dim == 2) return true;
  // z: [c][N2][G1][G0] -> [c][G2][G1][G0]
  AxisPass& z = ax[1];
  z = AxisPass{};
  z.in = tmp;
  z.out = grid;
  z.in_coil = (long long)G0 * G1 * N2;
  z.out_coil = (long long)G0 * G1 * g.G[2];
  z.in_step = z.out_step = (long long)G0 * G1;
  z.in_ob = z.out_ob = G0;
  z.nx = G0;
  z.nob = G1;
  z.L = g.G[2];
  z.in_len = N2;
  z.out_len = g.G[2];
  z.in_scatter = 1;
  z.in_off = g.crop[2] - g.shift_inv[2];
  z.sign = sign;
  z.tw = tw + tw_off[2][0];
  return finish_axis(&z, coils, stage_table(tw, tw_off, 2));
}

}  // namespace

bool fused_fft_supported(const Geom& g) {
  FftRadices r;
  for (int a = 0; a < g.dim; a++) {
    // one padded line fits a tile; the x tiling needs N0 <= G0 <= kMaxLen
    if (g.G[a] > kMaxLen || !make_radices(g.G[a], &r)) return false;
    if (g.N[a] > g.G[a]) return false;
  }
  return true;
}

// Default choice: lines of at most 1023 points, so that a y / z tile holds
// at least 16 lines (128-byte rows); longer lines leave shorter rows and
// rocFFT's kernels are faster there.
bool fused_fft_preferred(const Geom& g) {
  if (!fused_fft_supported(g)) return false;
  for (int a = 0; a < g.dim; a++)
    if (16 * pad_of(g.G[a]) > kBigTileElems) return false;
  return true;
}

bool fused_fft_prime(const Geom& g) {
  for (int a = 0; a < g.dim; a++)
    if (prime_radix(g.G[a])) return true;
  return false;
}

size_t fused_fft_tmp_elems(const Geom& g) {
  // adjoint z-pass / forward y-pass output (3-D), x-pass output (2-D)
  return g.dim == 3 ? (size_t)g.G[0] * g.G[1] * g.N[2] : (size_t)g.G[0] * g.N[1];
}

// Device tables of the fused FFT, one float2 array: per axis and sign the
// twiddles exp(sign 2 pi i k / G_a), k < G_a, computed in double
// (offsets[a][0 | 1]), then per axis its Stockham stage table
// (offsets[a][2], FftStage records; kernels read them with scalar loads).
std::vector<float2> fused_fft_twiddles(const Geom& g, size_t offsets[3][3]) {
  static_assert(sizeof(FftStage) % sizeof(float2) == 0, "stage records pack into float2 slots");
  std::vector<float2> t;
  for (int a = 0; a < 3; a++) offsets[a][0] = offsets[a][1] = offsets[a][2] = 0;
  for (int a = 0; a < g.dim; a++)
    for (int si = 0; si < 2; si++) {
      offsets[a][si] = t.size();
      const double sign = si == 0 ? -1.0 : 1.0;
      const int L = g.G[a];
      for (int k = 0; k < L; k++) {
        const double ang = sign * 2.0 * M_PI * (double)k / (double)L;
        t.push_back(make_float2((float)std::cos(ang), (float)std::sin(ang)));
      }
    }
  for (int a = 0; a < g.dim; a++) {
    FftRadices r;
    make_radices(g.G[a], &r);
    offsets[a][2] = t.size();
    const size_t slots = sizeof(r.st) / sizeof(float2);
    t.resize(t.size() + slots);
    std::memcpy(t.data() + offsets[a][2], r.st, sizeof(r.st));
  }
  return t;
}

// Adjoint: grid [coils][G2][G1][G0] (the INVERSE-shifted spectrum) -> image.
// tmp holds coils * fused_fft_tmp_elems(g) elements; the grid is
// overwritten. stop == 1 (FFT): per-coil images scaled by 1/sqrt(Ng) only.
// accumulate: add the coil sum to `out` (later coil batches).
hipError_t fused_fft_adjoint(const Geom& g, float2* grid, float2* tmp, int coils,
                             const float* deapo, const float2* sens, float2* out, int stop,
                             bool accumulate, const float2* tw, const size_t tw_off[3][3],
                             hipStream_t s) {
  AxisPass ax[2];
  LastPass last;
  if (!build_adjoint(g, grid, tmp, coils, deapo, sens, out, stop, accumulate, tw, tw_off, ax,
                     &last))
    return hipErrorInvalidValue;
  for (int i = 0; i < g.dim - 1; i++) {
    hipError_t e = launch_axis(ax[i], s);
    if (e != hipSuccess) return e;
  }
  return launch_last(last, s);
}

// Forward: image -> grid [coils][G2][G1][G0], the spectrum of the
// zero-padded, INVERSE-shifted, deapodized (and coil-weighted) image. tmp
// holds coils * fused_fft_tmp_elems(g) elements.
hipError_t fused_fft_forward(const Geom& g, const float2* img, bool img_per_coil,
                             const float2* sens, const float* deapo, float2* grid, float2* tmp,
                             int coils, const float2* tw, const size_t tw_off[3][3],
                             hipStream_t s) {
  FirstPass first;
  AxisPass ax[2];
  if (!build_forward(g, img, img_per_coil, sens, deapo, grid, tmp, coils, tw, tw_off, &first, ax))
    return hipErrorInvalidValue;
  hipError_t e = launch_first(first, s);
  for (int i = 0; i < g.dim - 1 && e == hipSuccess; i++) e = launch_axis(ax[i], s);
  return e;
}

// ---- batches --------------------------------------------------------------

size_t fused_fft_desc_bytes() { return sizeof(FftOpDesc); }

// Descriptor of one operator of a batch: its grid and intermediate, the
// full adjoint (DEAPODIZATION stop, coil sum when sens) and forward. The
// image pointers are set per call (fused_fft_desc_set_io).
bool fused_fft_desc_build(void* desc, const Geom& g, float2* grid, float2* tmp, int coils,
                          const float* deapo, const float2* sens, const float2* tw,
                          const size_t tw_off[3][3]) {
  FftOpDesc& d = *static_cast<FftOpDesc*>(desc);
  d = FftOpDesc{};
  d.n_ax = g.dim - 1;
  return build_adjoint(g, grid, tmp, coils, deapo, sens, nullptr, 2, false, tw, tw_off, d.adj_ax,
                       &d.adj_x) &&
         build_forward(g, nullptr, sens == nullptr, sens, deapo, grid, tmp, coils, tw, tw_off,
                       &d.fwd_x, d.fwd_ax);
}

void fused_fft_desc_set_io(void* desc, float2* adjoint_out, const float2* forward_img) {
  FftOpDesc& d = *static_cast<FftOpDesc*>(desc);
  d.adj_x.out = adjoint_out;
  d.fwd_x.img = forward_img;
}

// Passes in launch order: adjoint ax[0 .. d-2], x; forward x, ax[0 .. d-2].
int fused_fft_pass_count(int dim) { return dim; }

uint32_t fused_fft_pass_blocks(const void* desc, bool adjoint, int pass) {
  const FftOpDesc& d = *static_cast<const FftOpDesc*>(desc);
  if (adjoint) return pass < d.n_ax ? axis_blocks(d.adj_ax[pass]) : last_blocks(d.adj_x);
  return pass == 0 ? first_blocks(d.fwd_x) : axis_blocks(d.fwd_ax[pass - 1]);
}

// 1 when the pass of this operator runs on the large-tile kernel (a y / z
// pass of lines above 511 points)
int fused_fft_pass_big(const void* desc, bool adjoint, int pass) {
  const FftOpDesc& d = *static_cast<const FftOpDesc*>(desc);
  if (adjoint) return pass < d.n_ax ? d.adj_ax[pass].big : 0;
  return pass == 0 ? 0 : d.fwd_ax[pass - 1].big;
}

size_t fused_fft_pass_lds(const void* desc, bool adjoint, int pass) {
  const FftOpDesc& d = *static_cast<const FftOpDesc*>(desc);
  if (adjoint) return pass < d.n_ax ? axis_lds(d.adj_ax[pass]) : last_lds(d.adj_x);
  return pass == 0 ? first_lds(d.fwd_x) : axis_lds(d.fwd_ax[pass - 1]);
}

// One pass of every operator: items (operator, block), descs in device memory.
// PR: the kernel variants with the prime stages 7 .. 19.
template <bool PR>
hipError_t launch_fft_batch_pass(int dim, bool adjoint, int pass, bool big, const FftOpDesc* d,
                                 const uint2* items, uint32_t n_items, size_t lds, hipStream_t s) {
  const int n_ax = dim - 1;
  hipError_t e = hipSuccess;
  if (adjoint && pass == n_ax) {
    e = allow_lds(fft_adjoint_last_batch_kernel<PR>, lds);
    if (e == hipSuccess)
      hipLaunchKernelGGL(fft_adjoint_last_batch_kernel<PR>, dim3(n_items), dim3(kFftThreads), lds,
                         s, d, items);
  } else if (!adjoint && pass == 0) {
    e = allow_lds(fft_forward_first_batch_kernel<PR>, lds);
    if (e == hipSuccess)
      hipLaunchKernelGGL(fft_forward_first_batch_kernel<PR>, dim3(n_items), dim3(kFftThreads), lds,
                         s, d, items);
  } else if (big) {
    e = allow_lds(fft_axis_batch_big_kernel<PR>, lds);
    if (e == hipSuccess)
      hipLaunchKernelGGL(fft_axis_batch_big_kernel<PR>, dim3(n_items), dim3(kBigThreads), lds, s,
                         d, items, adjoint ? 1 : 0, adjoint ? pass : pass - 1);
  } else {
    e = allow_lds(fft_axis_batch_kernel<PR>, lds);
    if (e == hipSuccess)
      hipLaunchKernelGGL(fft_axis_batch_kernel<PR>, dim3(n_items), dim3(kFftThreads), lds, s, d,
                         items, adjoint ? 1 : 0, adjoint ? pass : pass - 1);
  }
  return e != hipSuccess ? e : hipGetLastError();
}

hipError_t launch_fused_fft_batch(int dim, bool adjoint, int pass, bool big, bool prime,
                                  const void* descs_d, const uint2* items, uint32_t n_items,
                                  size_t lds, hipStream_t s) {
  if (n_items == 0) return hipSuccess;
  const FftOpDesc* d = static_cast<const FftOpDesc*>(descs_d);
  if (prime) return launch_fft_batch_pass<true>(dim, adjoint, pass, big, d, items, n_items, lds, s);
  return launch_fft_batch_pass<false>(dim, adjoint, pass, big, d, items, n_items, lds, s);
}

}  // namespace gnufft
