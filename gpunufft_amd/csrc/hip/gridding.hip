// Kaiser-Bessel gridding kernels for CDNA4 (gfx950).
//
// spread (adjoint convolution, reference CUDA/src/gpu/atomic/
//   atomic_gpuNUFFT_kernels.cu:302-617 and texture_gpuNUFFT_kernels.cu:22-376)
// interp (forward convolution, atomic_gpuNUFFT_kernels.cu:966-1400,
//   texture_gpuNUFFT_kernels.cu:467-1039)
//
// Work decomposition: one workgroup per (sector chunk, coil block). A chunk
// is a contiguous run of at most `chunk` sorted samples of one sector, so
// heavy sectors near the k-space centre spread over many workgroups.
// Workgroup ids are remapped so that each of the 8 XCDs owns a contiguous
// range of chunks (neighbouring sectors share that XCD's L2).
//
// Every workgroup first stages its chunk's coordinates (and, in the spread,
// the gathered, density-weighted sample values) into LDS with coalesced
// loads issued all at once, then computes from LDS: a wave never waits on
// a dependent global load per sample.
//
// The spread accumulates the padded sector sub-grid in LDS in double
// precision, laid out [point][re,im][coil]: on gfx950 ds_add_f64 sustains
// ~20x the rate of ds_add_f32, and double sums are at least as accurate as
// the reference's float atomics. A wave's lanes are (sample slot, coil):
// each lane precomputes its sample's per-axis tap weights once (Taps), then
// walks the kw^DIM taps adding value*weight into its coil's words.

#include "gnufft_internal.hpp"

namespace gnufft {

constexpr int kSpreadThreads = 512;
constexpr int kInterpThreads = 256;

__device__ inline uint32_t xcd_chunk(uint32_t bid, uint32_t n_blocks) {
  // Blocks are dispatched round-robin over the 8 XCDs (bid % 8).
  uint32_t per = n_blocks / 8;
  return (bid % 8) * per + bid / 8;
}

template <int DIM>
__device__ inline void load_center(const uint32_t* centers, uint32_t sector, int* c) {
  c[0] = (int)centers[DIM * sector];
  c[1] = (int)centers[DIM * sector + 1];
  c[2] = DIM == 3 ? (int)centers[DIM * sector + 2] : 0;
}

// Global grid index of padded-sector point pt, with periodic wrap and an
// optional fftshift offset folded in (sign +1: FORWARD read, -1: INVERSE store).
template <int DIM>
__device__ inline size_t padded_to_grid(const Geom& g, const int* c3, int pt, int shift_sign) {
  const int P = g.P;
  const int yz = (int)udiv((uint32_t)pt, g.mP);
  const int x = pt - yz * P;
  const int z = DIM == 3 ? (int)udiv((uint32_t)pt, g.mP2) : 0;
  const int y = yz - z * P;
  int q0 = c3[0] - g.off + x;
  int q1 = c3[1] - g.off + y;
  int q2 = DIM == 3 ? c3[2] - g.off + z : 0;
  if (shift_sign > 0) {
    q0 += g.shift_fwd[0];
    q1 += g.shift_fwd[1];
    if (DIM == 3) q2 += g.shift_fwd[2];
  } else if (shift_sign < 0) {
    q0 -= g.shift_inv[0];
    q1 -= g.shift_inv[1];
    if (DIM == 3) q2 -= g.shift_inv[2];
  }
  q0 = wrap(q0, g.G[0]);
  q1 = wrap(q1, g.G[1]);
  q2 = DIM == 3 ? wrap(q2, g.G[2]) : 0;
  return (size_t)q0 + (size_t)g.G[0] * ((size_t)q1 + (size_t)g.G[1] * q2);
}

// Stage coordinates [DIM][n] of samples [b, b+n) into LDS.
template <int DIM, int NT>
__device__ inline void stage_coords(const float* crds, int M, uint32_t b, int n, float* sc, int CH) {
  for (int d = 0; d < DIM; d++)
    for (int i = threadIdx.x; i < n; i += NT) sc[d * CH + i] = crds[(size_t)d * M + b + i];
}

// Taps of one sample along every axis: taps lo[a]+t for t < NX, weight
// w[a][t], bit t of ok[a] set when the tap contributes (inside the padded
// sector and, in nearest mode, inside the kernel radius).
template <int NX>
struct Taps {
  int lo[3];
  unsigned ok[3];
  float w[3][NX];
};

// kp: k-space position of padded-sector index i along axis a, stored at
// kp[a * KPS + i] (precomputed per workgroup; equals map_grid_to_kspace).
template <int DIM, int NX>
__device__ inline void sample_taps(const Geom& g, const float* lut, const float* kp, int KPS,
                                   const int* c3, const float* p, Taps<NX>& T) {
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const float u = map_kspace_to_grid(p[d], g.G[d], c3[d], g.off);
    int lo, hi;
    tap_range(u, g.Pmax, g.kradius, &lo, &hi);
    T.lo[d] = lo;
    unsigned ok = 0;
#pragma unroll
    for (int t = 0; t < NX; t++) {
      float w = 0.0f;
      if (lo + t <= hi && axis_weight_kp(g, lut, d, kp[d * KPS + lo + t], p[d], &w))
        ok |= 1u << t;
      T.w[d][t] = w;
    }
    T.ok[d] = ok;
  }
  if (DIM == 2) {
    T.lo[2] = 0;
    T.ok[2] = 1u;
    T.w[2][0] = 1.0f;
  }
}

// k-space positions of the padded sector's indices, per axis.
template <int DIM, int NT>
__device__ inline void fill_kpos(const Geom& g, const int* c3, float* kp, int KPS) {
  for (int e = threadIdx.x; e < DIM * KPS; e += NT) {
    const int d = e / KPS;
    const int i = e - d * KPS;
    kp[e] = map_grid_to_kspace(i, g.G[d], c3[d], g.off);
  }
}

// ---------------------------------------------------------------------------
// Spread
// ---------------------------------------------------------------------------
template <int DIM, int NX, int CB>
__global__ void __launch_bounds__(kSpreadThreads) spread_kernel(Geom g, SpreadArgs a, uint32_t n_blocks,
                                                                int CH) {
  constexpr int NT = kSpreadThreads;
  constexpr int NW = NT / 64;
  constexpr int SPW = 64 / CB;  // samples per wave per step; one lane per (sample, coil)

  extern __shared__ float smem[];
  const int P = g.P;
  const int PD = DIM == 3 ? P * P * P : P * P;
  const int KPS = P + NX;
  const int Kpad = (g.K + 3) & ~3;
  float* lut = smem;
  float* kp = lut + Kpad;                      // 3 * KPS, padded to 4
  double* sub = reinterpret_cast<double*>(kp + ((3 * KPS + 3) & ~3));
  float* sc = reinterpret_cast<float*>(sub + (size_t)PD * CB * 2);
  float2* sv = reinterpret_cast<float2*>(sc + DIM * CH);

  const uint32_t ci = xcd_chunk(blockIdx.x, n_blocks);
  if (ci >= a.n_chunks) return;
  const Chunk ch = a.chunks[ci];
  const int n = (int)(ch.end - ch.begin);
  const int coil0 = blockIdx.y * CB;
  const int tid = threadIdx.x;
  int c3[3];
  load_center<DIM>(a.centers, ch.sector, c3);

  for (int i = tid; i < g.K; i += NT) lut[i] = a.lut[i];
  fill_kpos<DIM, NT>(g, c3, kp, KPS);
  for (int i = tid; i < PD * CB * 2; i += NT) sub[i] = 0.0;
  stage_coords<DIM, NT>(a.crds, g.M, ch.begin, n, sc, CH);
  for (int c = 0; c < CB; c++) {
    const int cg = coil0 + c;
    for (int i = tid; i < n; i += NT) {
      float2 v = make_float2(0.0f, 0.0f);
      if (cg < a.n_coils) {
        const uint32_t s = ch.begin + i;
        const uint32_t src = a.gather ? a.gather[s] : s;
        v = a.data[(size_t)cg * a.coil_stride + src];
        if (a.dens) {
          const float sq = sqrtf(a.dens[s]);
          v.x = v.x * sq;
          v.y = v.y * sq;
        }
      }
      sv[i * CB + c] = v;
    }
  }
  __syncthreads();

  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int slot = lane / CB;
  const int c = lane % CB;

  for (int base = wave * SPW; base < n; base += NW * SPW) {
    const int ls = base + slot;
    if (ls >= n) continue;
    float p[3];
    p[0] = sc[ls];
    p[1] = sc[CH + ls];
    p[2] = DIM == 3 ? sc[2 * CH + ls] : 0.0f;
    Taps<NX> T;
    sample_taps<DIM, NX>(g, lut, kp, KPS, c3, p, T);
    const float2 v = sv[ls * CB + c];
    // taps in (z, y, x) order, weight (wz*wy)*wx as the reference forms it
#pragma unroll
    for (int tz = 0; tz < (DIM == 3 ? NX : 1); tz++) {
      if (!((T.ok[2] >> tz) & 1u)) continue;
      const float wz = T.w[2][tz];
      const int k = T.lo[2] + tz;
#pragma unroll
      for (int ty = 0; ty < NX; ty++) {
        if (!((T.ok[1] >> ty) & 1u)) continue;
        const float wzy = DIM == 3 ? wz * T.w[1][ty] : T.w[1][ty];
        const int row = DIM == 3 ? ((k * P + T.lo[1] + ty) * P + T.lo[0]) : ((T.lo[1] + ty) * P + T.lo[0]);
#pragma unroll
        for (int tx = 0; tx < NX; tx++) {
          if (!((T.ok[0] >> tx) & 1u)) continue;
          const float w = wzy * T.w[0][tx];
          double* dst = &sub[((row + tx) * 2) * CB + c];
          atomicAdd(dst, (double)(w * v.x));
          atomicAdd(dst + CB, (double)(w * v.y));
        }
      }
    }
  }
  __syncthreads();

  // Flush: element order (coil, point) so consecutive lanes write
  // consecutive grid words of one coil. Periodic wrap, optional INVERSE
  // fftshift folded into the store index (std_gpuNUFFT_kernels.cu:629-654).
  float* gridf = reinterpret_cast<float*>(a.grid);
  for (int e = tid; e < PD * CB; e += NT) {
    const int cc = (int)udiv((uint32_t)e, g.mPD);
    const int pt = e - cc * PD;
    const int cgf = coil0 + cc;
    if (cgf >= a.n_coils) continue;
    const float vr = (float)sub[(pt * 2) * CB + cc];
    const float vi = (float)sub[(pt * 2 + 1) * CB + cc];
    if (vr == 0.0f && vi == 0.0f) continue;
    const size_t gi = padded_to_grid<DIM>(g, c3, pt, a.fold_shift ? -1 : 0);
    float* dst = &gridf[((size_t)cgf * g.Ng + gi) * 2];
    if (vr != 0.0f) atomicAdd(dst, vr);
    if (vi != 0.0f) atomicAdd(dst + 1, vi);
  }
}

// ---------------------------------------------------------------------------
// Interp
// ---------------------------------------------------------------------------
template <int DIM, int NX, int CB>
__global__ void __launch_bounds__(kInterpThreads) interp_kernel(Geom g, InterpArgs a, uint32_t n_blocks,
                                                                int CH) {
  constexpr int NT = kInterpThreads;
  constexpr int NW = NT / 64;
  constexpr int SPW = 64 / CB;  // samples per wave per step; one lane per (sample, coil)
  extern __shared__ float smem[];
  const int P = g.P;
  const int PD = DIM == 3 ? P * P * P : P * P;
  const int KPS = P + NX;
  const int Kpad = (g.K + 3) & ~3;
  float* lut = smem;
  float* kp = lut + Kpad;
  float2* sub = reinterpret_cast<float2*>(kp + ((3 * KPS + 3) & ~3));
  float* sc = reinterpret_cast<float*>(sub + (size_t)PD * CB);

  const uint32_t ci = xcd_chunk(blockIdx.x, n_blocks);
  if (ci >= a.n_chunks) return;
  const Chunk ch = a.chunks[ci];
  const int n = (int)(ch.end - ch.begin);
  const int coil0 = blockIdx.y * CB;
  const int tid = threadIdx.x;
  int c3[3];
  load_center<DIM>(a.centers, ch.sector, c3);

  for (int i = tid; i < g.K; i += NT) lut[i] = a.lut[i];
  fill_kpos<DIM, NT>(g, c3, kp, KPS);
  stage_coords<DIM, NT>(a.crds, g.M, ch.begin, n, sc, CH);
  // Stage the padded sector (periodic wrap, FORWARD fftshift folded into the
  // read index), coil-major so consecutive lanes read consecutive grid words.
  for (int e = tid; e < PD * CB; e += NT) {
    const int cc = (int)udiv((uint32_t)e, g.mPD);
    const int pt = e - cc * PD;
    const int cgf = coil0 + cc;
    float2 v = make_float2(0.0f, 0.0f);
    if (cgf < a.n_coils) v = a.grid[(size_t)cgf * g.Ng + padded_to_grid<DIM>(g, c3, pt, a.fold_shift ? 1 : 0)];
    sub[pt * CB + cc] = v;
  }
  __syncthreads();

  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int slot = lane / CB;
  const int c = lane % CB;
  const int cg = coil0 + c;

  for (int base = wave * SPW; base < n; base += NW * SPW) {
    const int ls = base + slot;
    if (ls >= n || cg >= a.n_coils) continue;
    float p[3];
    p[0] = sc[ls];
    p[1] = sc[CH + ls];
    p[2] = DIM == 3 ? sc[2 * CH + ls] : 0.0f;
    Taps<NX> T;
    sample_taps<DIM, NX>(g, lut, kp, KPS, c3, p, T);
    float accx = 0.0f, accy = 0.0f;
#pragma unroll
    for (int tz = 0; tz < (DIM == 3 ? NX : 1); tz++) {
      if (!((T.ok[2] >> tz) & 1u)) continue;
      const float wz = T.w[2][tz];
      const int k = T.lo[2] + tz;
#pragma unroll
      for (int ty = 0; ty < NX; ty++) {
        if (!((T.ok[1] >> ty) & 1u)) continue;
        const float wzy = DIM == 3 ? wz * T.w[1][ty] : T.w[1][ty];
        const int row = DIM == 3 ? ((k * P + T.lo[1] + ty) * P + T.lo[0]) : ((T.lo[1] + ty) * P + T.lo[0]);
#pragma unroll
        for (int tx = 0; tx < NX; tx++) {
          if (!((T.ok[0] >> tx) & 1u)) continue;
          const float val = wzy * T.w[0][tx];
          const float2 v = sub[(row + tx) * CB + c];
          accx += v.x * val;
          accy += v.y * val;
        }
      }
    }
    const uint32_t s = ch.begin + ls;
    float rx = accx * a.scale, ry = accy * a.scale;
    if (a.dens) {
      const float sq = sqrtf(a.dens[s]);
      rx = rx * sq;
      ry = ry * sq;
    }
    const uint32_t dst = a.scatter ? a.scatter[s] : s;
    a.out[(size_t)cg * a.coil_stride + dst] = make_float2(rx, ry);
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
// LDS bytes: lookup table, the padded sub-grid (bytes_per per point and
// coil: 16 = double complex in the spread, 8 = float complex in the interp),
// staged coordinates and, for the spread, staged sample values.
static size_t lds_bytes(const Geom& g, int cb, int bytes_per, int chunk, bool values) {
  const int P = g.P;
  const size_t PD = g.dim == 3 ? (size_t)P * P * P : (size_t)P * P;
  const int NX = g.kw + 1 <= 4 ? 4 : 8;
  size_t b = (size_t)((g.K + 3) & ~3) * sizeof(float) + PD * cb * bytes_per;
  b += (size_t)((3 * (P + NX) + 3) & ~3) * sizeof(float);
  b += (size_t)g.dim * chunk * sizeof(float);
  if (values) b += (size_t)chunk * cb * sizeof(float2);
  return b;
}

// Dynamic LDS a workgroup may use on the current device (160 KB per CU on
// gfx950; queried, since launches above 64 KB need the device to allow it).
static size_t lds_limit() {
  static size_t limit = 0;
  if (limit == 0) {
    int dev = 0, v = 0;
    size_t l = 64 * 1024;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
        v > 0)
      l = (size_t)v;
    limit = l;
  }
  return limit;
}

template <typename K>
static hipError_t allow_lds(K kernel, size_t lds) {
  if (lds <= 64 * 1024) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

template <int DIM, int NX, int CB>
static hipError_t spread_launch(const Geom& g, const SpreadArgs& a, int chunk, hipStream_t s) {
  const uint32_t n_blocks = ((a.n_chunks + 7) / 8) * 8;
  const int coil_blocks = (a.n_coils + CB - 1) / CB;
  const size_t lds = lds_bytes(g, CB, 16, chunk, true);
  hipError_t e = allow_lds(spread_kernel<DIM, NX, CB>, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((spread_kernel<DIM, NX, CB>), dim3(n_blocks, coil_blocks),
                     dim3(kSpreadThreads), lds, s, g, a, n_blocks, chunk);
  return hipGetLastError();
}

template <int DIM, int NX, int CB>
static hipError_t interp_launch(const Geom& g, const InterpArgs& a, int chunk, hipStream_t s) {
  const uint32_t n_blocks = ((a.n_chunks + 7) / 8) * 8;
  const int coil_blocks = (a.n_coils + CB - 1) / CB;
  const size_t lds = lds_bytes(g, CB, 8, chunk, false);
  hipError_t e = allow_lds(interp_kernel<DIM, NX, CB>, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((interp_kernel<DIM, NX, CB>), dim3(n_blocks, coil_blocks),
                     dim3(kInterpThreads), lds, s, g, a, n_blocks, chunk);
  return hipGetLastError();
}

// Largest coil block whose LDS footprint fits the device budget.
static int pick_cb(const Geom& g, int n_coils, int max_cb, int bytes_per, int chunk, bool values) {
  int cb = max_cb;
  while (cb > 1 &&
         (lds_bytes(g, cb, bytes_per, chunk, values) > lds_limit() || cb / 2 >= n_coils))
    cb /= 2;
  return cb;
}

template <int DIM>
static hipError_t spread_dispatch(const Geom& g, const SpreadArgs& a, int chunk, hipStream_t s) {
  if (g.kw + 1 <= 4) {
    switch (pick_cb(g, a.n_coils, 8, 16, chunk, true)) {
      case 8: return spread_launch<DIM, 4, 8>(g, a, chunk, s);
      case 4: return spread_launch<DIM, 4, 4>(g, a, chunk, s);
      case 2: return spread_launch<DIM, 4, 2>(g, a, chunk, s);
      default: return spread_launch<DIM, 4, 1>(g, a, chunk, s);
    }
  }
  switch (pick_cb(g, a.n_coils, 4, 16, chunk, true)) {
    case 4: return spread_launch<DIM, 8, 4>(g, a, chunk, s);
    case 2: return spread_launch<DIM, 8, 2>(g, a, chunk, s);
    default: return spread_launch<DIM, 8, 1>(g, a, chunk, s);
  }
}

template <int DIM>
static hipError_t interp_dispatch(const Geom& g, const InterpArgs& a, int chunk, hipStream_t s) {
  if (g.kw + 1 <= 4) {
    switch (pick_cb(g, a.n_coils, 8, 8, chunk, false)) {
      case 8: return interp_launch<DIM, 4, 8>(g, a, chunk, s);
      case 4: return interp_launch<DIM, 4, 4>(g, a, chunk, s);
      case 2: return interp_launch<DIM, 4, 2>(g, a, chunk, s);
      default: return interp_launch<DIM, 4, 1>(g, a, chunk, s);
    }
  }
  switch (pick_cb(g, a.n_coils, 8, 8, chunk, false)) {
    case 8: return interp_launch<DIM, 8, 8>(g, a, chunk, s);
    case 4: return interp_launch<DIM, 8, 4>(g, a, chunk, s);
    case 2: return interp_launch<DIM, 8, 2>(g, a, chunk, s);
    default: return interp_launch<DIM, 8, 1>(g, a, chunk, s);
  }
}

hipError_t launch_spread(const Geom& g, const SpreadArgs& a, int chunk, hipStream_t s) {
  if (a.n_chunks == 0 || a.n_coils == 0) return hipSuccess;
  if (g.kw > 7 || lds_bytes(g, 1, 16, chunk, true) > lds_limit()) return hipErrorInvalidValue;
  return g.dim == 3 ? spread_dispatch<3>(g, a, chunk, s) : spread_dispatch<2>(g, a, chunk, s);
}

hipError_t launch_interp(const Geom& g, const InterpArgs& a, int chunk, hipStream_t s) {
  if (a.n_chunks == 0 || a.n_coils == 0) return hipSuccess;
  if (g.kw > 7 || lds_bytes(g, 1, 8, chunk, false) > lds_limit()) return hipErrorInvalidValue;
  return g.dim == 3 ? interp_dispatch<3>(g, a, chunk, s) : interp_dispatch<2>(g, a, chunk, s);
}

}  // namespace gnufft
