// Kaiser-Bessel gridding kernels for CDNA4 (gfx950).
//
// spread (adjoint convolution, reference CUDA/src/gpu/atomic/
//   atomic_gpuNUFFT_kernels.cu:302-617 and texture_gpuNUFFT_kernels.cu:22-376)
// interp (forward convolution, atomic_gpuNUFFT_kernels.cu:966-1400,
//   texture_gpuNUFFT_kernels.cu:467-1039)
//
// Work decomposition: one workgroup (4 wave64) per (sector chunk, coil block).
// A chunk is a contiguous run of at most `chunk` sorted samples of one
// sector, so heavy sectors near the k-space centre spread over many
// workgroups. Workgroup ids are remapped so that each of the 8 XCDs owns a
// contiguous range of chunks (neighbouring sectors share that XCD's L2).
//
// LDS holds the lookup table and the padded sector sub-grid laid out
// [point][coil][re,im]. In the spread, the lanes of one sample cover
// (x tap, coil, re/im): for a fixed (z, y) row they address consecutive
// words, so every ds_add_f32 wave instruction is bank-conflict free.

#include "gnufft_internal.hpp"

namespace gnufft {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

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

__device__ inline float sample_coord(const float* crds, uint32_t s, int a, int M) {
  return crds[s + (size_t)a * M];
}

// ---------------------------------------------------------------------------
// Spread
// ---------------------------------------------------------------------------
template <int DIM, int NX, int CB>
__global__ void __launch_bounds__(kThreads) spread_kernel(Geom g, SpreadArgs a, uint32_t n_blocks) {
  constexpr int L = NX * CB * 2;  // lanes per sample
  constexpr int SPW = 64 / L;     // samples per wave per step
  static_assert(L <= 64 && (64 % L) == 0, "lane layout");

  extern __shared__ float smem[];
  const int Kpad = (g.K + 3) & ~3;
  float* lut = smem;
  float* sub = smem + Kpad;

  const uint32_t ci = xcd_chunk(blockIdx.x, n_blocks);
  if (ci >= a.n_chunks) return;
  const Chunk ch = a.chunks[ci];
  const int coil0 = blockIdx.y * CB;

  const int P = g.P;
  const int PD = DIM == 3 ? P * P * P : P * P;
  const int tid = threadIdx.x;

  for (int i = tid; i < g.K; i += kThreads) lut[i] = a.lut[i];
  for (int i = tid; i < PD * CB * 2; i += kThreads) sub[i] = 0.0f;
  __syncthreads();

  int c3[3];
  load_center<DIM>(a.centers, ch.sector, c3);

  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int slot = lane / L;
  const int r = lane % L;
  const int xt = r / (CB * 2);
  const int c = (r >> 1) % CB;
  const int ri = r & 1;
  const int cg = coil0 + c;
  const bool coil_ok = cg < a.n_coils;

  for (uint32_t base = ch.begin + (uint32_t)(wave * SPW); base < ch.end;
       base += (uint32_t)(kWaves * SPW)) {
    const uint32_t s = base + slot;
    if (s >= ch.end) continue;
    float p[3];
    p[0] = sample_coord(a.crds, s, 0, g.M);
    p[1] = sample_coord(a.crds, s, 1, g.M);
    p[2] = DIM == 3 ? sample_coord(a.crds, s, 2, g.M) : 0.0f;
    int lo[3], hi[3];
    for (int d = 0; d < DIM; d++) {
      float u = map_kspace_to_grid(p[d], g.G[d], c3[d], g.off);
      tap_range(u, g.Pmax, g.kradius, &lo[d], &hi[d]);
    }
    float val = 0.0f;
    if (coil_ok) {
      uint32_t src = a.gather ? a.gather[s] : s;
      float2 v = a.data[(size_t)cg * a.coil_stride + src];
      if (a.dens) {
        float sq = sqrtf(a.dens[s]);
        v.x = v.x * sq;
        v.y = v.y * sq;
      }
      val = ri ? v.y : v.x;
    }
    const int i = lo[0] + xt;
    float wx = 0.0f;
    const bool xok = (i <= hi[0]) && axis_weight(g, lut, 0, i, c3[0], p[0], &wx);
    if (DIM == 3) {
      for (int k = lo[2]; k <= hi[2]; k++) {
        float wz;
        if (!axis_weight(g, lut, 2, k, c3[2], p[2], &wz)) continue;
        for (int j = lo[1]; j <= hi[1]; j++) {
          float wy;
          if (!axis_weight(g, lut, 1, j, c3[1], p[1], &wy)) continue;
          if (xok) {
            float w = (wz * wy) * wx;
            atomicAdd(&sub[(((k * P + j) * P + i) * CB + c) * 2 + ri], w * val);
          }
        }
      }
    } else {
      for (int j = lo[1]; j <= hi[1]; j++) {
        float wy;
        if (!axis_weight(g, lut, 1, j, c3[1], p[1], &wy)) continue;
        if (xok) {
          float w = wy * wx;
          atomicAdd(&sub[((j * P + i) * CB + c) * 2 + ri], w * val);
        }
      }
    }
  }
  __syncthreads();

  // Flush: element order (coil, point, re/im) so consecutive lanes write
  // consecutive grid words of one coil. Periodic wrap, optional INVERSE
  // fftshift folded into the store index (std_gpuNUFFT_kernels.cu:629-654).
  const int E = PD * CB * 2;
  float* gridf = reinterpret_cast<float*>(a.grid);
  for (int e = tid; e < E; e += kThreads) {
    const int cc = e / (PD * 2);
    const int rem = e - cc * (PD * 2);
    const int pt = rem >> 1;
    const int rr = rem & 1;
    const int cgf = coil0 + cc;
    if (cgf >= a.n_coils) continue;
    const float v = sub[(pt * CB + cc) * 2 + rr];
    if (v == 0.0f) continue;
    const int x = pt % P;
    const int y = (pt / P) % P;
    const int z = DIM == 3 ? pt / (P * P) : 0;
    int q0 = wrap(c3[0] - g.off + x, g.G[0]);
    int q1 = wrap(c3[1] - g.off + y, g.G[1]);
    int q2 = DIM == 3 ? wrap(c3[2] - g.off + z, g.G[2]) : 0;
    if (a.fold_shift) {
      q0 = wrap(q0 - g.shift_inv[0], g.G[0]);
      q1 = wrap(q1 - g.shift_inv[1], g.G[1]);
      if (DIM == 3) q2 = wrap(q2 - g.shift_inv[2], g.G[2]);
    }
    const size_t gi = (size_t)q0 + (size_t)g.G[0] * ((size_t)q1 + (size_t)g.G[1] * q2);
    atomicAdd(&gridf[((size_t)cgf * g.Ng + gi) * 2 + rr], v);
  }
}

// ---------------------------------------------------------------------------
// Interp
// ---------------------------------------------------------------------------
template <int DIM, int NX, int CB>
__global__ void __launch_bounds__(kThreads) interp_kernel(Geom g, InterpArgs a, uint32_t n_blocks) {
  constexpr int SPW = 64 / CB;  // samples per wave per step; one lane per (sample, coil)
  extern __shared__ float smem[];
  const int Kpad = (g.K + 3) & ~3;
  float* lut = smem;
  float2* sub = reinterpret_cast<float2*>(smem + Kpad);

  const uint32_t ci = xcd_chunk(blockIdx.x, n_blocks);
  if (ci >= a.n_chunks) return;
  const Chunk ch = a.chunks[ci];
  const int coil0 = blockIdx.y * CB;

  const int P = g.P;
  const int PD = DIM == 3 ? P * P * P : P * P;
  const int tid = threadIdx.x;
  int c3[3];
  load_center<DIM>(a.centers, ch.sector, c3);

  for (int i = tid; i < g.K; i += kThreads) lut[i] = a.lut[i];
  // Stage the padded sector (periodic wrap, FORWARD fftshift folded into the
  // read index), coil-major so consecutive lanes read consecutive grid words.
  for (int e = tid; e < PD * CB; e += kThreads) {
    const int cc = e / PD;
    const int pt = e - cc * PD;
    const int cgf = coil0 + cc;
    float2 v = make_float2(0.0f, 0.0f);
    if (cgf < a.n_coils) {
      const int x = pt % P;
      const int y = (pt / P) % P;
      const int z = DIM == 3 ? pt / (P * P) : 0;
      int q0 = wrap(c3[0] - g.off + x, g.G[0]);
      int q1 = wrap(c3[1] - g.off + y, g.G[1]);
      int q2 = DIM == 3 ? wrap(c3[2] - g.off + z, g.G[2]) : 0;
      if (a.fold_shift) {
        q0 = wrap(q0 + g.shift_fwd[0], g.G[0]);
        q1 = wrap(q1 + g.shift_fwd[1], g.G[1]);
        if (DIM == 3) q2 = wrap(q2 + g.shift_fwd[2], g.G[2]);
      }
      const size_t gi = (size_t)q0 + (size_t)g.G[0] * ((size_t)q1 + (size_t)g.G[1] * q2);
      v = a.grid[(size_t)cgf * g.Ng + gi];
    }
    sub[pt * CB + cc] = v;
  }
  __syncthreads();

  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int slot = lane / CB;
  const int c = lane % CB;
  const int cg = coil0 + c;

  for (uint32_t base = ch.begin + (uint32_t)(wave * SPW); base < ch.end;
       base += (uint32_t)(kWaves * SPW)) {
    const uint32_t s = base + slot;
    if (s >= ch.end || cg >= a.n_coils) continue;
    float p[3];
    p[0] = sample_coord(a.crds, s, 0, g.M);
    p[1] = sample_coord(a.crds, s, 1, g.M);
    p[2] = DIM == 3 ? sample_coord(a.crds, s, 2, g.M) : 0.0f;
    int lo[3], hi[3];
    for (int d = 0; d < DIM; d++) {
      float u = map_kspace_to_grid(p[d], g.G[d], c3[d], g.off);
      tap_range(u, g.Pmax, g.kradius, &lo[d], &hi[d]);
    }
    float wx[NX];
    bool okx[NX];
#pragma unroll
    for (int t = 0; t < NX; t++) {
      const int i = lo[0] + t;
      wx[t] = 0.0f;
      okx[t] = (i <= hi[0]) && axis_weight(g, lut, 0, i, c3[0], p[0], &wx[t]);
    }
    float accx = 0.0f, accy = 0.0f;
    const int kz0 = DIM == 3 ? lo[2] : 0;
    const int kz1 = DIM == 3 ? hi[2] : 0;
    for (int k = kz0; k <= kz1; k++) {
      float wz = 1.0f;
      if (DIM == 3 && !axis_weight(g, lut, 2, k, c3[2], p[2], &wz)) continue;
      for (int j = lo[1]; j <= hi[1]; j++) {
        float wy;
        if (!axis_weight(g, lut, 1, j, c3[1], p[1], &wy)) continue;
        const float wzy = DIM == 3 ? wz * wy : wy;
        const int row = DIM == 3 ? (k * P + j) * P : j * P;
#pragma unroll
        for (int t = 0; t < NX; t++) {
          if (!okx[t]) continue;
          const float val = wzy * wx[t];
          const float2 v = sub[(row + lo[0] + t) * CB + c];
          accx += v.x * val;
          accy += v.y * val;
        }
      }
    }
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
static size_t lds_bytes(const Geom& g, int cb, int per_point_floats) {
  const int P = g.P;
  const size_t PD = g.dim == 3 ? (size_t)P * P * P : (size_t)P * P;
  return ((size_t)((g.K + 3) & ~3) + PD * cb * per_point_floats) * sizeof(float);
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
static hipError_t spread_launch(const Geom& g, const SpreadArgs& a, hipStream_t s) {
  const uint32_t n_blocks = ((a.n_chunks + 7) / 8) * 8;
  const int coil_blocks = (a.n_coils + CB - 1) / CB;
  const size_t lds = lds_bytes(g, CB, 2);
  hipError_t e = allow_lds(spread_kernel<DIM, NX, CB>, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((spread_kernel<DIM, NX, CB>), dim3(n_blocks, coil_blocks), dim3(kThreads),
                     lds, s, g, a, n_blocks);
  return hipGetLastError();
}

template <int DIM, int NX, int CB>
static hipError_t interp_launch(const Geom& g, const InterpArgs& a, hipStream_t s) {
  const uint32_t n_blocks = ((a.n_chunks + 7) / 8) * 8;
  const int coil_blocks = (a.n_coils + CB - 1) / CB;
  const size_t lds = lds_bytes(g, CB, 2);
  hipError_t e = allow_lds(interp_kernel<DIM, NX, CB>, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((interp_kernel<DIM, NX, CB>), dim3(n_blocks, coil_blocks), dim3(kThreads),
                     lds, s, g, a, n_blocks);
  return hipGetLastError();
}

// Largest coil block whose sub-grid fits the LDS budget.
static int pick_cb(const Geom& g, int n_coils, int max_cb) {
  int cb = max_cb;
  while (cb > 1 && (lds_bytes(g, cb, 2) > lds_limit() || cb / 2 >= n_coils)) cb /= 2;
  return cb;
}

template <int DIM>
static hipError_t spread_dispatch(const Geom& g, const SpreadArgs& a, hipStream_t s) {
  if (g.kw + 1 <= 4) {
    switch (pick_cb(g, a.n_coils, 8)) {
      case 8: return spread_launch<DIM, 4, 8>(g, a, s);
      case 4: return spread_launch<DIM, 4, 4>(g, a, s);
      case 2: return spread_launch<DIM, 4, 2>(g, a, s);
      default: return spread_launch<DIM, 4, 1>(g, a, s);
    }
  }
  switch (pick_cb(g, a.n_coils, 4)) {
    case 4: return spread_launch<DIM, 8, 4>(g, a, s);
    case 2: return spread_launch<DIM, 8, 2>(g, a, s);
    default: return spread_launch<DIM, 8, 1>(g, a, s);
  }
}

template <int DIM>
static hipError_t interp_dispatch(const Geom& g, const InterpArgs& a, hipStream_t s) {
  if (g.kw + 1 <= 4) {
    switch (pick_cb(g, a.n_coils, 8)) {
      case 8: return interp_launch<DIM, 4, 8>(g, a, s);
      case 4: return interp_launch<DIM, 4, 4>(g, a, s);
      case 2: return interp_launch<DIM, 4, 2>(g, a, s);
      default: return interp_launch<DIM, 4, 1>(g, a, s);
    }
  }
  switch (pick_cb(g, a.n_coils, 8)) {
    case 8: return interp_launch<DIM, 8, 8>(g, a, s);
    case 4: return interp_launch<DIM, 8, 4>(g, a, s);
    case 2: return interp_launch<DIM, 8, 2>(g, a, s);
    default: return interp_launch<DIM, 8, 1>(g, a, s);
  }
}

hipError_t launch_spread(const Geom& g, const SpreadArgs& a, hipStream_t s) {
  if (a.n_chunks == 0 || a.n_coils == 0) return hipSuccess;
  if (g.kw > 7 || lds_bytes(g, 1, 2) > lds_limit()) return hipErrorInvalidValue;
  return g.dim == 3 ? spread_dispatch<3>(g, a, s) : spread_dispatch<2>(g, a, s);
}

hipError_t launch_interp(const Geom& g, const InterpArgs& a, hipStream_t s) {
  if (a.n_chunks == 0 || a.n_coils == 0) return hipSuccess;
  if (g.kw > 7 || lds_bytes(g, 1, 2) > lds_limit()) return hipErrorInvalidValue;
  return g.dim == 3 ? interp_dispatch<3>(g, a, s) : interp_dispatch<2>(g, a, s);
}

}  // namespace gnufft
