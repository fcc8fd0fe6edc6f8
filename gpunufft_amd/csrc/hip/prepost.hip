// Bandwidth-bound stages around the oversampled FFT, each fused into one
// pass over memory (reference stages: CUDA/src/gpu/std_gpuNUFFT_kernels.cu
// fftScale :114-140, sens mul/sum :142-218, crop :341-372, fftshift
// :374-462/:569-654, precomputed deapodization :493-521, padding :699-793).
//
// The fftshifts are index permutations folded into these kernels and into
// the gridding kernels; the FFT itself is rocFFT.

#include "gnufft_internal.hpp"

namespace gnufft {

__global__ void __launch_bounds__(256) zero_kernel(float4* __restrict__ p, size_t n4,
                                                   float* __restrict__ tail, int ntail) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n4; i += stride) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  if (blockIdx.x == 0 && (int)threadIdx.x < ntail) tail[threadIdx.x] = 0.0f;
}

hipError_t launch_zero(void* p, size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  // callers pass float-granular, 16-byte aligned buffers
  size_t nf = bytes / 4;
  size_t n4 = nf / 4;
  int ntail = (int)(nf - n4 * 4);
  size_t blocks = (n4 + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (float4*)p, n4,
                     (float*)p + n4 * 4, ntail);
  return hipGetLastError();
}

// Last axis of the deapodization DFT (the plan's set-up): pixel (z, y, x)
// = norm / |sum_t B[t][y][x] E[t][z]| in double, B the transform of the
// separable tap tensor over x and y, E the phases along z (one row of ones
// in 2-D).
__global__ void __launch_bounds__(256) deapo_combine_kernel(const double2* __restrict__ B,
                                                            const double2* __restrict__ E, int T,
                                                            int Nx, int Ny, int Nz, double norm,
                                                            float* __restrict__ out) {
  const size_t plane = (size_t)Nx * Ny;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= plane * Nz) return;
  const int z = (int)(i / plane);
  const size_t yx = i - (size_t)z * plane;
  double re = 0.0, im = 0.0;
  for (int t = 0; t < T; t++) {
    const double2 b = B[(size_t)t * plane + yx];
    const double2 e = E[(size_t)t * Nz + z];
    re += b.x * e.x - b.y * e.y;
    im += b.x * e.y + b.y * e.x;
  }
  out[i] = (float)(norm / hypot(re, im));
}

hipError_t launch_deapo_combine(const double2* B, const double2* E, int T, int Nx, int Ny, int Nz,
                                double norm, float* out, hipStream_t s) {
  const size_t n = (size_t)Nx * Ny * Nz;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(deapo_combine_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, B, E,
                     T, Nx, Ny, Nz, norm, out);
  return hipGetLastError();
}

// Image pixel -> (x, y, z) and the grid index after crop offset and the
// INVERSE fftshift (the second shift of the adjoint, before crop).
__device__ inline void pixel_coords(const Geom& g, long long p, int* x, int* y, int* z) {
  *x = (int)(p % g.N[0]);
  long long r = p / g.N[0];
  *y = (int)(r % g.N[1]);
  *z = (int)(r / g.N[1]);
}

// Adjoint tail: crop + INVERSE fftshift (index map) + 1/sqrt(Ng) scale
// [+ deapodization] [+ conj(sens) multiply and coil sum].
//   stop == FFT:  out[c][p] = grid[c][m(p)] * scale
//   otherwise:    v_c = grid[c][m(p)] * scale * deapo[p]
//                 sens:   out[p] (+)= sum_c conj(S_c[p]) v_c   (accumulate: add to out)
//                 no sens: out[c][p] = v_c
__device__ inline void adjoint_post_body(const Geom& g, const float2* __restrict__ grid,
                                         int n_coils, const float* __restrict__ deapo,
                                         const float2* __restrict__ sens, float2* __restrict__ out,
                                         int accumulate, int stop, long long p) {
  if (p >= g.Ni) return;
  int x, y, z;
  pixel_coords(g, p, &x, &y, &z);
  const int m0 = wrap(x + g.crop[0] + g.shift_inv[0], g.G[0]);
  const int m1 = wrap(y + g.crop[1] + g.shift_inv[1], g.G[1]);
  const int m2 = g.dim == 3 ? wrap(z + g.crop[2] + g.shift_inv[2], g.G[2]) : 0;
  const size_t gi = (size_t)m0 + (size_t)g.G[0] * ((size_t)m1 + (size_t)g.G[1] * m2);
  const float sc = g.fft_scale;
  if (stop == 1) {  // FFT
    for (int c = 0; c < n_coils; c++) {
      float2 v = grid[(size_t)c * g.Ng + gi];
      out[(size_t)c * g.Ni + p] = make_float2(v.x * sc, v.y * sc);
    }
    return;
  }
  const float dp = deapo[p];
  if (sens) {
    float2 acc = accumulate ? out[p] : make_float2(0.f, 0.f);
    for (int c = 0; c < n_coils; c++) {
      float2 v = grid[(size_t)c * g.Ng + gi];
      v.x = (v.x * sc) * dp;
      v.y = (v.y * sc) * dp;
      const float2 sv = sens[(size_t)c * g.Ni + p];
      acc.x += v.x * sv.x + v.y * sv.y;
      acc.y += v.y * sv.x - v.x * sv.y;
    }
    out[p] = acc;
  } else {
    for (int c = 0; c < n_coils; c++) {
      float2 v = grid[(size_t)c * g.Ng + gi];
      out[(size_t)c * g.Ni + p] = make_float2((v.x * sc) * dp, (v.y * sc) * dp);
    }
  }
}

__global__ void __launch_bounds__(256) adjoint_post_kernel(Geom g, const float2* __restrict__ grid,
                                                           int n_coils,
                                                           const float* __restrict__ deapo,
                                                           const float2* __restrict__ sens,
                                                           float2* __restrict__ out,
                                                           int accumulate, int stop) {
  adjoint_post_body(g, grid, n_coils, deapo, sens, out, accumulate, stop,
                    (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

hipError_t launch_adjoint_post(const Geom& g, const float2* grid, int n_coils,
                               const float* deapo, const float2* sens, float2* out,
                               bool accumulate, int stop, hipStream_t s) {
  unsigned blocks = (unsigned)((g.Ni + 255) / 256);
  hipLaunchKernelGGL(adjoint_post_kernel, dim3(blocks), dim3(256), 0, s, g, grid, n_coils, deapo,
                     sens, out, accumulate ? 1 : 0, stop);
  return hipGetLastError();
}

// Forward head: [x S_c] x deapo, zero-pad, INVERSE fftshift: writes every
// grid point of every coil (no separate memset).
__device__ inline void forward_pre_body(const Geom& g, const float2* __restrict__ img,
                                        int img_per_coil, const float2* __restrict__ sens,
                                        const float* __restrict__ deapo, float2* __restrict__ grid,
                                        int n_coils, long long t) {
  if (t >= g.Ng) return;
  const int tx = (int)(t % g.G[0]);
  const long long r = t / g.G[0];
  const int ty = (int)(r % g.G[1]);
  const int tz = (int)(r / g.G[1]);
  const int px = wrap(tx + g.shift_inv[0], g.G[0]) - g.crop[0];
  const int py = wrap(ty + g.shift_inv[1], g.G[1]) - g.crop[1];
  const int pz = g.dim == 3 ? wrap(tz + g.shift_inv[2], g.G[2]) - g.crop[2] : 0;
  const bool inside = px >= 0 && px < g.N[0] && py >= 0 && py < g.N[1] && pz >= 0 && pz < g.N[2];
  if (!inside) {
    for (int c = 0; c < n_coils; c++) grid[(size_t)c * g.Ng + t] = make_float2(0.f, 0.f);
    return;
  }
  const long long p = px + (long long)g.N[0] * (py + (long long)g.N[1] * pz);
  const float dp = deapo[p];
  for (int c = 0; c < n_coils; c++) {
    float2 v = img[(img_per_coil ? (size_t)c * g.Ni : 0) + p];
    if (sens) {
      const float2 sv = sens[(size_t)c * g.Ni + p];
      float2 m;
      m.x = v.x * sv.x - v.y * sv.y;
      m.y = v.x * sv.y + v.y * sv.x;
      v = m;
    }
    grid[(size_t)c * g.Ng + t] = make_float2(v.x * dp, v.y * dp);
  }
}

__global__ void __launch_bounds__(256) forward_pre_kernel(Geom g, const float2* __restrict__ img,
                                                          int img_per_coil,
                                                          const float2* __restrict__ sens,
                                                          const float* __restrict__ deapo,
                                                          float2* __restrict__ grid, int n_coils) {
  forward_pre_body(g, img, img_per_coil, sens, deapo, grid, n_coils,
                   (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

hipError_t launch_forward_pre(const Geom& g, const float2* img, bool img_per_coil,
                              const float2* sens, const float* deapo, float2* grid, int n_coils,
                              hipStream_t s) {
  unsigned blocks = (unsigned)((g.Ng + 255) / 256);
  hipLaunchKernelGGL(forward_pre_kernel, dim3(blocks), dim3(256), 0, s, g, img,
                     img_per_coil ? 1 : 0, sens, deapo, grid, n_coils);
  return hipGetLastError();
}

// [x S_c] x deapo on the image (the DENSITY_ESTIMATION forward interpolates
// from this image-sized array directly, gpuNUFFT_operator.cpp:1169-1187).
__global__ void __launch_bounds__(256) image_prep_kernel(Geom g, const float2* __restrict__ img,
                                                         int img_per_coil,
                                                         const float2* __restrict__ sens,
                                                         const float* __restrict__ deapo,
                                                         float2* __restrict__ out, int n_coils) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.Ni) return;
  const float dp = deapo[p];
  for (int c = 0; c < n_coils; c++) {
    float2 v = img[(img_per_coil ? (size_t)c * g.Ni : 0) + p];
    if (sens) {
      const float2 sv = sens[(size_t)c * g.Ni + p];
      float2 m;
      m.x = v.x * sv.x - v.y * sv.y;
      m.y = v.x * sv.y + v.y * sv.x;
      v = m;
    }
    out[(size_t)c * g.Ni + p] = make_float2(v.x * dp, v.y * dp);
  }
}

hipError_t launch_image_prep(const Geom& g, const float2* img, bool img_per_coil,
                             const float2* sens, const float* deapo, float2* out, int n_coils,
                             hipStream_t s) {
  unsigned blocks = (unsigned)((g.Ni + 255) / 256);
  hipLaunchKernelGGL(image_prep_kernel, dim3(blocks), dim3(256), 0, s, g, img,
                     img_per_coil ? 1 : 0, sens, deapo, out, n_coils);
  return hipGetLastError();
}

// Batched: item (operator, block of 256 elements, 0, 0) per workgroup; the
// batch list is not XCD-partitioned (elementwise, no locality to keep).
__global__ void __launch_bounds__(256) post_batch_kernel(const BatchOp* __restrict__ ops,
                                                         const uint4* __restrict__ items) {
  const uint4 it = items[blockIdx.x];
  const BatchOp& op = ops[it.x];
  const PostArgs& a = op.post;
  adjoint_post_body(op.g, a.grid, a.n_coils, a.deapo, a.sens, a.out, 0, a.stop,
                    (long long)it.y * 256 + threadIdx.x);
}

__global__ void __launch_bounds__(256) pre_batch_kernel(const BatchOp* __restrict__ ops,
                                                        const uint4* __restrict__ items) {
  const uint4 it = items[blockIdx.x];
  const BatchOp& op = ops[it.x];
  const PreArgs& a = op.pre;
  forward_pre_body(op.g, a.img, a.img_per_coil, a.sens, a.deapo, a.grid, a.n_coils,
                   (long long)it.y * 256 + threadIdx.x);
}

hipError_t launch_post_batch(const BatchOp* ops, const BatchList& l, hipStream_t s) {
  if (l.n == 0) return hipSuccess;
  hipLaunchKernelGGL(post_batch_kernel, dim3(l.n), dim3(256), 0, s, ops, l.items);
  return hipGetLastError();
}

hipError_t launch_pre_batch(const BatchOp* ops, const BatchList& l, hipStream_t s) {
  if (l.n == 0) return hipSuccess;
  hipLaunchKernelGGL(pre_batch_kernel, dim3(l.n), dim3(256), 0, s, ops, l.items);
  return hipGetLastError();
}

}  // namespace gnufft
