// Sector precompute on the device: sample -> sector assignment, stable radix
// sort by sector, per-sector offsets, sector centres, sorted trajectory and
// density, and the gridding work list (sector chunks).
//
// Reference semantics: CUDA/inc/precomp_utils.hpp:34-153,
// CUDA/src/gpuNUFFT_operator_factory.cpp:128-270, CUDA/src/gpu/precomp_kernels.cu:8-177.

#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "gnufft_internal.hpp"

namespace gnufft {

__global__ void __launch_bounds__(256) assign_sectors_kernel(Geom g, const float* __restrict__ traj,
                                                             uint32_t M, uint32_t* __restrict__ keys,
                                                             uint32_t* __restrict__ vals) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M) return;
  int sx = sector_of(traj[t], g.G[0], g.sw);
  int sy = sector_of(traj[t + M], g.G[1], g.sw);
  int sz = 0;
  if (g.dim == 3) sz = sector_of(traj[t + 2 * (size_t)M], g.G[2], g.sw);
  keys[t] = (uint32_t)(sx + g.S[0] * (sy + g.S[1] * sz));
  vals[t] = t;
}

hipError_t launch_assign_sectors(const Geom& g, const float* traj, uint32_t M, uint32_t* keys,
                                 uint32_t* vals, hipStream_t s) {
  if (M == 0) return hipSuccess;
  unsigned blocks = (M + 255) / 256;
  hipLaunchKernelGGL(assign_sectors_kernel, dim3(blocks), dim3(256), 0, s, g, traj, M, keys, vals);
  return hipGetLastError();
}

// Stable LSD radix sort of (sector, original index) pairs: within a sector
// samples keep their original order, the canonical order of the reference's
// std::sort (equal per-sector index sets, CUDA/src/gpuNUFFT_operator_factory.cpp:56-72).
hipError_t sort_pairs(void* temp, size_t* temp_bytes, const uint32_t* keys_in,
                      uint32_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out,
                      uint32_t M, int key_bits, hipStream_t s) {
  return rocprim::radix_sort_pairs(temp, *temp_bytes, keys_in, keys_out, vals_in, vals_out,
                                   (size_t)M, 0, key_bits, s);
}

// sdc[s] = first sorted position with key >= s (lower bound), s = 0..S.
// Equals the reference's [0, n0, n0+n1, ...] (gpuNUFFT_operator_factory.cpp:183-205).
__global__ void __launch_bounds__(256) sector_count_kernel(const uint32_t* __restrict__ keys,
                                                           uint32_t M, uint32_t S,
                                                           uint32_t* __restrict__ sdc) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > S) return;
  uint32_t lo = 0, hi = M;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (keys[mid] < s)
      lo = mid + 1;
    else
      hi = mid;
  }
  sdc[s] = lo;
}

hipError_t launch_sector_data_count(const uint32_t* sorted_keys, uint32_t M, uint32_t S,
                                    uint32_t* sdc, hipStream_t s) {
  unsigned blocks = (S + 1 + 255) / 256;
  hipLaunchKernelGGL(sector_count_kernel, dim3(blocks), dim3(256), 0, s, sorted_keys, M, S, sdc);
  return hipGetLastError();
}

// Centres c_a = s_a*sw + floor(sw/2), interleaved per sector
// (gpuNUFFT_operator_factory.cpp:207-270).
__global__ void __launch_bounds__(256) sector_centers_kernel(Geom g, uint32_t* __restrict__ out) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t S = (uint32_t)g.S[0] * g.S[1] * g.S[2];
  if (s >= S) return;
  uint32_t x = s % g.S[0];
  uint32_t y = (s / g.S[0]) % g.S[1];
  uint32_t z = s / ((uint32_t)g.S[0] * g.S[1]);
  uint32_t half = (uint32_t)(g.sw / 2);
  if (g.dim == 2) {
    out[2 * s] = x * g.sw + half;
    out[2 * s + 1] = y * g.sw + half;
  } else {
    out[3 * s] = x * g.sw + half;
    out[3 * s + 1] = y * g.sw + half;
    out[3 * s + 2] = z * g.sw + half;
  }
}

hipError_t launch_sector_centers(const Geom& g, uint32_t* centers, hipStream_t s) {
  uint32_t S = (uint32_t)g.S[0] * g.S[1] * g.S[2];
  hipLaunchKernelGGL(sector_centers_kernel, dim3((S + 255) / 256), dim3(256), 0, s, g, centers);
  return hipGetLastError();
}

// Sorted trajectory / density gather (precomp_kernels.cu:85-113).
__global__ void __launch_bounds__(256) gather_traj_kernel(const float* __restrict__ traj,
                                                          const float* __restrict__ dens,
                                                          const uint32_t* __restrict__ idx,
                                                          uint32_t M, int dim,
                                                          float* __restrict__ traj_sorted,
                                                          float* __restrict__ dens_sorted) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M) return;
  uint32_t j = idx[t];
  for (int a = 0; a < dim; a++) traj_sorted[t + (size_t)a * M] = traj[j + (size_t)a * M];
  if (dens) dens_sorted[t] = dens[j];
}

hipError_t launch_gather_traj(const float* traj, const float* dens, const uint32_t* idx,
                              uint32_t M, int dim, float* traj_sorted, float* dens_sorted,
                              hipStream_t s) {
  if (M == 0) return hipSuccess;
  hipLaunchKernelGGL(gather_traj_kernel, dim3((M + 255) / 256), dim3(256), 0, s, traj, dens, idx,
                     M, dim, traj_sorted, dens_sorted);
  return hipGetLastError();
}

// Number of work chunks per sector, then an exclusive scan: chunk_off has
// S+1 entries, chunk_off[S] = total chunks.
__global__ void __launch_bounds__(256) chunk_count_kernel(const uint32_t* __restrict__ sdc,
                                                          uint32_t S, uint32_t chunk,
                                                          uint32_t* __restrict__ cnt) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > S) return;
  if (s == S) {
    cnt[s] = 0;
    return;
  }
  uint32_t n = sdc[s + 1] - sdc[s];
  cnt[s] = (n + chunk - 1) / chunk;
}

hipError_t count_chunks(const uint32_t* sdc, uint32_t S, uint32_t chunk, uint32_t* cnt,
                        uint32_t* chunk_off, void* temp, size_t* temp_bytes, hipStream_t s) {
  if (temp == nullptr) {
    return rocprim::exclusive_scan(nullptr, *temp_bytes, cnt, chunk_off, 0u, (size_t)S + 1,
                                   rocprim::plus<uint32_t>(), s);
  }
  hipLaunchKernelGGL(chunk_count_kernel, dim3((S + 1 + 255) / 256), dim3(256), 0, s, sdc, S,
                     chunk, cnt);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return rocprim::exclusive_scan(temp, *temp_bytes, cnt, chunk_off, 0u, (size_t)S + 1,
                                 rocprim::plus<uint32_t>(), s);
}

__global__ void __launch_bounds__(256) fill_chunks_kernel(const uint32_t* __restrict__ sdc,
                                                          const uint32_t* __restrict__ chunk_off,
                                                          uint32_t S, uint32_t chunk,
                                                          Chunk* __restrict__ chunks) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  uint32_t b = sdc[s], e = sdc[s + 1];
  uint32_t o = chunk_off[s];
  for (uint32_t p = b; p < e; p += chunk) {
    Chunk c;
    c.sector = s;
    c.begin = p;
    c.end = (e - p > chunk) ? p + chunk : e;
    c.pad = 0;
    chunks[o++] = c;
  }
}

hipError_t launch_fill_chunks(const uint32_t* sdc, const uint32_t* chunk_off, uint32_t S,
                              uint32_t chunk, Chunk* chunks, hipStream_t s) {
  hipLaunchKernelGGL(fill_chunks_kernel, dim3((S + 255) / 256), dim3(256), 0, s, sdc, chunk_off,
                     S, chunk, chunks);
  return hipGetLastError();
}

}  // namespace gnufft
