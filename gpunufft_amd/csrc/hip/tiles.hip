// Adjoint gridding on grid tiles (CDNA4, gfx950).
//
// The reference spreads each sector's samples into a padded sub-grid in
// shared memory and adds the sub-grid into the grid with global atomics
// (CUDA/src/gpu/atomic/atomic_gpuNUFFT_kernels.cu:302-451 in 3-D, :469-617
// in 2-D; texture lookup texture_gpuNUFFT_kernels.cu:22-376). The padded
// border of every sector overlaps its neighbours, so every grid point near
// a sector edge is written by several sub-grids.
//
// Here the grid is cut into tiles (about 16^3 points in 3-D, 64^2 in 2-D)
// and each tile is owned by one workgroup, which accumulates only the tile's
// own points in LDS and writes each of them to the grid exactly once, with
// plain coalesced stores and no zero-fill beforehand. A tile's work is every
// sample whose kernel footprint reaches into it: its own samples and the
// samples of neighbouring tiles within half a kernel width of its faces
// ("visits", listed once per trajectory). A visit applies only the taps that
// land inside the tile, so every (sample, tap) pair is accumulated by
// exactly one workgroup.
//
// Every visit applies exactly the reference's taps: the window is the
// reference's, relative to the sample's own sector (E3 assignment, padded
// sector window with its clamps, CUDA/inc/precomp_utils.hpp:34-48 and
// CUDA/inc/gpuNUFFT_utils.hpp:89-104, 194-199), carried to global grid
// indices; it matters at the top grid edge, where samples clamped into the
// last sector lose their taps past the padded window. The weight of global
// tap g uses the k-space position g / G - 1/2, the same float operations as
// the reference's map_grid_to_kspace (gpuNUFFT_utils.hpp:202-209) with
// i + center - offset == g exactly, so every weight is bit-identical.
//
// Tiles with more visits than one workgroup should take are split into
// parts; each part writes its partial tile to scratch and a reduction sums
// the parts in a fixed order.
//
// Accumulation is in float with LDS atomics (ds_add_f32, two per tap), as in
// the reference; the order of float additions within a tile is therefore
// not fixed (as in the reference's atomics).

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <vector>

#include "gnufft_internal.hpp"

namespace gnufft {

#ifndef GNUFFT_TILE_THREADS
#define GNUFFT_TILE_THREADS 512
#endif
#ifndef GNUFFT_TILE_PER_CU
#define GNUFFT_TILE_PER_CU 2
#endif
constexpr int kTileThreads = GNUFFT_TILE_THREADS;

// Per-axis tile of a grid coordinate from the boundary table:
// tiles [bound[t], bound[t+1]).
__device__ inline int tile_of_coord(const uint32_t* bound, int n, int q) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((int)bound[mid] <= q)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

#pragma clang fp contract(off)

// Global (unwrapped) grid position of a sample along one axis: the first
// two float operations of map_kspace_to_grid (gpuNUFFT_utils.hpp:194-199).
__device__ __host__ inline float grid_pos(float pos, int G) {
  return (pos * (float)G) + (0.5f * (float)G);
}

// k-space position of global grid index g: map_grid_to_kspace with
// i + center - offset == g.
__device__ __host__ inline float kspace_of(int g, int G) { return (float)g / (float)G - 0.5f; }

// The reference's taps of one axis as global (unwrapped) grid indices
// lo .. hi: the sample's sector (E3), its padded window and the clamped tap
// range inside it. The precompute and the kernel use this one definition,
// so they agree on every window.
__device__ inline void ref_taps(const Geom& g, int a, float pos, int* lo, int* hi) {
  const int G = g.G[a];
  const int c = sector_of(pos, G, g.sw) * g.sw + g.sw / 2;
  const float u = map_kspace_to_grid(pos, G, c, g.off);
  int l, h;
  tap_range(u, g.Pmax, g.kradius, &l, &h);
  if (h > l + g.kw) h = l + g.kw;
  *lo = l + c - g.off;
  *hi = h + c - g.off;
}

__device__ inline int home_coord(float A) { return (int)floorf(A + 0.5f); }

#pragma clang fp contract(on)

// ---------------------------------------------------------------------------
// Visit precompute
// ---------------------------------------------------------------------------

// Distinct tiles one axis of a footprint reaches, in tap order (at most 3),
// and the position of the home tile among them.
__device__ inline int axis_tiles(const Geom& g, const TileGrid& t, int a, float pos, int* list,
                                 int* home_pos) {
  const int G = g.G[a];
  const float A = grid_pos(pos, G);
  int lo, hi;
  ref_taps(g, a, pos, &lo, &hi);
  const int home = tile_of_coord(t.bound[a], t.n[a], wrap(home_coord(A), G));
  int n = 0;
  *home_pos = 0;
  for (int q = lo; q <= hi && n < 3; q++) {
    const int tt = tile_of_coord(t.bound[a], t.n[a], wrap(q, G));
    bool seen = false;
    for (int k = 0; k < n; k++) seen = seen || list[k] == tt;
    if (!seen) {
      if (tt == home) *home_pos = n;
      list[n++] = tt;
    }
  }
  if (n == 0) {  // empty window (cannot happen for kw >= 1): visit the home tile
    list[0] = home;
    n = 1;
  }
  return n;
}

__global__ void __launch_bounds__(256) visit_count_kernel(Geom g, TileGrid t,
                                                          const float* __restrict__ crds,
                                                          uint32_t M, uint32_t* __restrict__ cnt) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > M) return;
  if (s == M) {
    cnt[M] = 0;
    return;
  }
  uint32_t c = 1;
  for (int a = 0; a < g.dim; a++) {
    int list[3], hp;
    c *= (uint32_t)axis_tiles(g, t, a, crds[s + (size_t)a * M], list, &hp);
  }
  cnt[s] = c;
}

// key = tile * 27 + code: code orders a tile's visits as own samples first,
// then by the side they come from (similar clipping per wave).
__global__ void __launch_bounds__(256) visit_fill_kernel(Geom g, TileGrid t,
                                                         const float* __restrict__ crds, uint32_t M,
                                                         const uint32_t* __restrict__ off,
                                                         uint32_t* __restrict__ keys,
                                                         uint32_t* __restrict__ vals) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= M) return;
  int list[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, n[3] = {1, 1, 1}, hp[3] = {0, 0, 0};
  for (int a = 0; a < g.dim; a++) n[a] = axis_tiles(g, t, a, crds[s + (size_t)a * M], list[a], &hp[a]);
  uint32_t o = off[s];
  for (int z = 0; z < n[2]; z++)
    for (int y = 0; y < n[1]; y++)
      for (int x = 0; x < n[0]; x++) {
        const uint32_t tile =
            (uint32_t)list[0][x] + (uint32_t)t.n[0] * ((uint32_t)list[1][y] + (uint32_t)t.n[1] * list[2][z]);
        const int cx = x == hp[0] ? 0 : (x < hp[0] ? 1 : 2);
        const int cy = y == hp[1] ? 0 : (y < hp[1] ? 1 : 2);
        const int cz = z == hp[2] ? 0 : (z < hp[2] ? 1 : 2);
        keys[o] = tile * 27u + (uint32_t)(cx + 3 * cy + 9 * cz);
        vals[o] = s;
        o++;
      }
}

__global__ void __launch_bounds__(256) visit_offsets_kernel(const uint32_t* __restrict__ keys,
                                                            uint32_t V, uint32_t n_tiles,
                                                            uint32_t* __restrict__ vbeg) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > n_tiles) return;
  const uint32_t key = t * 27u;
  uint32_t lo = 0, hi = V;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (keys[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  vbeg[t] = lo;
}

__global__ void __launch_bounds__(256) visit_gather_kernel(const uint32_t* __restrict__ pos_of,
                                                           uint32_t V, uint32_t M, int dim,
                                                           const float* __restrict__ crds,
                                                           const uint32_t* __restrict__ idx,
                                                           const float* __restrict__ dens,
                                                           float* __restrict__ vcrd,
                                                           uint32_t* __restrict__ vidx,
                                                           float* __restrict__ vdens) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const uint32_t s = pos_of[v];
  for (int a = 0; a < dim; a++) vcrd[v + (size_t)a * V] = crds[s + (size_t)a * M];
  vidx[v] = idx[s];
  if (dens) vdens[v] = dens[s];
}

hipError_t build_visits(const Geom& g, const TileGrid& t, const float* crds_sorted,
                        const uint32_t* idx, const float* dens_sorted, uint32_t M,
                        TileVisits* out, hipStream_t s) {
  *out = TileVisits();
  uint32_t* cnt = nullptr;
  uint32_t* off = nullptr;
  uint32_t *keys = nullptr, *vals = nullptr, *keys_s = nullptr;
  void* temp = nullptr;
  size_t temp_bytes = 0;
  uint32_t V = 0;
  hipError_t e = hipMalloc(&cnt, (size_t)(M + 1) * 4);
  if (e == hipSuccess) e = hipMalloc(&off, (size_t)(M + 1) * 4);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(visit_count_kernel, dim3((M + 1 + 255) / 256), dim3(256), 0, s, g, t,
                       crds_sorted, M, cnt);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = rocprim::exclusive_scan(nullptr, temp_bytes, cnt, off, 0u, (size_t)M + 1,
                                rocprim::plus<uint32_t>(), s);
  if (e == hipSuccess) e = hipMalloc(&temp, temp_bytes ? temp_bytes : 4);
  if (e == hipSuccess)
    e = rocprim::exclusive_scan(temp, temp_bytes, cnt, off, 0u, (size_t)M + 1,
                                rocprim::plus<uint32_t>(), s);
  if (e == hipSuccess) e = hipMemcpyAsync(&V, off + M, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (temp) (void)hipFree(temp);
  temp = nullptr;
  const size_t Vn = V ? V : 1;
  if (e == hipSuccess) e = hipMalloc(&keys, Vn * 4);
  if (e == hipSuccess) e = hipMalloc(&vals, Vn * 4);
  if (e == hipSuccess) e = hipMalloc(&keys_s, Vn * 4);
  if (e == hipSuccess) e = hipMalloc(&out->pos, Vn * 4);
  if (e == hipSuccess && M) {
    hipLaunchKernelGGL(visit_fill_kernel, dim3((M + 255) / 256), dim3(256), 0, s, g, t,
                       crds_sorted, M, off, keys, vals);
    e = hipGetLastError();
  }
  int bits = 1;
  while ((1ull << bits) < (unsigned long long)t.n_tiles * 27ull) bits++;
  temp_bytes = 0;
  if (e == hipSuccess && V)
    e = rocprim::radix_sort_pairs(nullptr, temp_bytes, keys, keys_s, vals, out->pos, (size_t)V, 0,
                                  bits, s);
  if (e == hipSuccess) e = hipMalloc(&temp, temp_bytes ? temp_bytes : 4);
  if (e == hipSuccess && V)
    e = rocprim::radix_sort_pairs(temp, temp_bytes, keys, keys_s, vals, out->pos, (size_t)V, 0, bits,
                                  s);
  if (e == hipSuccess) e = hipMalloc(&out->vbeg, (size_t)(t.n_tiles + 1) * 4);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(visit_offsets_kernel, dim3((t.n_tiles + 1 + 255) / 256), dim3(256), 0, s,
                       keys_s, V, t.n_tiles, out->vbeg);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMalloc(&out->crd, Vn * 4 * g.dim);
  if (e == hipSuccess) e = hipMalloc(&out->idx, Vn * 4);
  if (e == hipSuccess && dens_sorted) e = hipMalloc(&out->dens, Vn * 4);
  if (e == hipSuccess && V) {
    hipLaunchKernelGGL(visit_gather_kernel, dim3((V + 255) / 256), dim3(256), 0, s, out->pos, V, M,
                       g.dim, crds_sorted, idx, dens_sorted, out->crd, out->idx, out->dens);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  out->V = V;
  if (cnt) (void)hipFree(cnt);
  if (off) (void)hipFree(off);
  if (keys) (void)hipFree(keys);
  if (vals) (void)hipFree(vals);
  if (keys_s) (void)hipFree(keys_s);
  if (temp) (void)hipFree(temp);
  if (e != hipSuccess) free_visits(out);
  return e;
}

void free_visits(TileVisits* v) {
  if (v->pos) (void)hipFree(v->pos);
  if (v->vbeg) (void)hipFree(v->vbeg);
  if (v->crd) (void)hipFree(v->crd);
  if (v->idx) (void)hipFree(v->idx);
  if (v->dens) (void)hipFree(v->dens);
  *v = TileVisits();
}

// Tile partition of one axis of G points into n tiles whose extents differ
// by at most one (no sliver tiles): tile a = [floor(a*G/n), floor((a+1)*G/n)).
void tile_axis(int G, int target, int* n, int* Tm, std::vector<uint32_t>* bound) {
  int k = (G + target - 1) / target;
  if (k < 1) k = 1;
  *n = k;
  bound->resize((size_t)k + 1);
  int mx = 0;
  for (int a = 0; a <= k; a++) (*bound)[(size_t)a] = (uint32_t)(((long long)a * G) / k);
  for (int a = 0; a < k; a++) mx = std::max(mx, (int)((*bound)[(size_t)a + 1] - (*bound)[(size_t)a]));
  *Tm = mx;
}

// ---------------------------------------------------------------------------
// Spread
// ---------------------------------------------------------------------------

// Visit record in LDS: word 0 packs the in-tile tap count per axis (4 bits
// each); words 1 + a*NX + t hold the weight of the t-th in-tile tap on axis
// a (zero past the count); then per axis NX/4 words of tile-local tap
// indices, one byte per tap (past the count: the first tap's index, so a
// zero-weight tap still addresses a valid word).
template <int NX>
struct TRec {
  static constexpr int IW = (NX + 3) / 4;
  static constexpr int W = ((1 + 3 * NX + 3 * IW) + 3) & ~3;
};

struct TileLayout {
  size_t acc_off, lut_off, rec_off, sv_off, bytes;
  __host__ __device__ TileLayout(int np, int K, int nx, int cb, int R) {
    const int iw = (nx + 3) / 4;
    const int rw = ((1 + 3 * nx + 3 * iw) + 3) & ~3;
    acc_off = 0;
    lut_off = ((size_t)np * cb * 8 + 15) & ~(size_t)15;  // whole float4s (zeroed as such)
    rec_off = lut_off + (size_t)((K + 3) & ~3) * 4;
    sv_off = rec_off + (size_t)R * rw * 4;
    bytes = sv_off + (size_t)R * cb * 8;
  }
};

__device__ inline int wave_max(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

template <int DIM, int NX>
__device__ inline void make_tile_record(const Geom& g, const float* lut, const int* o, const int* E,
                                        const float* p, float* rec) {
  constexpr int IW = TRec<NX>::IW;
  constexpr int RW = TRec<NX>::W;
  float w[RW];
#pragma unroll
  for (int q = 0; q < RW; q++) w[q] = 0.0f;
  uint32_t counts = 0;
#pragma unroll
  for (int a = 0; a < DIM; a++) {
    const int G = g.G[a];
    int lo, hi;
    ref_taps(g, a, p[a], &lo, &hi);
    uint32_t idx[IW];
#pragma unroll
    for (int k = 0; k < IW; k++) idx[k] = 0;
    int n = 0, first = 0;
#pragma unroll
    for (int t = 0; t < NX; t++) {
      const int q = lo + t;
      if (q > hi) break;
      const int l = wrap(q, G) - o[a];
      if (l < 0 || l >= E[a]) continue;
      float wt = 0.0f;
      if (!axis_weight_kp(g, lut, a, kspace_of(q, G), p[a], &wt)) wt = 0.0f;
      if (n == 0) first = l;
      w[1 + a * NX + n] = wt;
      idx[n >> 2] |= (uint32_t)l << (8 * (n & 3));
      n++;
    }
    for (int t = n; t < NX; t++) idx[t >> 2] |= (uint32_t)first << (8 * (t & 3));
    counts |= (uint32_t)n << (4 * a);
#pragma unroll
    for (int k = 0; k < IW; k++) w[1 + 3 * NX + a * IW + k] = __uint_as_float(idx[k]);
  }
  if (DIM == 2) {
    counts |= 1u << 8;  // one z tap at local 0, weight 1
    w[1 + 2 * NX] = 1.0f;
  }
  w[0] = __uint_as_float(counts);
  float4* r4 = reinterpret_cast<float4*>(rec);
#pragma unroll
  for (int q = 0; q < RW / 4; q++) r4[q] = make_float4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
}

template <int NX>
__device__ inline int rec_index(const float* w, int a, int t) {
  constexpr int IW = TRec<NX>::IW;
  const uint32_t word = __float_as_uint(w[1 + 3 * NX + a * IW + (t >> 2)]);
  return (int)((word >> (8 * (t & 3))) & 255u);
}

// Tile geometry of item `it`: origin and extent per axis.
template <int DIM>
__device__ inline void tile_box(const TileArgs& a, uint32_t tile, int* o, int* E) {
  uint32_t r = tile;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (d < DIM) {
      const uint32_t n = (uint32_t)a.tiles.n[d];
      const uint32_t td = r % n;
      r /= n;
      o[d] = (int)a.tiles.bound[d][td];
      E[d] = (int)(a.tiles.bound[d][td + 1] - a.tiles.bound[d][td]);
    } else {
      o[d] = 0;
      E[d] = 1;
    }
  }
}

template <int DIM, int NX, int CB, int NT>
__global__ void __launch_bounds__(NT) tile_spread_kernel(Geom g, TileArgs a, int R) {
  constexpr int NW = NT / 64;
  constexpr int SPW = 64 / CB;
  constexpr int RW = TRec<NX>::W;
  const uint32_t x = blockIdx.x & 7u;
  const uint32_t ii = a.xcd_begin[x] + (blockIdx.x >> 3);
  if (ii >= a.xcd_begin[x + 1]) return;
  const uint4 it = a.items[ii];  // (tile, visit begin, visit end, partial slot or ~0)
  const int coil0 = blockIdx.y * CB;

  extern __shared__ float4 smem4[];
  char* smem = reinterpret_cast<char*>(smem4);
  const int Tm0 = a.tiles.Tm[0], Tm1 = a.tiles.Tm[1], Tm2 = DIM == 3 ? a.tiles.Tm[2] : 1;
  const int NP = Tm0 * Tm1 * Tm2;
  const TileLayout L(NP, g.K, NX, CB, R);
  float* acc = reinterpret_cast<float*>(smem + L.acc_off);
  float* lut = reinterpret_cast<float*>(smem + L.lut_off);
  float* rec = reinterpret_cast<float*>(smem + L.rec_off);
  float2* sv = reinterpret_cast<float2*>(smem + L.sv_off);

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  int o[3], E[3];
  tile_box<DIM>(a, it.x, o, E);

  for (int i = tid; i < g.K; i += NT) lut[i] = a.lut[i];
  {
    float4* a4 = reinterpret_cast<float4*>(acc);
    const int n4 = (NP * CB * 2 + 3) / 4;
    for (int i = tid; i < n4; i += NT) a4[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }

  const int slot = lane / CB;
  const int c = lane % CB;
  const bool coil_ok = coil0 + c < a.n_coils;
  const int n = (int)(it.z - it.y);
  for (int r0 = 0; r0 < n; r0 += R) {
    const int rn = min(R, n - r0);
    __syncthreads();  // tables zeroed / previous round consumed
    for (int i = tid; i < rn; i += NT) {
      const uint32_t v = it.y + r0 + i;
      float p[3];
      p[0] = a.vcrd[v];
      p[1] = a.vcrd[(size_t)a.V + v];
      p[2] = DIM == 3 ? a.vcrd[2 * (size_t)a.V + v] : 0.0f;
      const uint32_t src = a.vsrc[v];
      const float sq = a.vdens ? sqrtf(a.vdens[v]) : 1.0f;
#pragma unroll
      for (int cc = 0; cc < CB; cc++) {
        float2 val = make_float2(0.0f, 0.0f);
        if (coil0 + cc < a.n_coils) {
          val = a.data[(size_t)(coil0 + cc) * a.coil_stride + src];
          val.x *= sq;
          val.y *= sq;
        }
        sv[i * CB + cc] = val;
      }
      make_tile_record<DIM, NX>(g, lut, o, E, p, rec + i * RW);
    }
    __syncthreads();

    for (int b = wave * SPW; b < rn; b += NW * SPW) {
      const int ls = b + slot;
      const bool act = ls < rn && coil_ok;
      float w[RW];
      if (act) {
        const float4* r4 = reinterpret_cast<const float4*>(rec + ls * RW);
#pragma unroll
        for (int q = 0; q < RW / 4; q++) {
          const float4 t4 = r4[q];
          w[4 * q] = t4.x;
          w[4 * q + 1] = t4.y;
          w[4 * q + 2] = t4.z;
          w[4 * q + 3] = t4.w;
        }
      } else {
#pragma unroll
        for (int q = 0; q < RW; q++) w[q] = 0.0f;
      }
      const uint32_t counts = __float_as_uint(w[0]);
      const int nx = wave_max((int)(counts & 15u));
      const int ny = wave_max((int)((counts >> 4) & 15u));
      const int nz = wave_max((int)((counts >> 8) & 15u));
      if (!act) continue;
      const float2 val = sv[ls * CB + c];
      // taps in (z, y, x) order, weight (wz*wy)*wx as the reference forms it
#pragma unroll
      for (int tz = 0; tz < (DIM == 3 ? NX : 1); tz++) {
        if (tz >= nz) break;
        const int iz = DIM == 3 ? rec_index<NX>(w, 2, tz) : 0;
        const float wz = DIM == 3 ? w[1 + 2 * NX + tz] : 1.0f;
#pragma unroll
        for (int ty = 0; ty < NX; ty++) {
          if (ty >= ny) break;
          const int iy = rec_index<NX>(w, 1, ty);
          const float wzy = DIM == 3 ? wz * w[1 + NX + ty] : w[1 + NX + ty];
          const int row = (iz * Tm1 + iy) * Tm0;
#pragma unroll
          for (int tx = 0; tx < NX; tx++) {
            if (tx >= nx) break;
            const int ix = rec_index<NX>(w, 0, tx);
            const float wt = wzy * w[1 + tx];
            float* dst = acc + 2 * ((row + ix) * CB + c);
            atomicAdd(dst, wt * val.x);
            atomicAdd(dst + 1, wt * val.y);
          }
        }
      }
    }
  }
  __syncthreads();

  // Store: every point of the tile exactly once, the coils of a point
  // consecutive in LDS; with a partial slot, to scratch [slot][point][coil].
  const int Etot = E[0] * E[1] * E[2];
  const float2* acc2 = reinterpret_cast<const float2*>(acc);
  for (int e = tid; e < Etot * CB; e += NT) {
    const int cc = e % CB;
    const int pe = e / CB;
    if (coil0 + cc >= a.n_coils) continue;
    const int lx = pe % E[0];
    const int lyz = pe / E[0];
    const int ly = lyz % E[1];
    const int lz = lyz / E[1];
    const int pt = (lz * Tm1 + ly) * Tm0 + lx;
    const float2 v = acc2[pt * CB + cc];
    if (it.w != ~0u) {
      a.partial[((size_t)it.w * NP + pt) * a.n_coils + coil0 + cc] = v;
    } else {
      int q0 = o[0] + lx, q1 = o[1] + ly, q2 = o[2] + lz;
      if (a.fold_shift) {
        q0 = wrap(q0 - g.shift_inv[0], g.G[0]);
        q1 = wrap(q1 - g.shift_inv[1], g.G[1]);
        if (DIM == 3) q2 = wrap(q2 - g.shift_inv[2], g.G[2]);
      }
      const size_t gi = (size_t)q0 + (size_t)g.G[0] * ((size_t)q1 + (size_t)g.G[1] * (size_t)q2);
      a.grid[(size_t)(coil0 + cc) * g.Ng + gi] = v;
    }
  }
}

// Split tiles: task (tile, first slot, parts) sums its parts in slot order
// and stores the tile. blockIdx.y: block of points.
template <int DIM>
__global__ void __launch_bounds__(256) tile_reduce_kernel(Geom g, TileArgs a) {
  const uint4 task = a.tasks[blockIdx.x];
  int o[3], E[3];
  tile_box<DIM>(a, task.x, o, E);
  const int Tm0 = a.tiles.Tm[0], Tm1 = a.tiles.Tm[1], Tm2 = DIM == 3 ? a.tiles.Tm[2] : 1;
  const size_t NP = (size_t)Tm0 * Tm1 * Tm2;
  const int Etot = E[0] * E[1] * E[2];
  const int C = a.n_coils;
  for (int e = blockIdx.y * 256 + threadIdx.x; e < Etot * C; e += gridDim.y * 256) {
    const int cc = e % C;
    const int pe = e / C;
    const int lx = pe % E[0];
    const int lyz = pe / E[0];
    const int ly = lyz % E[1];
    const int lz = lyz / E[1];
    const size_t pt = (size_t)(lz * Tm1 + ly) * Tm0 + lx;
    float2 s = make_float2(0.0f, 0.0f);
    for (uint32_t k = 0; k < task.z; k++) {
      const float2 v = a.partial[((size_t)(task.y + k) * NP + pt) * C + cc];
      s.x += v.x;
      s.y += v.y;
    }
    int q0 = o[0] + lx, q1 = o[1] + ly, q2 = o[2] + lz;
    if (a.fold_shift) {
      q0 = wrap(q0 - g.shift_inv[0], g.G[0]);
      q1 = wrap(q1 - g.shift_inv[1], g.G[1]);
      if (DIM == 3) q2 = wrap(q2 - g.shift_inv[2], g.G[2]);
    }
    const size_t gi = (size_t)q0 + (size_t)g.G[0] * ((size_t)q1 + (size_t)g.G[1] * (size_t)q2);
    a.grid[(size_t)cc * g.Ng + gi] = s;
  }
}

static size_t tile_lds_limit() {
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

template <int DIM, int NX, int CB, int NT>
static hipError_t tile_launch(const Geom& g, const TileArgs& a, int R, hipStream_t s) {
  const int NP = a.tiles.Tm[0] * a.tiles.Tm[1] * (DIM == 3 ? a.tiles.Tm[2] : 1);
  const size_t lds = TileLayout(NP, g.K, NX, CB, R).bytes;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(tile_spread_kernel<DIM, NX, CB, NT>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  const uint32_t blocks = 8u * (a.max_range ? a.max_range : 1u);
  const int coil_blocks = (a.n_coils + CB - 1) / CB;
  hipLaunchKernelGGL((tile_spread_kernel<DIM, NX, CB, NT>), dim3(blocks, coil_blocks), dim3(NT), lds,
                     s, g, a, R);
  return hipGetLastError();
}

// Coil block and round size: the largest coil block (not above the coil
// count) whose tile accumulator plus a round of at least 64 visits fits
// `per_cu` workgroups in a CU's LDS.
static int tile_pick(const Geom& g, const TileArgs& a, int nx, int per_cu, int* R) {
  const int NP = a.tiles.Tm[0] * a.tiles.Tm[1] * (g.dim == 3 ? a.tiles.Tm[2] : 1);
  const size_t limit = tile_lds_limit() / (size_t)per_cu - 512;
  for (int cb = 8; cb >= 1; cb /= 2) {
    if (cb > 1 && cb / 2 >= a.n_coils) continue;
    const size_t fixed = TileLayout(NP, g.K, nx, cb, 0).bytes;
    const size_t per = TileLayout(NP, g.K, nx, cb, 1).bytes - fixed;
    if (fixed + 64 * per > limit) continue;
    size_t r = (limit - fixed) / per;
    r = std::min<size_t>(r, kTileThreads);
    r &= ~(size_t)63;
    *R = (int)r;
    return cb;
  }
  return 0;
}

// Workgroups per CU the LDS footprint is sized for.
int tile_spread_blocks_per_cu(const Geom&) { return GNUFFT_TILE_PER_CU; }

hipError_t launch_tile_spread(const Geom& g, const TileArgs& a, hipStream_t s) {
  if (a.n_coils == 0 || a.n_items == 0) return hipSuccess;
  const int nx = g.kw + 1 <= 4 ? 4 : 8;
  int R = 0;
  int cb = tile_pick(g, a, nx, tile_spread_blocks_per_cu(g), &R);
  if (cb == 0) cb = tile_pick(g, a, nx, 1, &R);
  hipError_t e = hipErrorInvalidValue;
#define TILE_CASE(D, X, C)                                 \
  if (g.dim == D && nx == X && cb == C) e = tile_launch<D, X, C, kTileThreads>(g, a, R, s);
  TILE_CASE(3, 4, 8) TILE_CASE(3, 4, 4) TILE_CASE(3, 4, 2) TILE_CASE(3, 4, 1)
  TILE_CASE(3, 8, 8) TILE_CASE(3, 8, 4) TILE_CASE(3, 8, 2) TILE_CASE(3, 8, 1)
  TILE_CASE(2, 4, 8) TILE_CASE(2, 4, 4) TILE_CASE(2, 4, 2) TILE_CASE(2, 4, 1)
  TILE_CASE(2, 8, 8) TILE_CASE(2, 8, 4) TILE_CASE(2, 8, 2) TILE_CASE(2, 8, 1)
#undef TILE_CASE
  if (e != hipSuccess) return e;
  if (a.n_tasks) {
    const int NP = a.tiles.Tm[0] * a.tiles.Tm[1] * (g.dim == 3 ? a.tiles.Tm[2] : 1);
    const int yb = std::max(1, std::min(64, (NP * a.n_coils + 255) / 256));
    if (g.dim == 3)
      hipLaunchKernelGGL(tile_reduce_kernel<3>, dim3(a.n_tasks, yb), dim3(256), 0, s, g, a);
    else
      hipLaunchKernelGGL(tile_reduce_kernel<2>, dim3(a.n_tasks, yb), dim3(256), 0, s, g, a);
    e = hipGetLastError();
  }
  return e;
}

}  // namespace gnufft
