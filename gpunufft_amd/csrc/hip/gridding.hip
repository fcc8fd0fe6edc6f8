// Kaiser-Bessel gridding kernels for CDNA4 (gfx950).
//
// spread (adjoint convolution, reference CUDA/src/gpu/atomic/
//   atomic_gpuNUFFT_kernels.cu:302-617 and texture_gpuNUFFT_kernels.cu:22-376)
// interp (forward convolution, atomic_gpuNUFFT_kernels.cu:966-1400,
//   texture_gpuNUFFT_kernels.cu:467-1039)
//
// Work decomposition: one workgroup per (sector chunk, coil block). A chunk
// is a contiguous run of at most `chunk` sorted samples of one sector, so
// heavy sectors near the k-space centre spread over many workgroups while
// ordinary sectors are one chunk each (one sub-grid flush per sector).
// Workgroup ids are remapped so that each of the 8 XCDs owns a contiguous
// range of chunks (neighbouring sectors share that XCD's L2).
//
// A chunk is processed in rounds of R samples (R sized to the LDS left over
// by the sub-grid). Each round first computes, once per sample, a tap record
// in LDS: the first tap index along each axis, whether the sample has kw+1
// taps along it, and the per-axis kernel weights (zero where the reference
// skips a tap). The tap loop then runs with lanes = (sample slot, coil);
// its trip counts are wave-uniform (kw, or kw+1 when any lane of the wave
// needs it), so no lane idles through masked iterations.
//
// The spread accumulates the padded sector sub-grid in LDS in double
// precision, two planes (re, im) laid out [point][coil]: on gfx950
// ds_add_f64 sustains ~20x the rate of ds_add_f32, and double sums are at
// least as accurate as the reference's float atomics. The slots of one wave
// take samples that lie far apart in the round, so their taps rarely hit
// the same LDS words.

#include "gnufft_internal.hpp"

namespace gnufft {

constexpr int kThreads = 512;

// Chunk of workgroup bid: the (bid / 8)-th chunk of XCD (bid % 8)'s range,
// or ~0u past the end of that range.
__device__ inline uint32_t work_chunk(const WorkList& w, uint32_t bid) {
  const uint32_t x = bid & 7u;
  const uint32_t ci = w.xcd_begin[x] + (bid >> 3);
  return ci < w.xcd_begin[x + 1] ? ci : ~0u;
}

// Workgroups along x for a work list.
static inline uint32_t work_blocks(const WorkList& w) { return 8u * (w.max_range ? w.max_range : 1u); }

void partition_work(const uint32_t* counts, uint32_t n, WorkList* w) {
  // Cost of a chunk: its samples plus a fixed share for staging and
  // flushing the padded sub-grid (about 300 samples' worth on gfx950).
  constexpr double kChunkCost = 300.0;
  double total = 0.0;
  for (uint32_t i = 0; i < n; i++) total += (double)counts[i] + kChunkCost;
  uint32_t x = 0;
  double acc = 0.0;
  w->xcd_begin[0] = 0;
  for (uint32_t i = 0; i < n && x < 7; i++) {
    acc += (double)counts[i] + kChunkCost;
    while (x < 7 && acc >= total * (double)(x + 1) / 8.0) w->xcd_begin[++x] = i + 1;
  }
  while (x < 8) w->xcd_begin[++x] = n;
  w->n_chunks = n;
  w->max_range = 0;
  for (int k = 0; k < 8; k++) {
    const uint32_t len = w->xcd_begin[k + 1] - w->xcd_begin[k];
    if (len > w->max_range) w->max_range = len;
  }
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

// k-space positions of the padded sector's indices, per axis:
// kp[a * KPS + i] == map_grid_to_kspace(i, G[a], center[a], off).
template <int DIM, int NT>
__device__ inline void fill_kpos(const Geom& g, const int* c3, float* kp, int KPS) {
  for (int e = threadIdx.x; e < DIM * KPS; e += NT) {
    const int d = e / KPS;
    const int i = e - d * KPS;
    kp[e] = map_grid_to_kspace(i, g.G[d], c3[d], g.off);
  }
}

// ---------------------------------------------------------------------------
// Tap records. Word 0 packs the first tap index per axis (bits 8a..8a+7) and
// a flag per axis (bit 24+a) set when the sample has kw+1 taps along it;
// word 1 + a*NX + t holds the weight of tap first+t on axis a, zero for taps
// the reference does not apply (beyond the sample's range, or failing the
// nearest-mode radius test, atomic_gpuNUFFT_kernels.cu:343-368).
// ---------------------------------------------------------------------------
template <int NX>
struct Rec {
  static constexpr int W = ((1 + 3 * NX) + 3) & ~3;  // words, float4 aligned
};

template <int DIM, int NX>
__device__ inline void make_record(const Geom& g, const float* lut, const float* kp, int KPS,
                                   const int* c3, const float* p, float* rec) {
  constexpr int RW = Rec<NX>::W;
  float w[RW];
#pragma unroll
  for (int q = 0; q < RW; q++) w[q] = 0.0f;
  uint32_t packed = 0;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const float u = map_kspace_to_grid(p[d], g.G[d], c3[d], g.off);
    int lo, hi;
    tap_range(u, g.Pmax, g.kradius, &lo, &hi);
    packed |= (uint32_t)lo << (8 * d);
    if (hi - lo + 1 > g.kw) packed |= 1u << (24 + d);
#pragma unroll
    for (int t = 0; t < NX; t++) {
      float wt = 0.0f;
      if (lo + t <= hi && axis_weight_kp(g, lut, d, kp[d * KPS + lo + t], p[d], &wt))
        w[1 + d * NX + t] = wt;
    }
  }
  w[0] = __uint_as_float(packed);
  float4* r4 = reinterpret_cast<float4*>(rec);
#pragma unroll
  for (int q = 0; q < RW / 4; q++) r4[q] = make_float4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
}

template <int NX>
__device__ inline void load_record(const float* rec, float* w) {
  const float4* r4 = reinterpret_cast<const float4*>(rec);
#pragma unroll
  for (int q = 0; q < Rec<NX>::W / 4; q++) {
    const float4 v = r4[q];
    w[4 * q] = v.x;
    w[4 * q + 1] = v.y;
    w[4 * q + 2] = v.z;
    w[4 * q + 3] = v.w;
  }
}

// Wave-uniform number of taps along axis a: kw, or kw+1 when any active
// lane's sample has kw+1 taps there.
__device__ inline int wave_trip(uint32_t packed, int a, int kw) {
  return kw + (__ballot((packed >> (24 + a)) & 1u) != 0ull ? 1 : 0);
}

// Sample coordinate of sorted sample s (SoA [DIM][M]).
template <int DIM>
__device__ inline void load_coord(const float* crds, int M, uint32_t s, float* p) {
  p[0] = crds[s];
  p[1] = crds[(size_t)M + s];
  p[2] = DIM == 3 ? crds[2 * (size_t)M + s] : 0.0f;
}

// Shared-memory layout of the spread (offsets in bytes from the dynamic
// LDS base; identical on host and device).
struct SpreadLayout {
  int kpad, kpw, plane, sub_n;
  size_t kp_off, sub_off, rec_off, sv_off, bytes;
  __host__ __device__ SpreadLayout(const Geom& g, int nx, int cb, int R) {
    const int P = g.P;
    const int PD = g.dim == 3 ? P * P * P : P * P;
    const int rw = ((1 + 3 * nx) + 3) & ~3;
    kpad = (g.K + 3) & ~3;
    kpw = (3 * (P + nx) + 3) & ~3;
    plane = PD * cb;
    sub_n = 2 * plane + nx * cb;  // + slack for the last row's window overrun
    kp_off = (size_t)kpad * 4;
    sub_off = kp_off + (size_t)kpw * 4;
    rec_off = sub_off + (size_t)sub_n * 8;
    sv_off = rec_off + (size_t)R * rw * 4;
    bytes = sv_off + (size_t)R * cb * 8;
  }
};

struct InterpLayout {
  int kpad, kpw, sub_n;
  size_t kp_off, sub_off, rec_off, bytes;
  __host__ __device__ InterpLayout(const Geom& g, int nx, int cb, int R) {
    const int P = g.P;
    const int PD = g.dim == 3 ? P * P * P : P * P;
    const int rw = ((1 + 3 * nx) + 3) & ~3;
    kpad = (g.K + 3) & ~3;
    kpw = (3 * (P + nx) + 3) & ~3;
    sub_n = PD * cb + nx * cb;  // + slack for the last row's window overrun
    kp_off = (size_t)kpad * 4;
    sub_off = kp_off + (size_t)kpw * 4;
    rec_off = sub_off + (size_t)((sub_n + 1) & ~1) * 8;
    bytes = rec_off + (size_t)R * rw * 4;
  }
};

// ---------------------------------------------------------------------------
// Spread
// ---------------------------------------------------------------------------
template <int DIM, int NX, int CB>
__global__ void __launch_bounds__(kThreads) spread_kernel(Geom g, SpreadArgs a, int R) {
  constexpr int NT = kThreads;
  constexpr int NW = NT / 64;
  constexpr int SPW = 64 / CB;  // sample slots per wave; one lane per (slot, coil)
  constexpr int RW = Rec<NX>::W;

  extern __shared__ float4 smem4[];
  char* smem = reinterpret_cast<char*>(smem4);
  const SpreadLayout L(g, NX, CB, R);
  float* lut = reinterpret_cast<float*>(smem);
  float* kp = reinterpret_cast<float*>(smem + L.kp_off);
  double* sub = reinterpret_cast<double*>(smem + L.sub_off);
  float* rec = reinterpret_cast<float*>(smem + L.rec_off);
  float2* sv = reinterpret_cast<float2*>(smem + L.sv_off);
  const int P = g.P;
  const int Pmax = g.Pmax;
  const int KPS = P + NX;
  const int PD = DIM == 3 ? P * P * P : P * P;

  const uint32_t ci = work_chunk(a.work, blockIdx.x);
  if (ci == ~0u) return;
  const Chunk ch = a.work.chunks[ci];
  const int n = (int)(ch.end - ch.begin);
  const int coil0 = blockIdx.y * CB;
  const int tid = threadIdx.x;
  int c3[3];
  load_center<DIM>(a.centers, ch.sector, c3);

  for (int i = tid; i < g.K; i += NT) lut[i] = a.lut[i];
  fill_kpos<DIM, NT>(g, c3, kp, KPS);
  {
    float4* z4 = reinterpret_cast<float4*>(sub);
    for (int i = tid; i < L.sub_n / 2; i += NT) z4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }

  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int slot = lane / CB;
  const int c = lane % CB;
  const bool coil_ok = coil0 + c < a.n_coils;

  for (int r0 = 0; r0 < n; r0 += R) {
    const int rn = min(R, n - r0);
    __syncthreads();  // tables ready / previous round consumed
    for (int i = tid; i < rn; i += NT) {
      const uint32_t s = ch.begin + r0 + i;
      float p[3];
      load_coord<DIM>(a.crds, g.M, s, p);
      const uint32_t src = a.gather ? a.gather[s] : s;
      const float sq = a.dens ? sqrtf(a.dens[s]) : 1.0f;
#pragma unroll
      for (int cc = 0; cc < CB; cc++) {
        float2 v = make_float2(0.0f, 0.0f);
        if (coil0 + cc < a.n_coils) {
          v = a.data[(size_t)(coil0 + cc) * a.coil_stride + src];
          if (a.dens) {
            v.x = v.x * sq;
            v.y = v.y * sq;
          }
        }
        sv[i * CB + cc] = v;
      }
      make_record<DIM, NX>(g, lut, kp, KPS, c3, p, rec + i * RW);
    }
    __syncthreads();

    // Slot k of a wave takes samples [k*seg, (k+1)*seg) of the round.
    const int seg = (rn + SPW - 1) / SPW;
    for (int j = wave; j < seg; j += NW) {
      const int ls = slot * seg + j;
      if (ls >= rn || !coil_ok) continue;
      float w[RW];
      load_record<NX>(rec + ls * RW, w);
      const uint32_t packed = __float_as_uint(w[0]);
      const int bx = (int)(packed & 255u);
      const int by = (int)((packed >> 8) & 255u);
      const int bz = (int)((packed >> 16) & 255u);
      const int nx = wave_trip(packed, 0, g.kw);
      const int ny = wave_trip(packed, 1, g.kw);
      const int nz = DIM == 3 ? wave_trip(packed, 2, g.kw) : 1;
      const float2 v = sv[ls * CB + c];
      const double vx = (double)v.x, vy = (double)v.y;
      // taps in (z, y, x) order, weight (wz*wy)*wx as the reference forms it
#pragma unroll
      for (int tz = 0; tz < (DIM == 3 ? NX : 1); tz++) {
        if (tz >= nz) break;
        const int iz = DIM == 3 ? min(bz + tz, Pmax) : 0;
        const float wz = DIM == 3 ? w[1 + 2 * NX + tz] : 1.0f;
#pragma unroll
        for (int ty = 0; ty < NX; ty++) {
          if (ty >= ny) break;
          const int iy = min(by + ty, Pmax);
          const float wzy = DIM == 3 ? wz * w[1 + NX + ty] : w[1 + NX + ty];
          double* re = sub + ((iz * P + iy) * P + bx) * CB + c;
          double* im = re + L.plane;
#pragma unroll
          for (int tx = 0; tx < NX; tx++) {
            if (tx >= nx) break;
            const double wt = (double)(wzy * w[1 + tx]);
            atomicAdd(re + tx * CB, wt * vx);
            atomicAdd(im + tx * CB, wt * vy);
          }
        }
      }
    }
  }
  __syncthreads();

  // Store the chunk's sub-grid, [coil][padded point], with plain coalesced
  // stores; fold_kernel sums the sub-grids that overlap on the grid.
  float2* dst = a.scratch + ((size_t)ci * a.n_coils + coil0) * PD;
  for (int e = tid; e < PD * CB; e += NT) {
    const int pt = e / CB;
    const int cc = e % CB;
    if (coil0 + cc >= a.n_coils) continue;
    dst[(size_t)cc * PD + pt] = make_float2((float)sub[e], (float)sub[L.plane + e]);
  }
}

// ---------------------------------------------------------------------------
// Fold: grid point q of coil c = sum over the padded sectors covering q and
// over each such sector's chunks of the sub-grid value at q's padded index.
// One workgroup per (home sector, coil); the home sector of q is q / sw per
// axis, so every grid point is written exactly once, in a fixed summation
// order (deterministic). Periodic wrap and the optional INVERSE fftshift of
// the store (std_gpuNUFFT_kernels.cu:629-654) are in the cover lists and
// the store index.
// ---------------------------------------------------------------------------
template <int DIM>
__global__ void __launch_bounds__(256) fold_kernel(Geom g, SpreadArgs a) {
  const int P = g.P;
  const int PD = DIM == 3 ? P * P * P : P * P;
  const int sw = g.sw;
  const uint32_t s = blockIdx.x;
  const int c = blockIdx.y;
  const int sx = (int)(s % (uint32_t)g.S[0]);
  const uint32_t r = s / (uint32_t)g.S[0];
  const int sy = (int)(r % (uint32_t)g.S[1]);
  const int sz = DIM == 3 ? (int)(r / (uint32_t)g.S[1]) : 0;
  const int x0 = sx * sw, y0 = sy * sw, z0 = sz * sw;
  const int hx = min(sw, g.G[0] - x0);
  const int hy = min(sw, g.G[1] - y0);
  const int hz = DIM == 3 ? min(sw, g.G[2] - z0) : 1;
  const int HP = hx * hy * hz;
  const float2* scr = a.scratch + (size_t)c * PD;
  const size_t cstride = (size_t)a.n_coils * PD;
  for (int e = threadIdx.x; e < HP; e += blockDim.x) {
    const int lx = e % hx;
    const int lyz = e / hx;
    const int ly = lyz % hy;
    const int lz = lyz / hy;
    const int q0 = x0 + lx, q1 = y0 + ly, q2 = z0 + lz;
    float accx = 0.0f, accy = 0.0f;
    const uint32_t bz = DIM == 3 ? a.cover.off[2][q2] : 0, ez = DIM == 3 ? a.cover.off[2][q2 + 1] : 1;
    const uint32_t by = a.cover.off[1][q1], ey = a.cover.off[1][q1 + 1];
    const uint32_t bx = a.cover.off[0][q0], ex = a.cover.off[0][q0 + 1];
    for (uint32_t kz = bz; kz < ez; kz++) {
      const uint32_t vz = DIM == 3 ? a.cover.ent[2][kz] : 0u;
      for (uint32_t ky = by; ky < ey; ky++) {
        const uint32_t vy = a.cover.ent[1][ky];
        const uint32_t nyz = (vy >> 8) + (uint32_t)g.S[1] * (vz >> 8);
        const int iyz = (int)(vy & 255u) + P * (int)(vz & 255u);
        for (uint32_t kx = bx; kx < ex; kx++) {
          const uint32_t vx = a.cover.ent[0][kx];
          const uint32_t n = (vx >> 8) + (uint32_t)g.S[0] * nyz;
          const int pt = (int)(vx & 255u) + P * iyz;
          const uint32_t k1 = a.sector_chunks[n + 1];
          for (uint32_t k = a.sector_chunks[n]; k < k1; k++) {
            const float2 v = scr[(size_t)k * cstride + pt];
            accx += v.x;
            accy += v.y;
          }
        }
      }
    }
    int o0 = q0, o1 = q1, o2 = q2;
    if (a.fold_shift) {
      o0 = wrap(o0 - g.shift_inv[0], g.G[0]);
      o1 = wrap(o1 - g.shift_inv[1], g.G[1]);
      if (DIM == 3) o2 = wrap(o2 - g.shift_inv[2], g.G[2]);
    }
    a.grid[(size_t)c * g.Ng + (size_t)o0 + (size_t)g.G[0] * ((size_t)o1 + (size_t)g.G[1] * o2)] =
        make_float2(accx, accy);
  }
}

void build_cover(const Geom& g, std::vector<uint32_t>* words, size_t base[6]) {
  std::vector<uint32_t>& w = *words;
  w.clear();
  std::vector<std::vector<uint32_t>> lists[3];
  for (int a = 0; a < 3; a++) {
    const int G = g.G[a];
    lists[a].assign((size_t)G, {});
    if (a >= g.dim) {
      lists[a][0].push_back(0u);  // 2-D: the single z plane, padded index 0
      continue;
    }
    for (int n = 0; n < g.S[a]; n++) {
      const int first = n * g.sw + g.sw / 2 - g.off;  // sector center - off
      for (int i = 0; i < g.P; i++) {
        const int q = ((first + i) % G + G) % G;
        lists[a][(size_t)q].push_back(((uint32_t)n << 8) | (uint32_t)i);
      }
    }
  }
  for (int a = 0; a < 3; a++) {
    base[a] = w.size();
    uint32_t acc = 0;
    for (const auto& l : lists[a]) {
      w.push_back(acc);
      acc += (uint32_t)l.size();
    }
    w.push_back(acc);
  }
  for (int a = 0; a < 3; a++) {
    base[3 + a] = w.size();
    for (const auto& l : lists[a]) w.insert(w.end(), l.begin(), l.end());
  }
}

// ---------------------------------------------------------------------------
// Interp
// ---------------------------------------------------------------------------
template <int DIM, int NX, int CB>
__global__ void __launch_bounds__(kThreads) interp_kernel(Geom g, InterpArgs a, int R) {
  constexpr int NT = kThreads;
  constexpr int NW = NT / 64;
  constexpr int SPW = 64 / CB;  // sample slots per wave; one lane per (slot, coil)
  constexpr int RW = Rec<NX>::W;

  extern __shared__ float4 smem4[];
  char* smem = reinterpret_cast<char*>(smem4);
  const InterpLayout L(g, NX, CB, R);
  float* lut = reinterpret_cast<float*>(smem);
  float* kp = reinterpret_cast<float*>(smem + L.kp_off);
  float2* sub = reinterpret_cast<float2*>(smem + L.sub_off);
  float* rec = reinterpret_cast<float*>(smem + L.rec_off);
  const int P = g.P;
  const int Pmax = g.Pmax;
  const int KPS = P + NX;
  const int PD = DIM == 3 ? P * P * P : P * P;

  const uint32_t ci = work_chunk(a.work, blockIdx.x);
  if (ci == ~0u) return;
  const Chunk ch = a.work.chunks[ci];
  const int n = (int)(ch.end - ch.begin);
  const int coil0 = blockIdx.y * CB;
  const int tid = threadIdx.x;
  int c3[3];
  load_center<DIM>(a.centers, ch.sector, c3);

  for (int i = tid; i < g.K; i += NT) lut[i] = a.lut[i];
  fill_kpos<DIM, NT>(g, c3, kp, KPS);
  // Stage the padded sector (periodic wrap, FORWARD fftshift folded into the
  // read index), element e = pt*CB + coil; zero the overrun slack.
  for (int e = tid; e < L.sub_n; e += NT) {
    const int pt = e / CB;
    const int cc = e % CB;
    const int cgf = coil0 + cc;
    float2 v = make_float2(0.0f, 0.0f);
    if (pt < PD && cgf < a.n_coils)
      v = a.grid[(size_t)cgf * g.Ng + padded_to_grid<DIM>(g, c3, pt, a.fold_shift ? 1 : 0)];
    sub[e] = v;
  }

  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int slot = lane / CB;
  const int c = lane % CB;
  const int cg = coil0 + c;

  for (int r0 = 0; r0 < n; r0 += R) {
    const int rn = min(R, n - r0);
    __syncthreads();  // tables ready / previous round consumed
    for (int i = tid; i < rn; i += NT) {
      float p[3];
      load_coord<DIM>(a.crds, g.M, ch.begin + r0 + i, p);
      make_record<DIM, NX>(g, lut, kp, KPS, c3, p, rec + i * RW);
    }
    __syncthreads();

    // consecutive samples per wave step: the scattered stores of one wave
    // cover runs of neighbouring k-space samples
    for (int b = wave * SPW; b < rn; b += NW * SPW) {
      const int ls = b + slot;
      if (ls >= rn || cg >= a.n_coils) continue;
      float w[RW];
      load_record<NX>(rec + ls * RW, w);
      const uint32_t packed = __float_as_uint(w[0]);
      const int bx = (int)(packed & 255u);
      const int by = (int)((packed >> 8) & 255u);
      const int bz = (int)((packed >> 16) & 255u);
      const int nx = wave_trip(packed, 0, g.kw);
      const int ny = wave_trip(packed, 1, g.kw);
      const int nz = DIM == 3 ? wave_trip(packed, 2, g.kw) : 1;
      float accx = 0.0f, accy = 0.0f;
#pragma unroll
      for (int tz = 0; tz < (DIM == 3 ? NX : 1); tz++) {
        if (tz >= nz) break;
        const int iz = DIM == 3 ? min(bz + tz, Pmax) : 0;
        const float wz = DIM == 3 ? w[1 + 2 * NX + tz] : 1.0f;
#pragma unroll
        for (int ty = 0; ty < NX; ty++) {
          if (ty >= ny) break;
          const int iy = min(by + ty, Pmax);
          const float wzy = DIM == 3 ? wz * w[1 + NX + ty] : w[1 + NX + ty];
          const float2* row = sub + ((iz * P + iy) * P + bx) * CB + c;
#pragma unroll
          for (int tx = 0; tx < NX; tx++) {
            if (tx >= nx) break;
            const float val = wzy * w[1 + tx];
            const float2 v = row[tx * CB];
            accx += v.x * val;
            accy += v.y * val;
          }
        }
      }
      const uint32_t s = ch.begin + r0 + ls;
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
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
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

static int nx_of(const Geom& g) { return g.kw + 1 <= 4 ? 4 : 8; }

// Samples per round: as many as fit next to the sub-grid (budget `limit`),
// at most one per thread and at most the chunk; multiple of 16.
template <typename Layout>
static int round_size(const Geom& g, int cb, int chunk, size_t limit) {
  const int nx = nx_of(g);
  const size_t fixed = Layout(g, nx, cb, 0).bytes;
  const size_t per = Layout(g, nx, cb, 1).bytes - fixed;
  if (fixed + 16 * per > limit) return 0;
  size_t r = (limit - fixed) / per;
  r = r > (size_t)kThreads ? (size_t)kThreads : r;
  r &= ~(size_t)15;
  const int want = (chunk + 15) & ~15;
  return (int)r < want ? (int)r : want;
}

// Largest coil block (<= max_cb, not above the coil count) whose round size
// is at least min(64, chunk) samples.
template <typename Layout>
static int pick_cb(const Geom& g, int n_coils, int max_cb, int chunk, size_t limit, int* R) {
  const int need = chunk < 64 ? chunk : 64;
  for (int cb = max_cb; cb >= 1; cb /= 2) {
    if (cb > 1 && cb / 2 >= n_coils) continue;
    const int r = round_size<Layout>(g, cb, chunk, limit);
    if (r >= need || (cb == 1 && r > 0)) {
      *R = r;
      return cb;
    }
  }
  *R = 0;
  return 0;
}

template <int DIM, int NX, int CB>
static hipError_t spread_launch(const Geom& g, const SpreadArgs& a, int R, hipStream_t s) {
  const uint32_t n_blocks = work_blocks(a.work);
  const int coil_blocks = (a.n_coils + CB - 1) / CB;
  const size_t lds = SpreadLayout(g, NX, CB, R).bytes;
  hipError_t e = allow_lds(spread_kernel<DIM, NX, CB>, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((spread_kernel<DIM, NX, CB>), dim3(n_blocks, coil_blocks), dim3(kThreads), lds,
                     s, g, a, R);
  return hipGetLastError();
}

template <int DIM, int NX, int CB>
static hipError_t interp_launch(const Geom& g, const InterpArgs& a, int R, hipStream_t s) {
  const uint32_t n_blocks = work_blocks(a.work);
  const int coil_blocks = (a.n_coils + CB - 1) / CB;
  const size_t lds = InterpLayout(g, NX, CB, R).bytes;
  hipError_t e = allow_lds(interp_kernel<DIM, NX, CB>, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((interp_kernel<DIM, NX, CB>), dim3(n_blocks, coil_blocks), dim3(kThreads), lds,
                     s, g, a, R);
  return hipGetLastError();
}

template <int DIM, int NX>
static hipError_t spread_nx(const Geom& g, const SpreadArgs& a, int chunk, hipStream_t s) {
  int R = 0;
  switch (pick_cb<SpreadLayout>(g, a.n_coils, 8, chunk, lds_limit(), &R)) {
    case 8: return spread_launch<DIM, NX, 8>(g, a, R, s);
    case 4: return spread_launch<DIM, NX, 4>(g, a, R, s);
    case 2: return spread_launch<DIM, NX, 2>(g, a, R, s);
    case 1: return spread_launch<DIM, NX, 1>(g, a, R, s);
    default: return hipErrorInvalidValue;  // padded sector does not fit in LDS
  }
}

template <int DIM, int NX>
static hipError_t interp_nx(const Geom& g, const InterpArgs& a, int chunk, hipStream_t s) {
  // Prefer a footprint that lets two workgroups share a CU.
  int R = 0;
  int cb = pick_cb<InterpLayout>(g, a.n_coils, 8, chunk, lds_limit() / 2, &R);
  if (cb == 0) cb = pick_cb<InterpLayout>(g, a.n_coils, 8, chunk, lds_limit(), &R);
  switch (cb) {
    case 8: return interp_launch<DIM, NX, 8>(g, a, R, s);
    case 4: return interp_launch<DIM, NX, 4>(g, a, R, s);
    case 2: return interp_launch<DIM, NX, 2>(g, a, R, s);
    case 1: return interp_launch<DIM, NX, 1>(g, a, R, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_spread(const Geom& g, const SpreadArgs& a, int chunk, hipStream_t s) {
  if (a.n_coils == 0) return hipSuccess;
  if (g.kw > 7 || g.P > 255) return hipErrorInvalidValue;
  hipError_t e = hipSuccess;
  if (a.work.n_chunks > 0) {
    if (g.dim == 3)
      e = nx_of(g) == 4 ? spread_nx<3, 4>(g, a, chunk, s) : spread_nx<3, 8>(g, a, chunk, s);
    else
      e = nx_of(g) == 4 ? spread_nx<2, 4>(g, a, chunk, s) : spread_nx<2, 8>(g, a, chunk, s);
    if (e != hipSuccess) return e;
  }
  // the fold writes every grid point, zeros where no sample contributes
  const uint32_t S = (uint32_t)g.S[0] * g.S[1] * g.S[2];
  const int hp = g.dim == 3 ? g.sw * g.sw * g.sw : g.sw * g.sw;
  const int threads = hp >= 256 ? 256 : ((hp + 63) / 64) * 64;
  if (g.dim == 3)
    hipLaunchKernelGGL(fold_kernel<3>, dim3(S, a.n_coils), dim3(threads), 0, s, g, a);
  else
    hipLaunchKernelGGL(fold_kernel<2>, dim3(S, a.n_coils), dim3(threads), 0, s, g, a);
  return hipGetLastError();
}

hipError_t launch_interp(const Geom& g, const InterpArgs& a, int chunk, hipStream_t s) {
  if (a.work.n_chunks == 0 || a.n_coils == 0) return hipSuccess;
  if (g.kw > 7 || g.P > 255) return hipErrorInvalidValue;
  if (g.dim == 3) return nx_of(g) == 4 ? interp_nx<3, 4>(g, a, chunk, s) : interp_nx<3, 8>(g, a, chunk, s);
  return nx_of(g) == 4 ? interp_nx<2, 4>(g, a, chunk, s) : interp_nx<2, 8>(g, a, chunk, s);
}

}  // namespace gnufft
