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
// The spread accumulates the padded sector sub-grid in LDS in 64-bit fixed
// point, one word per (point, coil) laid out [point][coil] (re and im packed
// as two 32-bit halves, ds_add_u64; see the Spread section): on gfx950 a
// wave's ds_add_f32 costs ~190 clocks against ~11 for ds_add_u64 with
// random addresses, and integer sums are exact and order-independent.
//
// Two kernel families: for kernel width 3 (every shipped workload) the
// "lane" kernels form each sample's tap record in registers and share it
// across the four lanes of a quad (one lane per coil) with DPP moves; other
// widths use the round-based kernels below, which stage tap records in LDS
// per round of samples.

#include <algorithm>
#include <cstdlib>
#include <cstring>

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

// Offset of (padded point pt, coil c) inside one chunk's sub-grids: coil
// block c / sb (sb = the spread's coil block) holds its min(sb, C - block
// start) coils of every point contiguously, so each spread workgroup writes
// one contiguous run.
__device__ inline size_t sub_offset(int c, int pt, int sb, int C, int PD) {
  const int b0 = (c / sb) * sb;
  const int w = min(sb, C - b0);
  return (size_t)b0 * PD + (size_t)pt * w + (c - b0);
}

// Workgroups along x for a work list.
static inline uint32_t work_blocks(const WorkList& w) { return 8u * (w.max_range ? w.max_range : 1u); }

// Split n items of the given costs into 8 contiguous ranges of about equal
// total cost: begin[x] .. begin[x+1] is range x.
static void split_ranges(const uint32_t* costs, uint32_t n, uint32_t begin[9], uint32_t* max_range) {
  double total = 0.0;
  for (uint32_t i = 0; i < n; i++) total += (double)costs[i];
  uint32_t x = 0;
  double acc = 0.0;
  begin[0] = 0;
  for (uint32_t i = 0; i < n && x < 7; i++) {
    acc += (double)costs[i];
    while (x < 7 && acc >= total * (double)(x + 1) / 8.0) begin[++x] = i + 1;
  }
  while (x < 8) begin[++x] = n;
  *max_range = 0;
  for (int k = 0; k < 8; k++) *max_range = std::max(*max_range, begin[k + 1] - begin[k]);
}

void partition_work(const uint32_t* counts, uint32_t n, WorkList* w) {
  // Cost of a chunk: its samples plus a fixed share for staging and
  // storing the padded sub-grid (about 300 samples' worth on gfx950).
  std::vector<uint32_t> cost(n);
  for (uint32_t i = 0; i < n; i++) cost[i] = counts[i] + 300u;
  split_ranges(cost.data(), n, w->xcd_begin, &w->max_range);
  w->n_chunks = n;
}

void partition_items(const uint32_t* costs, uint32_t n, BatchList* l) {
  split_ranges(costs, n, l->xcd_begin, &l->max_range);
  l->n = n;
}

// Item of workgroup bid in a batch list (same XCD mapping as work_chunk);
// false past the end of its XCD's range.
__device__ inline bool batch_item(const BatchList& l, uint32_t bid, uint4* it) {
  const uint32_t x = bid & 7u;
  const uint32_t i = l.xcd_begin[x] + (bid >> 3);
  if (i >= l.xcd_begin[x + 1]) return false;
  const uint4 v = l.items[i];
  // workgroup-uniform: scalar registers, so the item's plan descriptor is
  // read with scalar loads (as vector values every descriptor field would
  // take VGPRs)
  *it = make_uint4(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y),
                   __builtin_amdgcn_readfirstlane(v.z), __builtin_amdgcn_readfirstlane(v.w));
  return true;
}

static inline uint32_t batch_blocks(const BatchList& l) { return 8u * (l.max_range ? l.max_range : 1u); }

template <int DIM>
__device__ inline void load_center(const Geom& g, uint32_t sector, int* c) {
  // the standard center layout (index * sw + sw / 2 per axis, the only one
  // plans accept, gpuNUFFT_operator_factory.cpp:207-214): computed, not
  // loaded, which saves a dependent memory round trip at workgroup start
  const uint32_t x = sector % (uint32_t)g.S[0];
  const uint32_t yz = sector / (uint32_t)g.S[0];
  const int half = g.sw / 2;
  c[0] = (int)x * g.sw + half;
  c[1] = (int)(yz % (uint32_t)g.S[1]) * g.sw + half;
  c[2] = DIM == 3 ? (int)(yz / (uint32_t)g.S[1]) * g.sw + half : 0;
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
// LDS base; identical on host and device). The sub-grid holds one 64-bit
// fixed-point (re, im) pair per (padded point, coil).
struct SpreadLayout {
  int kpad, kpw, sub_n;
  size_t kp_off, sub_off, rec_off, sv_off, bytes;
  __host__ __device__ SpreadLayout(const Geom& g, int nx, int cb, int R) {
    const int P = g.P;
    const int PD = g.dim == 3 ? P * P * P : P * P;
    const int rw = ((1 + 3 * nx) + 3) & ~3;
    kpad = (g.K + 3) & ~3;
    kpw = (3 * (P + nx) + 3) & ~3;
    sub_n = PD * cb + nx * cb;  // + slack for the last row's window overrun
    kp_off = (size_t)kpad * 4;
    sub_off = kp_off + (size_t)kpw * 4;
    rec_off = sub_off + (size_t)((sub_n + 1) & ~1) * 8;
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
//
// The sub-grid accumulates in 64-bit fixed point: each (point, coil) word is
// im * 2^32 + re as one two's-complement integer, re and im each a 32-bit
// multiple of 1/scale, with scale = 2^30 / L1 and L1 the chunk's sum over
// samples of max(|re|, |im|) (density-weighted) for that coil. Every kernel
// weight is at most 1, so no partial sum of either component can reach 2^31
// and the packed sum never carries wrongly between the halves. One
// ds_add_u64 per tap replaces two float atomics; integer sums are exact and
// order-independent, so the result is bit-reproducible. The quantisation
// step is L1 / 2^30 per term (relative L2 error of a chunk's sub-grid about
// 5e-10 * samples in the chunk).
// ---------------------------------------------------------------------------
constexpr double kRoundMagic = 6755399441055744.0;  // 1.5 * 2^52: x + magic rounds x to an integer
constexpr double kFixedOne = 1073741824.0;          // 2^30

__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Copy n floats to LDS with U loads in flight per thread (a plain loop
// waits for each load before its LDS write).
template <int NT, int U>
__device__ inline void copy_to_lds(float* dst, const float* __restrict__ src, int n) {
  for (int i0 = threadIdx.x; i0 < n; i0 += U * NT) {
    float t[U];
#pragma unroll
    for (int u = 0; u < U; u++) t[u] = i0 + u * NT < n ? src[i0 + u * NT] : 0.0f;
#pragma unroll
    for (int u = 0; u < U; u++)
      if (i0 + u * NT < n) dst[i0 + u * NT] = t[u];
  }
}

// One workgroup's share of the spread: chunk ci of the work list, coils
// [coil_block * CB, coil_block * CB + CB).
template <int DIM, int NX, int CB, int NT = kThreads>
__device__ inline void spread_body(const Geom& g, const SpreadArgs& a, uint32_t ci, int coil_block,
                                   int R) {
  constexpr int NW = NT / 64;
  constexpr int SPW = 64 / CB;  // sample slots per wave; one lane per (slot, coil)
  constexpr int RW = Rec<NX>::W;

  extern __shared__ float4 smem4[];
  __shared__ float l1_part[NW][CB];
  __shared__ double scale_s[CB];
  __shared__ double inv_s[CB];
  char* smem = reinterpret_cast<char*>(smem4);
  const SpreadLayout L(g, NX, CB, R);
  float* lut = reinterpret_cast<float*>(smem);
  float* kp = reinterpret_cast<float*>(smem + L.kp_off);
  unsigned long long* sub = reinterpret_cast<unsigned long long*>(smem + L.sub_off);
  float* rec = reinterpret_cast<float*>(smem + L.rec_off);
  float2* sv = reinterpret_cast<float2*>(smem + L.sv_off);
  const int P = g.P;
  const int Pmax = g.Pmax;
  const int KPS = P + NX;
  const int PD = DIM == 3 ? P * P * P : P * P;

  const Chunk ch = a.work.chunks[ci];
  const int n = (int)(ch.end - ch.begin);
  const int coil0 = coil_block * CB;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  int c3[3];
  load_center<DIM>(g, ch.sector, c3);

  copy_to_lds<NT, 4>(lut, a.lut, g.K);
  fill_kpos<DIM, NT>(g, c3, kp, KPS);
  {
    // 16-byte stores; the allocation rounds sub_n up to an even count
    ulonglong2* sub2 = reinterpret_cast<ulonglong2*>(sub);
    for (int i = tid; i < (L.sub_n + 1) / 2; i += NT) sub2[i] = make_ulonglong2(0ull, 0ull);
  }

  // per-coil fixed-point scale from the chunk's L1 bound; the raw values of
  // the first round's samples stay in sv (round 0 scales them in place
  // instead of reading them again)
  {
    float l1[CB];
#pragma unroll
    for (int cc = 0; cc < CB; cc++) l1[cc] = 0.0f;
    for (int i = tid; i < n; i += NT) {
      const uint32_t s = ch.begin + i;
      const uint32_t src = a.gather ? a.gather[s] : s;
      const float sq = a.dens ? sqrtf(a.dens[s]) : 1.0f;
#pragma unroll
      for (int cc = 0; cc < CB; cc++)
        if (coil0 + cc < a.n_coils) {
          const float2 v = a.data[(size_t)(coil0 + cc) * a.coil_stride + src];
          l1[cc] += fmaxf(fabsf(v.x), fabsf(v.y)) * sq;
          if (i < R) sv[i * CB + cc] = v;
        }
    }
#pragma unroll
    for (int cc = 0; cc < CB; cc++) {
      const float t = wave_sum(l1[cc]);
      if (lane == 0) l1_part[wave][cc] = t;
    }
    __syncthreads();
    if (tid < CB) {
      double t = 0.0;
      for (int w = 0; w < NW; w++) t += (double)l1_part[w][tid];
      const double sc = t > 0.0 ? kFixedOne / t : 0.0;
      scale_s[tid] = sc;
      inv_s[tid] = sc > 0.0 ? 1.0 / sc : 0.0;
    }
  }

  const int slot = lane / CB;
  const int c = lane % CB;
  const bool coil_ok = coil0 + c < a.n_coils;

  for (int r0 = 0; r0 < n; r0 += R) {
    const int rn = min(R, n - r0);
    __syncthreads();  // tables and scales ready / previous round consumed
    for (int i = tid; i < rn; i += NT) {
      const uint32_t s = ch.begin + r0 + i;
      float p[3];
      load_coord<DIM>(a.crds, g.M, s, p);
      const uint32_t src = a.gather ? a.gather[s] : s;
      const double sq = a.dens ? (double)sqrtf(a.dens[s]) : 1.0;
#pragma unroll
      for (int cc = 0; cc < CB; cc++) {
        float2 v = make_float2(0.0f, 0.0f);
        if (coil0 + cc < a.n_coils) {
          // round 0: this thread staged sample i's raw value in the L1 pass
          v = r0 == 0 ? sv[i * CB + cc] : a.data[(size_t)(coil0 + cc) * a.coil_stride + src];
          const double k = sq * scale_s[cc];
          v = make_float2((float)(v.x * k), (float)(v.y * k));
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
          unsigned long long* row = sub + ((iz * P + iy) * P + bx) * CB + c;
#pragma unroll
          for (int tx = 0; tx < NX; tx++) {
            if (tx >= nx) break;
            const double wt = (double)(wzy * w[1 + tx]);
            // round(wt * v) in the low word of (wt * v + magic)
            const uint32_t re = (uint32_t)__double_as_longlong(fma(wt, vx, kRoundMagic));
            const uint32_t im = (uint32_t)__double_as_longlong(fma(wt, vy, kRoundMagic));
            const uint32_t hi = im + (uint32_t)((int32_t)re >> 31);  // borrow of a negative re
            atomicAdd(row + tx * CB, ((unsigned long long)hi << 32) | re);
          }
        }
      }
    }
  }
  __syncthreads();

  // Store the chunk's sub-grid, [padded point][coil] like the LDS copy, with
  // plain coalesced stores; fold_kernel sums the sub-grids that overlap on
  // the grid (the coils of a point are one contiguous run for its reads).
  // this block's min(CB, C - coil0) coils of every point, one contiguous run
  float2* dst = a.scratch + ((size_t)ci * a.n_coils + coil0) * PD;
  const int wb = min(CB, a.n_coils - coil0);
  for (int e = tid; e < PD * CB; e += NT) {
    const int pt = e / CB;
    const int cc = e % CB;
    if (cc >= wb) continue;
    const unsigned long long u = sub[e];
    const int32_t re = (int32_t)(uint32_t)u;
    const int32_t im = (int32_t)(uint32_t)(u >> 32) - (re >> 31);
    const double inv = inv_s[cc];
    dst[(size_t)pt * wb + cc] = make_float2((float)(re * inv), (float)(im * inv));
  }
}

template <int DIM, int NX, int CB>
__global__ void __launch_bounds__(kThreads) spread_kernel(Geom g, SpreadArgs a, int R) {
  const uint32_t ci = work_chunk(a.work, blockIdx.x);
  if (ci == ~0u) return;
  spread_body<DIM, NX, CB>(g, a, ci, blockIdx.y, R);
}

// ---------------------------------------------------------------------------
// Spread with tap records in registers (kernel widths with a compile-time
// tap count KW; KW = 3 is instantiated).
//
// Each thread owns whole samples (sample j = round * NT + tid): it loads
// their coordinates and density and forms their tap records (first tap per
// axis, KW or KW+1 weights per axis) in registers. The four lanes of a quad
// are one sample's coils (CB = 4; with CB = 2 two samples of two coils, with
// CB = 1 four samples): in pass ps the quad works on the record of its lane
// ps, fetched with DPP quad_perm moves (VALU, no LDS), and each lane adds
// all KW^d taps of its coil. The CB coils of a point are consecutive LDS
// words, so one atomic instruction touches 64 / CB points in runs of CB
// words (as few bank conflicts as the [point][coil] layout allows), and the
// samples of one instruction are four apart in the chunk.
//
// Nothing but the sub-grid, the kernel table and the per-axis k-space
// positions lives in LDS; there are no LDS round trips of records or values
// and three barriers per chunk (tables and L1 partials, scales, taps done).
//
// Per tap: two f64 fmas (the value pre-multiplied by the row weight wz*wy,
// then by wx, rounded to an integer by the 1.5*2^52 trick), two integer
// operations and one ds_add_u64 whose x offset is an immediate.
//
// Taps run KW per axis with no clamping. A window starts at most at Pmax-1
// (the clamps of the reference's tap range), so the last of KW+1 taps is at
// most index P+1; the sub-grid has a slack of 2P^2+2P+2 words (3-D; 2P+2 in
// 2-D) past its end for such indices, whose taps carry weight zero (they lie
// past the reference's last tap). A wave with any sample needing KW+1 taps
// along an axis (a position exactly half-way between grid points) runs its
// samples with KW+1 taps per axis.
// ---------------------------------------------------------------------------
struct LaneLayout {
  int kpad, kpw, sub_n;
  size_t kp_off, sub_off, bytes;
  // lut: the kernel table is copied to LDS (else read from global memory)
  __host__ __device__ LaneLayout(const Geom& g, int kw, int cb, bool lut = true) {
    const int P = g.P;
    const int PD = g.dim == 3 ? P * P * P : P * P;
    const int slack = g.dim == 3 ? 2 * P * P + 2 * P + 2 : 2 * P + 2;
    kpad = lut ? (g.K + 3) & ~3 : 0;
    kpw = (3 * (P + kw + 1) + 3) & ~3;
    sub_n = (PD + slack) * cb;
    kp_off = (size_t)kpad * 4;
    sub_off = (kp_off + (size_t)kpw * 4 + 15) & ~(size_t)15;
    bytes = sub_off + (size_t)((sub_n + 1) & ~1) * 8;
  }
};

// Fixed-point add of v * wt (both real parts pre-scaled) to one sub-grid word.
__device__ inline void fx_add(unsigned long long* w, double wt, double vx, double vy) {
  const uint32_t re = (uint32_t)__double_as_longlong(fma(wt, vx, kRoundMagic));
  const uint32_t im = (uint32_t)__double_as_longlong(fma(wt, vy, kRoundMagic));
  const uint32_t hi = im + (uint32_t)((int32_t)re >> 31);  // borrow of a negative re
  atomicAdd(w, ((unsigned long long)hi << 32) | re);
}

// Taps of one sample and one coil: T per axis, row-major (z, y, x) as the
// reference forms the weight (wz*wy)*wx. Indices are not clamped: taps past
// the sample's window carry weight zero and land in the sub-grid or its
// slack.
template <int DIM, int T, int CB>
__device__ inline void lane_taps(unsigned long long* sub, int P, const int* b, const float (*w)[4],
                                 double vx, double vy, int cc) {
  double wxd[T];
#pragma unroll
  for (int t = 0; t < T; t++) wxd[t] = (double)w[0][t];
  unsigned long long* base = sub + ((b[2] * P + b[1]) * P + b[0]) * CB + cc;
#pragma unroll
  for (int tz = 0; tz < (DIM == 3 ? T : 1); tz++) {
    const float wz = DIM == 3 ? w[2][tz] : 1.0f;
#pragma unroll
    for (int ty = 0; ty < T; ty++) {
      const float wzy = DIM == 3 ? wz * w[1][ty] : w[1][ty];
      const double rx = (double)wzy * vx, ry = (double)wzy * vy;
      unsigned long long* row = base + (tz * P + ty) * P * CB;
#pragma unroll
      for (int tx = 0; tx < T; tx++) fx_add(row + tx * CB, wxd[tx], rx, ry);
      // one row at a time: the scheduler would otherwise hoist the next
      // rows' operands and spill
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// Value of lane `src` of this lane's quad, for each of the four lanes of the
// quad (DPP quad_perm: a VALU move, no LDS).
template <int A, int B, int C, int D>
__device__ inline uint32_t quad_get(uint32_t v) {
  if (A == 0 && B == 1 && C == 2 && D == 3) return v;
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, A | (B << 2) | (C << 4) | (D << 6), 0xF, 0xF, false);
}

// Source lane pattern of pass p: lane qi of a quad works on the sample owned
// by lane p * (4 / CB) + qi / CB, coil qi % CB.
template <int CB, int PS>
__device__ inline uint32_t quad_pass(uint32_t v) {
  constexpr int SPQ = 4 / CB;
  return quad_get<PS * SPQ + 0 / CB, PS * SPQ + 1 / CB, PS * SPQ + 2 / CB, PS * SPQ + 3 / CB>(v);
}

template <int CB>
__device__ inline uint32_t quad_pass_rt(uint32_t v, int ps) {
  // passes are unrolled; this maps the runtime index onto the compile-time
  // patterns
  switch (ps) {
    case 0: return quad_pass<CB, 0>(v);
    case 1: return CB > 1 ? quad_pass<CB, (CB > 1 ? 1 : 0)>(v) : v;
    case 2: return CB > 2 ? quad_pass<CB, (CB > 2 ? 2 : 0)>(v) : v;
    default: return CB > 3 ? quad_pass<CB, (CB > 3 ? 3 : 0)>(v) : v;
  }
}

// LUTG: kernel table read from global memory (cached) instead of an LDS
// copy, for small workgroups whose few samples make fewer lookups than the
// copy has words.
template <int DIM, int KW, int CB, int NT = kThreads, bool LUTG = false>
__device__ inline void spread_lane_body(const Geom& g, const SpreadArgs& a, uint32_t ci,
                                        int coil_block) {
  constexpr int NW = NT / 64;
  constexpr int NX = KW + 1;
  constexpr int RND = 2;        // rounds of NT owned samples: chunks hold at most 2 * NT
  constexpr int PASSES = CB;    // passes per round: one per coil of a quad's sample
  static_assert(CB == 1 || CB == 2 || CB == 4, "a sample's coils share one quad");
  static_assert(NX <= 4, "tap records hold at most four taps per axis");

  extern __shared__ float4 smem4[];
  __shared__ float l1_part[NW][CB];
  __shared__ double scale_s[CB];
  __shared__ double inv_s[CB];
  char* smem = reinterpret_cast<char*>(smem4);
  const LaneLayout L(g, KW, CB, !LUTG);
  float* lut = reinterpret_cast<float*>(smem);
  float* kp = reinterpret_cast<float*>(smem + L.kp_off);
  unsigned long long* sub = reinterpret_cast<unsigned long long*>(smem + L.sub_off);
  const int P = g.P;
  const int Pmax = g.Pmax;
  const int KPS = P + NX;
  const int PD = DIM == 3 ? P * P * P : P * P;

  const Chunk ch = a.work.chunks[ci];
  const int n = (int)(ch.end - ch.begin);
  const int coil0 = coil_block * CB;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int qc = (lane & 3) % CB;           // coil (in the block) of this lane
  const bool coil_ok = coil0 + qc < a.n_coils;

  // owned samples: j = r * NT + tid (a quad owns four consecutive samples,
  // so one atomic's lanes take samples four apart; waves past the chunk's
  // end own none and skip the taps)
  float p[RND][3];
  float sq[RND];
  uint32_t src[RND];
  bool live[RND];
#pragma unroll
  for (int r = 0; r < RND; r++) {
    const int j = r * NT + tid;
    live[r] = j < n;
    p[r][0] = p[r][1] = p[r][2] = 0.0f;
    sq[r] = 0.0f;
    src[r] = 0u;
    if (!live[r]) continue;
    const uint32_t s = ch.begin + (uint32_t)j;
    load_coord<DIM>(a.crds, g.M, s, p[r]);
    sq[r] = a.dens ? sqrtf(a.dens[s]) : 1.0f;
    src[r] = a.gather ? a.gather[s] : s;
  }
  // values: in pass ps this lane works on the sample of quad lane
  // ps * (4 / CB) + qi / CB, coil qc
  // (value times the square root of the density weight, in float)
  float2 v[RND][PASSES];
  bool livep[RND][PASSES];
#pragma unroll
  for (int r = 0; r < RND; r++) {
#pragma unroll
    for (int ps = 0; ps < PASSES; ps++) {
      const uint32_t so = quad_pass_rt<CB>(src[r], ps);
      livep[r][ps] = quad_pass_rt<CB>(live[r] ? 1u : 0u, ps) != 0u;
      v[r][ps] = make_float2(0.0f, 0.0f);
      if (livep[r][ps] && coil_ok) v[r][ps] = a.data[(size_t)(coil0 + qc) * a.coil_stride + so];
    }
  }
#pragma unroll
  for (int r = 0; r < RND; r++)
#pragma unroll
    for (int ps = 0; ps < PASSES; ps++) {
      const float sqo = __uint_as_float(quad_pass_rt<CB>(__float_as_uint(sq[r]), ps));
      v[r][ps].x *= sqo;
      v[r][ps].y *= sqo;
    }

  int c3[3];
  load_center<DIM>(g, ch.sector, c3);
  if (!LUTG) copy_to_lds<NT, 4>(lut, a.lut, g.K);
  const float* wlut = LUTG ? a.lut : lut;
  fill_kpos<DIM, NT>(g, c3, kp, KPS);
  {
    ulonglong2* sub2 = reinterpret_cast<ulonglong2*>(sub);
    for (int i = tid; i < (L.sub_n + 1) / 2; i += NT) sub2[i] = make_ulonglong2(0ull, 0ull);
  }

  // per-coil L1 of the chunk (density-weighted): the fixed-point scale
  {
    float l1 = 0.0f;
#pragma unroll
    for (int r = 0; r < RND; r++)
#pragma unroll
      for (int ps = 0; ps < PASSES; ps++)
        l1 += fmaxf(fabsf(v[r][ps].x), fabsf(v[r][ps].y));
    // sum over the lanes of the same coil (lane bits from log2(CB) up)
#pragma unroll
    for (int o = CB; o < 64; o <<= 1) l1 += __shfl_xor(l1, o);
    if (lane < CB) l1_part[wave][lane] = l1;
  }
  __syncthreads();  // tables, zeroed sub-grid and L1 partials ready
  if (tid < CB) {
    double t = 0.0;
    for (int w = 0; w < NW; w++) t += (double)l1_part[w][tid];
    const double sc = t > 0.0 ? kFixedOne / t : 0.0;
    scale_s[tid] = sc;
    inv_s[tid] = sc > 0.0 ? 1.0 / sc : 0.0;
  }
  __syncthreads();  // scales ready
  const double scale = scale_s[qc];

#pragma unroll
  for (int r = 0; r < RND; r++) {
    if (r * NT >= n) break;  // workgroup-uniform: no thread owns a sample of round r
    if (r * NT + wave * 64 >= n) continue;  // wave-uniform: this wave owns none
    // own tap record: first tap per axis (packed, 8 bits each) and NX
    // weights per axis (zero past the sample's taps or failing the radius
    // test)
    uint32_t bp = 0;
    float w[3][4];
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int t = 0; t < 4; t++) w[d][t] = 0.0f;
    bool wide = false;
    if (live[r]) {
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        const float u = map_kspace_to_grid(p[r][d], g.G[d], c3[d], g.off);
        int lo, hi;
        tap_range(u, Pmax, g.kradius, &lo, &hi);
        bp |= (uint32_t)lo << (8 * d);
        wide = wide || hi - lo + 1 > KW;
#pragma unroll
        for (int t = 0; t < NX; t++) {
          float wt = 0.0f;
          if (lo + t <= hi && axis_weight_kp(g, wlut, d, kp[d * KPS + lo + t], p[r][d], &wt))
            w[d][t] = wt;
        }
      }
    }
    const bool any_wide = __ballot(wide) != 0ull;
#pragma unroll
    for (int ps = 0; ps < PASSES; ps++) {
      // the worked sample's record from its owner lane (every lane takes
      // part in the moves)
      const uint32_t bo = quad_pass_rt<CB>(bp, ps);
      float wo[3][4];
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int t = 0; t < 4; t++)
          wo[d][t] = (d < DIM && t < (any_wide ? NX : KW))
                         ? __uint_as_float(quad_pass_rt<CB>(__float_as_uint(w[d][t]), ps))
                         : 0.0f;
      if (!livep[r][ps] || !coil_ok) continue;
      const int b[3] = {(int)(bo & 255u), (int)((bo >> 8) & 255u), (int)((bo >> 16) & 255u)};
      const double vx = (double)v[r][ps].x * scale, vy = (double)v[r][ps].y * scale;
      if (any_wide)
        lane_taps<DIM, NX, CB>(sub, P, b, wo, vx, vy, qc);
      else
        lane_taps<DIM, KW, CB>(sub, P, b, wo, vx, vy, qc);
    }
  }
  __syncthreads();

  // store the chunk's sub-grid (as spread_body)
  // this block's min(CB, C - coil0) coils of every point, one contiguous run
  float2* dst = a.scratch + ((size_t)ci * a.n_coils + coil0) * PD;
  const int wb = min(CB, a.n_coils - coil0);
  for (int e = tid; e < PD * CB; e += NT) {
    const int pt = e / CB;
    const int cc = e % CB;
    if (cc >= wb) continue;
    const unsigned long long u = sub[e];
    const int32_t re = (int32_t)(uint32_t)u;
    const int32_t im = (int32_t)(uint32_t)(u >> 32) - (re >> 31);
    const double inv = inv_s[cc];
    dst[(size_t)pt * wb + cc] = make_float2((float)(re * inv), (float)(im * inv));
  }
}

template <int DIM, int KW, int CB>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(6))) spread_lane_kernel(Geom g, SpreadArgs a) {
  const uint32_t ci = work_chunk(a.work, blockIdx.x);
  if (ci == ~0u) return;
  spread_lane_body<DIM, KW, CB>(g, a, ci, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Fold: grid point q of coil c = sum over the padded sectors covering q and
// over each such sector's chunks of the sub-grid value at q's padded index.
// One workgroup per (home sector, coil); the home sector of q is q / sw per
// axis, so every grid point is written exactly once, in a fixed summation
// order (deterministic). The loops over candidate sectors and their chunks
// are uniform across the workgroup; each thread tests its own points
// arithmetically, so the scratch reads carry no dependent-load chains.
// Periodic wrap and the optional INVERSE fftshift of the store
// (std_gpuNUFFT_kernels.cu:629-654) are folded into the indices.
// ---------------------------------------------------------------------------
constexpr size_t kInterpStaticLds = 4096;  // interp's static shared tables, rounded up
constexpr int kFoldThreads = 256;
// home points per thread per pass: one in 3-D (8 coils x 2 points would
// take 125 VGPRs, four waves per SIMD; one point: 71, seven), two in 2-D
// (batches of sparse 2-D sectors are faster with two)
__host__ __device__ constexpr int fold_pts(int dim) { return dim == 3 ? 1 : 2; }
constexpr int kFoldCands = 256;  // candidate sectors per home sector (host-checked)

// Sum every chunk sub-grid of a multi-chunk sector into the slot of its
// first chunk, in chunk order (deterministic); the fold then reads one
// sub-grid per sector. Task (first chunk, chunk count) per blockIdx.y.
__device__ inline void reduce_body(float2* scratch, size_t kstride, uint32_t k0, uint32_t cnt,
                                   size_t e) {
  if (e >= kstride) return;
  float2* base = scratch + (size_t)k0 * kstride + e;
  float ax = base[0].x, ay = base[0].y;
#pragma unroll 8
  for (uint32_t k = 1; k < cnt; k++) {
    const float2 v = base[(size_t)k * kstride];
    ax += v.x;
    ay += v.y;
  }
  base[0] = make_float2(ax, ay);
}

__global__ void __launch_bounds__(256) reduce_chunks_kernel(const uint2* __restrict__ tasks,
                                                            float2* __restrict__ scratch,
                                                            size_t kstride) {
  const uint2 t = tasks[blockIdx.y];
  reduce_body(scratch, kstride, t.x, t.y, (size_t)blockIdx.x * 256 + threadIdx.x);
}

// Fold of one home sector, driven by per-axis coverage tables: for each
// grid coordinate of the sector along an axis, the (candidate, padded index)
// pairs whose padded region covers it. A point visits exactly the sub-grids
// covering it (one for most interior points, up to 2^d at sector corners
// for P <= 2 sw), in a fixed (z, y, x) order: deterministic sums, and loads
// only of words that contribute. Needs max_cov <= kCovMax (host-checked).
constexpr int kCovMax = 4;
constexpr int kFoldMaxH = 256;  // sector width bound (P <= 255)

template <int DIM, int CB>
__device__ __forceinline__ void fold_body_cov(const Geom& g, const SpreadArgs& a, uint32_t s, int coil_block) {
  __shared__ int kofs[kFoldCands];                   // chunk of candidate (jz, jy, jx), -1 empty
  __shared__ uint16_t cov[3][kFoldMaxH * kCovMax];   // candidate | padded index << 8
  __shared__ uint8_t ncov[3][kFoldMaxH];
  const int P = g.P;
  const int PD = DIM == 3 ? P * P * P : P * P;
  const int sw = g.sw;
  const int coil0 = coil_block * CB;
  const int sx = (int)(s % (uint32_t)g.S[0]);
  const uint32_t r = s / (uint32_t)g.S[0];
  const int sy = (int)(r % (uint32_t)g.S[1]);
  const int sz = DIM == 3 ? (int)(r / (uint32_t)g.S[1]) : 0;
  const int hs[3] = {sx, sy, sz};
  const int h0[3] = {sx * sw, sy * sw, sz * sw};
  const int hx = min(sw, g.G[0] - h0[0]);
  const int hy = min(sw, g.G[1] - h0[1]);
  const int hz = DIM == 3 ? min(sw, g.G[2] - h0[2]) : 1;
  const int HP = hx * hy * hz;
  const uint32_t cx0 = a.cover.off[0][sx], ncx = a.cover.off[0][sx + 1] - cx0;
  const uint32_t cy0 = a.cover.off[1][sy], ncy = a.cover.off[1][sy + 1] - cy0;
  const uint32_t cz0 = DIM == 3 ? a.cover.off[2][sz] : 0;
  const uint32_t ncz = DIM == 3 ? a.cover.off[2][sz + 1] - cz0 : 1;
  const int J = (int)(ncx * ncy * ncz);
  const int tid = threadIdx.x, nt = (int)blockDim.x;
  for (int j = tid; j < J; j += nt) {
    const int jx = j % (int)ncx;
    const int jy = (j / (int)ncx) % (int)ncy;
    const int jz = j / (int)(ncx * ncy);
    const int nx = (int)a.cover.ent[0][cx0 + jx];
    const int ny = (int)a.cover.ent[1][cy0 + jy];
    const int nz = DIM == 3 ? (int)a.cover.ent[2][cz0 + jz] : 0;
    const int n = nx + g.S[0] * (ny + g.S[1] * nz);
    const int k0 = (int)a.sector_chunks[n];
    kofs[j] = (int)a.sector_chunks[n + 1] > k0 ? k0 : -1;
  }
  for (int e = tid; e < hx + hy + (DIM == 3 ? hz : 0); e += nt) {
    const int ax = e < hx ? 0 : (e < hx + hy ? 1 : 2);
    const int t = e - (ax == 0 ? 0 : (ax == 1 ? hx : hx + hy));
    const int q = h0[ax] + t;
    const uint32_t c0 = a.cover.off[ax][hs[ax]], nc = a.cover.off[ax][hs[ax] + 1] - c0;
    int count = 0;
    for (uint32_t j = 0; j < nc; j++) {
      const int origin = (int)a.cover.ent[ax][c0 + j] * sw + sw / 2 - g.off;
      for (int i = wrap(q - origin, g.G[ax]); i < P && count < kCovMax; i += g.G[ax])
        cov[ax][t * kCovMax + count++] = (uint16_t)(j | ((uint32_t)i << 8));
    }
    ncov[ax][t] = (uint8_t)count;
  }
  __syncthreads();
  const size_t kstride = (size_t)a.n_coils * PD;
  const int sb = a.sub_cb;
  // full block, even coil count and even spread blocks: every coil pair of
  // a point is one 16-byte word of one spread block
  const bool vec = CB % 2 == 0 && coil0 + CB <= a.n_coils && a.n_coils % 2 == 0 && sb % 2 == 0;
  for (int base = 0; base < HP; base += nt * fold_pts(DIM)) {
#pragma unroll
    for (int p = 0; p < fold_pts(DIM); p++) {
      const int e = base + p * nt + tid;
      if (e >= HP) continue;
      const int tx = e % hx;
      const int tyz = e / hx;
      const int ty = tyz % hy;
      const int tz = DIM == 3 ? tyz / hy : 0;
      float ax_[CB], ay_[CB];
#pragma unroll
      for (int c = 0; c < CB; c++) {
        ax_[c] = 0.0f;
        ay_[c] = 0.0f;
      }
      const int nzc = DIM == 3 ? ncov[2][tz] : 1;
      const int nyc = ncov[1][ty], nxc = ncov[0][tx];
      for (int iz = 0; iz < nzc; iz++) {
        const uint32_t vz = DIM == 3 ? cov[2][tz * kCovMax + iz] : 0u;
        const int jz = (int)(vz & 255u), pz = (int)(vz >> 8);
        for (int iy = 0; iy < nyc; iy++) {
          const uint32_t vy = cov[1][ty * kCovMax + iy];
          const int jyz = jz * (int)ncy + (int)(vy & 255u);
          const int pyz = P * ((int)(vy >> 8) + P * pz);
          for (int ix = 0; ix < nxc; ix++) {
            const uint32_t vx = cov[0][tx * kCovMax + ix];
            const int k = kofs[jyz * (int)ncx + (int)(vx & 255u)];
            if (k < 0) continue;
            const float2* src = a.scratch + (size_t)k * kstride;
            const int pt = pyz + (int)(vx >> 8);
            if (vec) {
              // the block's CB coils: CB/2 16-byte loads (16-byte aligned)
#pragma unroll
              for (int c2 = 0; c2 < CB / 2; c2++) {
                const float4 val = *reinterpret_cast<const float4*>(
                    src + sub_offset(coil0 + 2 * c2, pt, sb, a.n_coils, PD));
                ax_[2 * c2] += val.x;
                ay_[2 * c2] += val.y;
                ax_[2 * c2 + 1] += val.z;
                ay_[2 * c2 + 1] += val.w;
              }
            } else {
#pragma unroll
              for (int c = 0; c < CB; c++) {
                const float2 val = src[sub_offset(min(coil0 + c, a.n_coils - 1), pt, sb, a.n_coils, PD)];
                ax_[c] += val.x;
                ay_[c] += val.y;
              }
            }
          }
        }
      }
      int o0 = h0[0] + tx, o1 = h0[1] + ty, o2 = h0[2] + tz;
      if (a.fold_shift) {
        o0 = wrap(o0 - g.shift_inv[0], g.G[0]);
        o1 = wrap(o1 - g.shift_inv[1], g.G[1]);
        if (DIM == 3) o2 = wrap(o2 - g.shift_inv[2], g.G[2]);
      }
      const size_t gi = (size_t)o0 + (size_t)g.G[0] * ((size_t)o1 + (size_t)g.G[1] * o2);
#pragma unroll
      for (int c = 0; c < CB; c++)
        if (coil0 + c < a.n_coils) a.grid[(size_t)(coil0 + c) * g.Ng + gi] = make_float2(ax_[c], ay_[c]);
    }
  }
}

// Fold driven by the item list of all (candidate, periodic image) pairs;
// every point reads every item (masked). For geometries beyond kCovMax.
template <int DIM, int CB>
__device__ __forceinline__ void fold_body_generic(const Geom& g, const SpreadArgs& a, uint32_t s, int coil_block) {
  // item per non-empty candidate sector: padded origin per axis and the
  // chunk holding the sector's (reduced) sub-grid; rep: packed recurrence
  // offsets (P > G only)
  __shared__ int4 items[kFoldCands];
  __shared__ uint32_t reps_of[kFoldCands];
  __shared__ int slot[kFoldCands + 1];
  const int P = g.P;
  const int PD = DIM == 3 ? P * P * P : P * P;
  const int sw = g.sw;
  const int coil0 = coil_block * CB;
  const int sx = (int)(s % (uint32_t)g.S[0]);
  const uint32_t r = s / (uint32_t)g.S[0];
  const int sy = (int)(r % (uint32_t)g.S[1]);
  const int sz = DIM == 3 ? (int)(r / (uint32_t)g.S[1]) : 0;
  const int h0[3] = {sx * sw, sy * sw, sz * sw};
  const int hx = min(sw, g.G[0] - h0[0]);
  const int hy = min(sw, g.G[1] - h0[1]);
  const int hz = DIM == 3 ? min(sw, g.G[2] - h0[2]) : 1;
  const int HP = hx * hy * hz;
  const int rx_n = (P + g.G[0] - 1) / g.G[0];
  const int ry_n = (P + g.G[1] - 1) / g.G[1];
  const int rz_n = DIM == 3 ? (P + g.G[2] - 1) / g.G[2] : 1;
  const uint32_t cz0 = DIM == 3 ? a.cover.off[2][sz] : 0, cz1 = DIM == 3 ? a.cover.off[2][sz + 1] : 1;
  const uint32_t cy0 = a.cover.off[1][sy], cy1 = a.cover.off[1][sy + 1];
  const uint32_t cx0 = a.cover.off[0][sx], cx1 = a.cover.off[0][sx + 1];
  const int ncx = (int)(cx1 - cx0), ncy = (int)(cy1 - cy0), ncz = (int)(cz1 - cz0);
  const int nreps = rx_n * ry_n * rz_n;
  const int J = ncx * ncy * ncz * nreps;  // candidates, fixed (z, y, x, rep) order
  const int tid = threadIdx.x;

  int4 mine = make_int4(0, 0, 0, 0);
  uint32_t my_rep = 0;
  bool my_live = false;
  if (tid < J) {
    int j = tid;
    const int rr = j % nreps;
    j /= nreps;
    const int jx = j % ncx;
    j /= ncx;
    const int jy = j % ncy;
    const int jz = j / ncy;
    const int nx = (int)a.cover.ent[0][cx0 + jx];
    const int ny = (int)a.cover.ent[1][cy0 + jy];
    const int nz = DIM == 3 ? (int)a.cover.ent[2][cz0 + jz] : 0;
    const int n = nx + g.S[0] * (ny + g.S[1] * nz);
    const int k0 = (int)a.sector_chunks[n];
    my_live = (int)a.sector_chunks[n + 1] > k0;
    mine = make_int4(nx * sw + sw / 2 - g.off, ny * sw + sw / 2 - g.off, nz * sw + sw / 2 - g.off, k0);
    my_rep = (uint32_t)(rr % rx_n) | ((uint32_t)((rr / rx_n) % ry_n) << 8) |
             ((uint32_t)(rr / (rx_n * ry_n)) << 16);
    slot[tid] = my_live ? 1 : 0;
  }
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int j = 0; j < J; j++) {
      const int c = slot[j];
      slot[j] = acc;
      acc += c;
    }
    slot[J] = acc;
  }
  __syncthreads();
  if (tid < J && my_live) {
    items[slot[tid]] = mine;
    reps_of[slot[tid]] = my_rep;
  }
  __syncthreads();
  const int n_items = slot[J];

  const int nt = (int)blockDim.x;
  const size_t kstride = (size_t)a.n_coils * PD;
  for (int base = 0; base < HP; base += nt * fold_pts(DIM)) {
    int q[fold_pts(DIM)][3];
    bool live[fold_pts(DIM)];
    float ax[fold_pts(DIM)][CB], ay[fold_pts(DIM)][CB];
#pragma unroll
    for (int p = 0; p < fold_pts(DIM); p++) {
      const int e = base + p * nt + tid;
      live[p] = e < HP;
      const int lx = e % hx;
      const int lyz = e / hx;
      q[p][0] = h0[0] + lx;
      q[p][1] = h0[1] + lyz % hy;
      q[p][2] = DIM == 3 ? h0[2] + lyz / hy : 0;
#pragma unroll
      for (int c = 0; c < CB; c++) {
        ax[p][c] = 0.0f;
        ay[p][c] = 0.0f;
      }
    }
    // Loads are unconditional (points outside a candidate's padded region
    // read its first word and add zero), so the loads of several items issue
    // back to back instead of one branch-guarded round trip each.
#pragma unroll 2
    for (int it = 0; it < n_items; it++) {
      const int4 v = items[it];
      const uint32_t rp = reps_of[it];
      const float2* src = a.scratch + (size_t)v.w * kstride;
#pragma unroll
      for (int p = 0; p < fold_pts(DIM); p++) {
        const int ix = wrap(q[p][0] - v.x, g.G[0]) + (int)(rp & 255u) * g.G[0];
        const int iy = wrap(q[p][1] - v.y, g.G[1]) + (int)((rp >> 8) & 255u) * g.G[1];
        const int iz = DIM == 3 ? wrap(q[p][2] - v.z, g.G[2]) + (int)(rp >> 16) * g.G[2] : 0;
        const bool in = live[p] && ix < P && iy < P && iz < P;
        const int pt = in ? ix + P * (iy + P * iz) : 0;
        const float m = in ? 1.0f : 0.0f;
#pragma unroll
        for (int c = 0; c < CB; c++) {
          const float2 val = src[sub_offset(min(coil0 + c, a.n_coils - 1), pt, a.sub_cb, a.n_coils, PD)];
          ax[p][c] += m * val.x;
          ay[p][c] += m * val.y;
        }
      }
    }
#pragma unroll
    for (int p = 0; p < fold_pts(DIM); p++) {
      if (!live[p]) continue;
      int o0 = q[p][0], o1 = q[p][1], o2 = q[p][2];
      if (a.fold_shift) {
        o0 = wrap(o0 - g.shift_inv[0], g.G[0]);
        o1 = wrap(o1 - g.shift_inv[1], g.G[1]);
        if (DIM == 3) o2 = wrap(o2 - g.shift_inv[2], g.G[2]);
      }
      const size_t gi = (size_t)o0 + (size_t)g.G[0] * ((size_t)o1 + (size_t)g.G[1] * o2);
#pragma unroll
      for (int c = 0; c < CB; c++)
        if (coil0 + c < a.n_coils) a.grid[(size_t)(coil0 + c) * g.Ng + gi] = make_float2(ax[p][c], ay[p][c]);
    }
  }
}

// GEN: also compiled with the generic body, for geometries beyond kCovMax.
// The generic body's registers would otherwise size every fold: 2-D with 8
// coils takes 63 VGPRs without it and 116 with it.
template <int DIM, int CB, bool GEN>
__device__ __forceinline__ void fold_body(const Geom& g, const SpreadArgs& a, uint32_t s, int coil_block) {
  if constexpr (GEN) {
    if (a.cover.max_cov > kCovMax) {
      fold_body_generic<DIM, CB>(g, a, s, coil_block);
      return;
    }
  }
  fold_body_cov<DIM, CB>(g, a, s, coil_block);
}

bool fold_generic(const Cover& cover) { return cover.max_cov > kCovMax; }

template <int DIM, int CB, bool GEN>
__global__ void __launch_bounds__(kFoldThreads) fold_kernel(Geom g, SpreadArgs a) {
  fold_body<DIM, CB, GEN>(g, a, blockIdx.x, blockIdx.y);
}

void build_cover(const Geom& g, std::vector<uint32_t>* words, size_t base[6], int max_len[3],
                 int* max_cov) {
  std::vector<uint32_t>& w = *words;
  w.clear();
  std::vector<std::vector<uint32_t>> lists[3];
  *max_cov = 1;
  for (int a = 0; a < 3; a++) {
    const int G = g.G[a], S = g.S[a];
    lists[a].assign((size_t)S, {});
    if (a >= g.dim) {
      lists[a][0].push_back(0u);  // 2-D: the single z plane
      continue;
    }
    for (int s = 0; s < S; s++) {
      const int lo = s * g.sw, hi = std::min(s * g.sw + g.sw, G);
      for (int n = 0; n < S; n++) {
        const int first = n * g.sw + g.sw / 2 - g.off;  // sector center - off
        bool meets = false;
        for (int i = 0; i < g.P && !meets; i++) {
          const int q = ((first + i) % G + G) % G;
          meets = q >= lo && q < hi;
        }
        if (meets) lists[a][(size_t)s].push_back((uint32_t)n);
      }
      // (candidate, periodic image) pairs per coordinate of this sector
      for (int q = lo; q < hi; q++) {
        int count = 0;
        for (uint32_t n : lists[a][(size_t)s]) {
          const int first = (int)n * g.sw + g.sw / 2 - g.off;
          for (int i = (((q - first) % G) + G) % G; i < g.P; i += G) count++;
        }
        *max_cov = std::max(*max_cov, count);
      }
    }
  }
  for (int a = 0; a < 3; a++) {
    base[a] = w.size();
    max_len[a] = 0;
    uint32_t acc = 0;
    for (const auto& l : lists[a]) {
      w.push_back(acc);
      acc += (uint32_t)l.size();
      max_len[a] = std::max(max_len[a], (int)l.size());
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
// Interp taps of one (sample, coil): ny x nz rows (wave-uniform trip
// counts) of TX words each, read in one batch per row (taps past the
// sample's range have weight 0; the row window stays inside the sub-grid
// and its zeroed slack). Weight ((wz*wy)*wx) as the reference forms it.
template <int DIM, int NX, int CB, int TX>
__device__ inline void interp_rows(const float2* sub, int P, int Pmax, int bx, int by, int bz, int ny,
                                   int nz, const float* w, int c, float* accx, float* accy) {
  float ax = 0.0f, ay = 0.0f;
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
      float2 v[TX];
#pragma unroll
      for (int tx = 0; tx < TX; tx++) v[tx] = row[tx * CB];
#pragma unroll
      for (int tx = 0; tx < TX; tx++) {
        const float val = wzy * w[1 + tx];
        ax += v[tx].x * val;
        ay += v[tx].y * val;
      }
    }
  }
  *accx = ax;
  *accy = ay;
}

// interp_rows with T rows per axis known at compile time (no branch per
// row, so the LDS reads of every row can be in flight together). Callers
// use it when no lane of the wave has more than T taps on any axis; the
// same operations in the same order as interp_rows with those trips.
template <int DIM, int NX, int CB, int T>
__device__ inline void interp_rows_fixed(const float2* sub, int P, int Pmax, int bx, int by, int bz,
                                         const float* w, int c, float* accx, float* accy) {
  float ax = 0.0f, ay = 0.0f;
#pragma unroll
  for (int tz = 0; tz < (DIM == 3 ? T : 1); tz++) {
    const int iz = DIM == 3 ? min(bz + tz, Pmax) : 0;
    const float wz = DIM == 3 ? w[1 + 2 * NX + tz] : 1.0f;
#pragma unroll
    for (int ty = 0; ty < T; ty++) {
      const int iy = min(by + ty, Pmax);
      const float wzy = DIM == 3 ? wz * w[1 + NX + ty] : w[1 + NX + ty];
      const float2* row = sub + ((iz * P + iy) * P + bx) * CB + c;
      float2 v[T];
#pragma unroll
      for (int tx = 0; tx < T; tx++) v[tx] = row[tx * CB];
#pragma unroll
      for (int tx = 0; tx < T; tx++) {
        const float val = wzy * w[1 + tx];
        ax += v[tx].x * val;
        ay += v[tx].y * val;
      }
    }
  }
  *accx = ax;
  *accy = ay;
}

template <int DIM, int NX, int CB, int NT = kThreads>
__device__ inline void interp_body(const Geom& g, const InterpArgs& a, uint32_t ci, int coil_block,
                                   int R) {
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

  const Chunk ch = a.work.chunks[ci];
  const int n = (int)(ch.end - ch.begin);
  const int coil0 = coil_block * CB;
  const int tid = threadIdx.x;
  int c3[3];
  load_center<DIM>(g, ch.sector, c3);

  // grid offset of padded index i along axis d (periodic wrap, FORWARD
  // fftshift folded in): the padded point (x, y, z) reads grid element
  // gix[0][x] + gix[1][y] + gix[2][z]
  __shared__ int gix[3][256];
  // round 0's coordinates first: their loads overlap the table set-up
  // (p0 then carries each next round's coordinates, loaded a round ahead)
  float p0[3] = {0.0f, 0.0f, 0.0f};
  if (tid < min(R, n)) load_coord<DIM>(a.crds, g.M, ch.begin + tid, p0);
  copy_to_lds<NT, 4>(lut, a.lut, g.K);
  fill_kpos<DIM, NT>(g, c3, kp, KPS);
  for (int e = tid; e < DIM * P; e += NT) {
    const int d = e >= 2 * P ? 2 : (e >= P ? 1 : 0);
    const int i = e - d * P;
    const int q = c3[d] - g.off + i + (a.fold_shift ? g.shift_fwd[d] : 0);
    const int stride = d == 0 ? 1 : (d == 1 ? g.G[0] : g.G[0] * g.G[1]);
    gix[d][i] = wrap(q, g.G[d]) * stride;
  }
  if (DIM == 2 && tid == 0) gix[2][0] = 0;
  __syncthreads();
  if (tid < min(R, n)) make_record<DIM, NX>(g, lut, kp, KPS, c3, p0, rec + tid * RW);
  // Stage the padded sector, element e = pt*CB + coil (the coil of thread
  // tid is tid % CB in every iteration); zero the overrun slack. Four
  // loads per thread in flight before their LDS writes.
  {
    constexpr int U = 4;
    const int cc = tid % CB;
    const bool coil_in = coil0 + cc < a.n_coils;
    const float2* gbase = a.grid + (size_t)(coil_in ? coil0 + cc : 0) * g.Ng;
    for (int e0 = tid; e0 < L.sub_n; e0 += U * NT) {
      float2 v[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int pt = (e0 + u * NT) / CB;
        v[u] = make_float2(0.0f, 0.0f);
        if (pt < PD && coil_in) {
          const int yz = (int)udiv((uint32_t)pt, g.mP);
          const int x = pt - yz * P;
          const int z = DIM == 3 ? (int)udiv((uint32_t)pt, g.mP2) : 0;
          const int y = yz - z * P;
          v[u] = gbase[(uint32_t)(gix[0][x] + gix[1][y] + gix[2][z])];
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++)
        if (e0 + u * NT < L.sub_n) sub[e0 + u * NT] = v[u];
    }
  }

  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int slot = lane / CB;
  const int c = lane % CB;
  const int cg = coil0 + c;

  for (int r0 = 0; r0 < n; r0 += R) {
    const int rn = min(R, n - r0);
    if (r0 > 0) {
      __syncthreads();  // previous round consumed
      // sample tid of this round (R <= NT); its coordinates were loaded
      // during the previous round's taps
      if (tid < rn) make_record<DIM, NX>(g, lut, kp, KPS, c3, p0, rec + tid * RW);
    }
    __syncthreads();  // sub-grid staged, the round's records ready
    // the next round's coordinates, in flight during this round's taps
    if (r0 + R + tid < n) load_coord<DIM>(a.crds, g.M, ch.begin + r0 + R + tid, p0);

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
      const int ny = wave_trip(packed, 1, g.kw);
      const int nz = DIM == 3 ? wave_trip(packed, 2, g.kw) : 1;
      float accx = 0.0f, accy = 0.0f;
      // x taps: NX, or 3 when NX is 4 and no lane of the wave has a fourth
      // (wave-uniform; one branch per sample, the rows unrolled); at kernel
      // width 3 with three taps on every axis the rows run without a branch
      // per row (narrower kernels keep their shorter trips)
      // the store's density weight and destination, loaded before the taps
      // so that their latency overlaps them
      const uint32_t s = ch.begin + r0 + ls;
      const float dn = a.dens ? a.dens[s] : 1.0f;
      const uint32_t dst = a.scatter ? a.scatter[s] : s;
      const int nxw = wave_trip(packed, 0, g.kw);
      if (NX == 4 && g.kw == 3 && nxw <= 3 && ny <= 3 && nz <= 3)
        interp_rows_fixed<DIM, NX, CB, 3>(sub, P, Pmax, bx, by, bz, w, c, &accx, &accy);
      else if (NX == 4 && nxw <= 3)
        interp_rows<DIM, NX, CB, (NX == 4 ? 3 : NX)>(sub, P, Pmax, bx, by, bz, ny, nz, w, c, &accx, &accy);
      else
        interp_rows<DIM, NX, CB, NX>(sub, P, Pmax, bx, by, bz, ny, nz, w, c, &accx, &accy);
      float rx = accx * a.scale, ry = accy * a.scale;
      if (a.dens) {
        const float sq = sqrtf(dn);
        rx = rx * sq;
        ry = ry * sq;
      }
      a.out[(size_t)cg * a.coil_stride + dst] = make_float2(rx, ry);
    }
  }
}

template <int DIM, int NX, int CB>
__global__ void __launch_bounds__(kThreads) interp_kernel(Geom g, InterpArgs a, int R) {
  const uint32_t ci = work_chunk(a.work, blockIdx.x);
  if (ci == ~0u) return;
  interp_body<DIM, NX, CB>(g, a, ci, blockIdx.y, R);
}

// ---------------------------------------------------------------------------
// Interp with tap records in registers (kernel width KW = 3), the forward
// twin of spread_lane_body: every thread forms the record of the samples it
// owns, the lanes of a quad share them with DPP moves, and each lane sums
// the KW^d taps of one (sample, coil) from the staged padded sector. The
// samples of a quad are consecutive (the scattered output stores of a wave
// cover runs of neighbouring samples). No LDS round trip of records and one
// barrier after the staging.
// ---------------------------------------------------------------------------
template <int DIM, int T, int CB>
__device__ inline void lane_gather(const float2* sub, int P, const int* b, const float (*w)[4],
                                   int cc, float* ax, float* ay) {
  const float2* base = sub + ((b[2] * P + b[1]) * P + b[0]) * CB + cc;
  float sx = 0.0f, sy = 0.0f;
#pragma unroll
  for (int tz = 0; tz < (DIM == 3 ? T : 1); tz++) {
    const float wz = DIM == 3 ? w[2][tz] : 1.0f;
#pragma unroll
    for (int ty = 0; ty < T; ty++) {
      const float wzy = DIM == 3 ? wz * w[1][ty] : w[1][ty];
      const float2* row = base + (tz * P + ty) * P * CB;
      float2 v[T];
#pragma unroll
      for (int tx = 0; tx < T; tx++) v[tx] = row[tx * CB];
#pragma unroll
      for (int tx = 0; tx < T; tx++) {
        const float val = wzy * w[0][tx];
        sx += v[tx].x * val;
        sy += v[tx].y * val;
      }
    }
  }
  *ax = sx;
  *ay = sy;
}

// LUTG as in spread_lane_body.
template <int DIM, int KW, int CB, int NT = kThreads, bool LUTG = false>
__device__ inline void interp_lane_body(const Geom& g, const InterpArgs& a, uint32_t ci,
                                        int coil_block) {
  constexpr int NX = KW + 1;
  constexpr int RND = 2;
  constexpr int PASSES = CB;
  static_assert(CB == 1 || CB == 2 || CB == 4, "a sample's coils share one quad");

  extern __shared__ float4 smem4[];
  char* smem = reinterpret_cast<char*>(smem4);
  const LaneLayout L(g, KW, CB, !LUTG);
  float* lut = reinterpret_cast<float*>(smem);
  float* kp = reinterpret_cast<float*>(smem + L.kp_off);
  float2* sub = reinterpret_cast<float2*>(smem + L.sub_off);
  const int P = g.P;
  const int Pmax = g.Pmax;
  const int KPS = P + NX;
  const int PD = DIM == 3 ? P * P * P : P * P;

  const Chunk ch = a.work.chunks[ci];
  const int n = (int)(ch.end - ch.begin);
  const int coil0 = coil_block * CB;
  const int tid = threadIdx.x;
  const int qc = (tid & 3) % CB;
  const bool coil_ok = coil0 + qc < a.n_coils;
  int c3[3];
  load_center<DIM>(g, ch.sector, c3);

  // owned samples: j = r * NT + tid (a quad owns four consecutive samples)
  float p[RND][3];
  bool live[RND];
#pragma unroll
  for (int r = 0; r < RND; r++) {
    const int j = r * NT + tid;
    live[r] = j < n;
    p[r][0] = p[r][1] = p[r][2] = 0.0f;
    if (live[r]) load_coord<DIM>(a.crds, g.M, ch.begin + (uint32_t)j, p[r]);
  }

  if (!LUTG) copy_to_lds<NT, 4>(lut, a.lut, g.K);
  const float* wlut = LUTG ? a.lut : lut;
  fill_kpos<DIM, NT>(g, c3, kp, KPS);
  // Stage the padded sector as [point][coil] (periodic wrap and the FORWARD
  // fftshift folded into the grid index), zeros in the slack past it (its
  // words meet only zero weights, but must not hold NaNs). Four loads per
  // thread in flight before their LDS writes.
  {
    constexpr int U = 4;
    const int cc = tid % CB;
    const bool cin = coil0 + cc < a.n_coils;
    const float2* gbase = a.grid + (size_t)(cin ? coil0 + cc : 0) * g.Ng;
    int q0[3];
#pragma unroll
    for (int d = 0; d < 3; d++)
      q0[d] = d < DIM ? c3[d] - g.off + (a.fold_shift ? g.shift_fwd[d] : 0) : 0;
    for (int e0 = tid; e0 < L.sub_n; e0 += U * NT) {
      float2 v[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int pt = (e0 + u * NT) / CB;
        v[u] = make_float2(0.0f, 0.0f);
        if (pt < PD && cin) {
          const int yz = (int)udiv((uint32_t)pt, g.mP);
          const int x = pt - yz * P;
          const int z = DIM == 3 ? (int)udiv((uint32_t)pt, g.mP2) : 0;
          const int y = yz - z * P;
          const int gx = wrap(q0[0] + x, g.G[0]);
          const int gy = wrap(q0[1] + y, g.G[1]);
          const int gz = DIM == 3 ? wrap(q0[2] + z, g.G[2]) : 0;
          v[u] = gbase[(size_t)gx + (size_t)g.G[0] * ((size_t)gy + (size_t)g.G[1] * gz)];
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++)
        if (e0 + u * NT < L.sub_n) sub[e0 + u * NT] = v[u];
    }
  }
  __syncthreads();

#pragma unroll
  for (int r = 0; r < RND; r++) {
    if (r * NT >= n) break;  // workgroup-uniform
    if (r * NT + (tid & ~63) >= n) continue;  // wave-uniform: this wave owns none
    uint32_t bp = 0;
    float w[3][4];
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int t = 0; t < 4; t++) w[d][t] = 0.0f;
    bool wide = false;
    if (live[r]) {
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        const float u = map_kspace_to_grid(p[r][d], g.G[d], c3[d], g.off);
        int lo, hi;
        tap_range(u, Pmax, g.kradius, &lo, &hi);
        bp |= (uint32_t)lo << (8 * d);
        wide = wide || hi - lo + 1 > KW;
#pragma unroll
        for (int t = 0; t < NX; t++) {
          float wt = 0.0f;
          if (lo + t <= hi && axis_weight_kp(g, wlut, d, kp[d * KPS + lo + t], p[r][d], &wt))
            w[d][t] = wt;
        }
      }
    }
    const bool any_wide = __ballot(wide) != 0ull;
    const uint32_t lv = live[r] ? 1u : 0u;
#pragma unroll
    for (int ps = 0; ps < PASSES; ps++) {
      const uint32_t bo = quad_pass_rt<CB>(bp, ps);
      const bool lo_live = quad_pass_rt<CB>(lv, ps) != 0u;
      float wo[3][4];
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int t = 0; t < 4; t++)
          wo[d][t] = (d < DIM && t < (any_wide ? NX : KW))
                         ? __uint_as_float(quad_pass_rt<CB>(__float_as_uint(w[d][t]), ps))
                         : 0.0f;
      // the worked sample: owner lane (tid & ~3) + ps * (4 / CB) + qi / CB
      const int jo = r * NT + (tid & ~3) + ps * (4 / CB) + (tid & 3) / CB;
      if (!lo_live || !coil_ok) continue;
      const int b[3] = {(int)(bo & 255u), (int)((bo >> 8) & 255u), (int)((bo >> 16) & 255u)};
      // the store's density weight and destination, loaded before the taps
      const uint32_t s = ch.begin + (uint32_t)jo;
      const float dn = a.dens ? a.dens[s] : 1.0f;
      const uint32_t dst = a.scatter ? a.scatter[s] : s;
      float ax, ay;
      if (any_wide)
        lane_gather<DIM, NX, CB>(sub, P, b, wo, qc, &ax, &ay);
      else
        lane_gather<DIM, KW, CB>(sub, P, b, wo, qc, &ax, &ay);
      float rx = ax * a.scale, ry = ay * a.scale;
      if (a.dens) {
        const float sq = sqrtf(dn);
        rx = rx * sq;
        ry = ry * sq;
      }
      a.out[(size_t)(coil0 + qc) * a.coil_stride + dst] = make_float2(rx, ry);
    }
  }
}

template <int DIM, int KW, int CB>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(6))) interp_lane_kernel(Geom g, InterpArgs a) {
  const uint32_t ci = work_chunk(a.work, blockIdx.x);
  if (ci == ~0u) return;
  interp_lane_body<DIM, KW, CB>(g, a, ci, blockIdx.y);
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
// at most one per thread of a `threads`-thread workgroup and at most the
// chunk; multiple of 16.
template <typename Layout>
static int round_size(const Geom& g, int cb, int chunk, size_t limit, int threads = kThreads) {
  const int nx = nx_of(g);
  const size_t fixed = Layout(g, nx, cb, 0).bytes;
  const size_t per = Layout(g, nx, cb, 1).bytes - fixed;
  if (fixed + 16 * per > limit) return 0;
  size_t r = (limit - fixed) / per;
  r = r > (size_t)threads ? (size_t)threads : r;
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

// Lane-per-sample spread: kernel width 3 with a padded sector of at least
// kw+1 points (the slack bound), chunks of at most two samples per thread.
static bool lane_spread_ok(const Geom& g, int chunk) {
  static int mode = -1;  // GNUFFT_SPREAD=round forces the round-based kernel
  if (mode < 0) {
    const char* e = getenv("GNUFFT_SPREAD");
    mode = e && strcmp(e, "round") == 0 ? 0 : 1;
  }
  return mode == 1 && g.kw == 3 && g.P >= 4 && chunk <= 2 * kThreads;
}

// Lane interp: 2-D only by default (at 128^3 osf 1.5 the 3-D lane interp
// takes 0.81 ms against 0.65 ms for the round-based kernel; 2-D batches are
// faster with it). GNUFFT_INTERP=round / =lane override, for tuning.
static bool lane_interp_ok(const Geom& g, int chunk) {
  static int mode = -1;
  if (mode < 0) {
    const char* e = getenv("GNUFFT_INTERP");
    mode = e && strcmp(e, "round") == 0 ? 0 : (e && strcmp(e, "lane") == 0 ? 2 : 1);
  }
  const bool shape_ok = g.kw == 3 && g.P >= 4 && chunk <= 2 * kThreads;
  return shape_ok && (mode == 2 || (mode == 1 && g.dim == 2));
}

template <int DIM, int CB>
static hipError_t spread_lane_launch(const Geom& g, const SpreadArgs& a, hipStream_t s) {
  const uint32_t n_blocks = work_blocks(a.work);
  const int coil_blocks = (a.n_coils + CB - 1) / CB;
  const size_t lds = LaneLayout(g, 3, CB).bytes;
  hipError_t e = allow_lds(spread_lane_kernel<DIM, 3, CB>, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((spread_lane_kernel<DIM, 3, CB>), dim3(n_blocks, coil_blocks),
                     dim3(kThreads), lds, s, g, a);
  return hipGetLastError();
}

// Largest coil block (not above the coil count) whose footprint lets
// `per_cu` workgroups share a CU; 0 if none.
static int lane_cb(const Geom& g, int n_coils, int per_cu) {
  const size_t limit = lds_limit() / (size_t)per_cu - 1024;
  for (int cb = 4; cb >= 1; cb /= 2) {
    if (cb > 1 && cb / 2 >= n_coils) continue;
    if (LaneLayout(g, 3, cb).bytes <= limit) return cb;
  }
  return 0;
}

// Workgroups per CU the lane kernels' coil block is sized for (default 3:
// at 128^3 osf 1.5 three workgroups of 4 coils take the spread kernel
// 1.25 -> 0.72 ms against four of 2; GNUFFT_LANE_PER_CU overrides, for
// tuning).
static int lane_per_cu() {
  static int v = 0;
  if (v == 0) {
    const char* e = getenv("GNUFFT_LANE_PER_CU");
    v = e ? atoi(e) : 3;
    if (v < 1 || v > 8) v = 3;
  }
  return v;
}

template <int DIM>
static hipError_t spread_lane(const Geom& g, SpreadArgs a, hipStream_t s, int* cb_out) {
  int cb = 0;
  const int first = a.lane_per_cu > 0 ? a.lane_per_cu : lane_per_cu();
  for (int per = first; per >= 1 && cb == 0; per--) cb = lane_cb(g, a.n_coils, per);
  a.sub_cb = cb;
  *cb_out = cb;
  switch (cb) {
    case 4: return spread_lane_launch<DIM, 4>(g, a, s);
    case 2: return spread_lane_launch<DIM, 2>(g, a, s);
    case 1: return spread_lane_launch<DIM, 1>(g, a, s);
    default: return hipErrorInvalidValue;
  }
}

template <int DIM, int NX>
static hipError_t spread_nx(const Geom& g, SpreadArgs a, int chunk, hipStream_t s, int* cb_out) {
  // Prefer a footprint that lets three workgroups share a CU (the spread
  // waits on memory most of the time; at 128^3 osf 1.5 three workgroups of
  // 4 coils beat two of 8 by 14 %), then two, then one.
  int R = 0;
  int cb = pick_cb<SpreadLayout>(g, a.n_coils, 8, chunk, lds_limit() / 3 - 1024, &R);
  if (cb == 0) cb = pick_cb<SpreadLayout>(g, a.n_coils, 8, chunk, lds_limit() / 2 - 1024, &R);
  if (cb == 0) cb = pick_cb<SpreadLayout>(g, a.n_coils, 8, chunk, lds_limit() - 1024, &R);
  a.sub_cb = cb;
  *cb_out = cb;
  switch (cb) {
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
  // three workgroups per CU if they fit (as for the spread), then two, then
  // one; room for the static tables (grid offsets, 3 KB) beside the
  // dynamic LDS
  int cb = pick_cb<InterpLayout>(g, a.n_coils, 8, chunk, lds_limit() / 3 - kInterpStaticLds, &R);
  if (cb == 0) cb = pick_cb<InterpLayout>(g, a.n_coils, 8, chunk, lds_limit() / 2 - kInterpStaticLds, &R);
  if (cb == 0) cb = pick_cb<InterpLayout>(g, a.n_coils, 8, chunk, lds_limit() - kInterpStaticLds, &R);
  switch (cb) {
    case 8: return interp_launch<DIM, NX, 8>(g, a, R, s);
    case 4: return interp_launch<DIM, NX, 4>(g, a, R, s);
    case 2: return interp_launch<DIM, NX, 2>(g, a, R, s);
    case 1: return interp_launch<DIM, NX, 1>(g, a, R, s);
    default: return hipErrorInvalidValue;
  }
}

// Fold workgroup size: enough threads for fold_pts home points each and one
// per candidate sector while the item list is built; 0 if too many candidates.
int fold_threads(const Geom& g, const Cover& cover) {
  const int hp = g.dim == 3 ? g.sw * g.sw * g.sw : g.sw * g.sw;
  const int per = (hp + fold_pts(g.dim) - 1) / fold_pts(g.dim);
  int threads = per >= kFoldThreads ? kFoldThreads : ((per + 63) / 64) * 64;
  int cands = 1;
  for (int d = 0; d < g.dim; d++) cands *= cover.max_len[d] * ((g.P + g.G[d] - 1) / g.G[d]);
  if (cands > kFoldCands) return 0;
  if (cands > threads) threads = ((cands + 63) / 64) * 64;
  return threads;
}

template <int DIM, bool GEN>
static void fold_launch(int fcb, dim3 grid, int threads, const Geom& g, const SpreadArgs& a,
                        hipStream_t s) {
  switch (fcb) {
    case 8: hipLaunchKernelGGL((fold_kernel<DIM, 8, GEN>), grid, dim3(threads), 0, s, g, a); break;
    case 4: hipLaunchKernelGGL((fold_kernel<DIM, 4, GEN>), grid, dim3(threads), 0, s, g, a); break;
    case 2: hipLaunchKernelGGL((fold_kernel<DIM, 2, GEN>), grid, dim3(threads), 0, s, g, a); break;
    default: hipLaunchKernelGGL((fold_kernel<DIM, 1, GEN>), grid, dim3(threads), 0, s, g, a); break;
  }
}

hipError_t launch_spread(const Geom& g, const SpreadArgs& args, int chunk, hipStream_t s) {
  if (args.n_coils == 0) return hipSuccess;
  if (g.kw > 7 || g.P > 255) return hipErrorInvalidValue;
  hipError_t e = hipSuccess;
  SpreadArgs a = args;
  a.sub_cb = 1;  // set below to the spread's coil block (the sub-grid layout)
  int cb = 1;
  if (a.work.n_chunks > 0 && lane_spread_ok(g, chunk)) {
    e = g.dim == 3 ? spread_lane<3>(g, a, s, &cb) : spread_lane<2>(g, a, s, &cb);
    if (e != hipSuccess) return e;
  } else if (a.work.n_chunks > 0) {
    if (g.dim == 3)
      e = nx_of(g) == 4 ? spread_nx<3, 4>(g, a, chunk, s, &cb) : spread_nx<3, 8>(g, a, chunk, s, &cb);
    else
      e = nx_of(g) == 4 ? spread_nx<2, 4>(g, a, chunk, s, &cb) : spread_nx<2, 8>(g, a, chunk, s, &cb);
    if (e != hipSuccess) return e;
  }
  a.sub_cb = cb;
  const int PDn = g.dim == 3 ? g.P * g.P * g.P : g.P * g.P;
  if (a.n_reduce > 0) {
    const size_t kstride = (size_t)a.n_coils * PDn;
    const dim3 rgrid((unsigned)((kstride + 255) / 256), a.n_reduce);
    hipLaunchKernelGGL(reduce_chunks_kernel, rgrid, dim3(256), 0, s, a.reduce_tasks, a.scratch,
                       kstride);
  }
  // the fold writes every grid point, zeros where no sample contributes
  const uint32_t S = (uint32_t)g.S[0] * g.S[1] * g.S[2];
  const int threads = fold_threads(g, a.cover);
  if (threads == 0) return hipErrorInvalidValue;
  const int fcb = a.n_coils >= 8 ? 8 : a.n_coils >= 4 ? 4 : a.n_coils >= 2 ? 2 : 1;
  const dim3 grid(S, (a.n_coils + fcb - 1) / fcb);
  const bool gen = fold_generic(a.cover);
  if (g.dim == 3)
    gen ? fold_launch<3, true>(fcb, grid, threads, g, a, s) : fold_launch<3, false>(fcb, grid, threads, g, a, s);
  else
    gen ? fold_launch<2, true>(fcb, grid, threads, g, a, s) : fold_launch<2, false>(fcb, grid, threads, g, a, s);
  return hipGetLastError();
}

template <int DIM, int CB>
static hipError_t interp_lane_launch(const Geom& g, const InterpArgs& a, hipStream_t s) {
  const uint32_t n_blocks = work_blocks(a.work);
  const int coil_blocks = (a.n_coils + CB - 1) / CB;
  const size_t lds = LaneLayout(g, 3, CB).bytes;
  hipError_t e = allow_lds(interp_lane_kernel<DIM, 3, CB>, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((interp_lane_kernel<DIM, 3, CB>), dim3(n_blocks, coil_blocks), dim3(kThreads),
                     lds, s, g, a);
  return hipGetLastError();
}

template <int DIM>
static hipError_t interp_lane(const Geom& g, const InterpArgs& a, hipStream_t s) {
  int cb = 0;
  const int first = a.lane_per_cu > 0 ? a.lane_per_cu : lane_per_cu();
  for (int per = first; per >= 1 && cb == 0; per--) cb = lane_cb(g, a.n_coils, per);
  switch (cb) {
    case 4: return interp_lane_launch<DIM, 4>(g, a, s);
    case 2: return interp_lane_launch<DIM, 2>(g, a, s);
    case 1: return interp_lane_launch<DIM, 1>(g, a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_interp(const Geom& g, const InterpArgs& a, int chunk, hipStream_t s) {
  if (a.work.n_chunks == 0 || a.n_coils == 0) return hipSuccess;
  if (g.kw > 7 || g.P > 255) return hipErrorInvalidValue;
  if (lane_interp_ok(g, chunk)) return g.dim == 3 ? interp_lane<3>(g, a, s) : interp_lane<2>(g, a, s);
  InterpArgs b = a;
  if (a.work_runs.n_chunks > 0) b.work = a.work_runs;
  if (g.dim == 3) return nx_of(g) == 4 ? interp_nx<3, 4>(g, b, chunk, s) : interp_nx<3, 8>(g, b, chunk, s);
  return nx_of(g) == 4 ? interp_nx<2, 4>(g, b, chunk, s) : interp_nx<2, 8>(g, b, chunk, s);
}

// ---------------------------------------------------------------------------
// Batched launches: one launch per stage for a batch of operators
// ---------------------------------------------------------------------------
// Batch rounds are sized for kBatchPerCU workgroups of kBatchThreads per
// CU (eight waves per SIMD); the 2-D 4-tap spread fits that in 64 VGPRs
// without spilling, so it is compiled for it (unconstrained it takes 70).
template <int DIM, int NX>
__global__ void __launch_bounds__(kBatchThreads, DIM == 2 && NX == 4 ? 8 : 1)
    spread_batch_kernel(const BatchOp* __restrict__ ops, BatchList l) {
  uint4 it;
  if (!batch_item(l, blockIdx.x, &it)) return;
  const BatchOp& op = ops[it.x];
  spread_body<DIM, NX, kBatchCB, kBatchThreads>(op.g, op.s, it.y, (int)it.z, op.R_spread);
}

template <int DIM, int NX>
__global__ void __launch_bounds__(kBatchThreads) interp_batch_kernel(const BatchOp* __restrict__ ops,
                                                                     BatchList l) {
  uint4 it;
  if (!batch_item(l, blockIdx.x, &it)) return;
  const BatchOp& op = ops[it.x];
  interp_body<DIM, NX, kBatchCB, kBatchThreads>(op.g, op.i, it.y, (int)it.z, op.R_interp);
}

template <int DIM, bool GEN>
__global__ void __launch_bounds__(kFoldThreads) fold_batch_kernel(const BatchOp* __restrict__ ops,
                                                                  BatchList l) {
  uint4 it;
  if (!batch_item(l, blockIdx.x, &it)) return;
  const BatchOp& op = ops[it.x];
  fold_body<DIM, kBatchCB, GEN>(op.g, op.s, it.y, (int)it.z);
}

size_t lds_budget() { return lds_limit(); }

size_t spread_lds_bytes(const Geom& g, int cb, int R) {
  return SpreadLayout(g, nx_of(g), cb, R).bytes;
}

size_t interp_lds_bytes(const Geom& g, int cb, int R) {
  return InterpLayout(g, nx_of(g), cb, R).bytes;
}

int spread_round(const Geom& g, int cb, int chunk, size_t limit) {
  return round_size<SpreadLayout>(g, cb, chunk, limit, kBatchThreads);
}

int interp_round(const Geom& g, int cb, int chunk, size_t limit) {
  return round_size<InterpLayout>(g, cb, chunk, limit, kBatchThreads);
}

template <typename K>
static hipError_t launch_items(K kernel, const BatchOp* ops, const BatchList& l, int threads,
                               size_t lds, hipStream_t s) {
  if (l.n == 0) return hipSuccess;
  hipError_t e = allow_lds(kernel, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kernel, dim3(batch_blocks(l)), dim3(threads), lds, s, ops, l);
  return hipGetLastError();
}

// task (operator, first chunk, chunk count, 0) per blockIdx.y
__global__ void __launch_bounds__(256) reduce_batch_kernel(const BatchOp* __restrict__ ops,
                                                           const uint4* __restrict__ tasks) {
  const uint4 t = tasks[blockIdx.y];
  const BatchOp& op = ops[t.x];
  const size_t PD = op.g.dim == 3 ? (size_t)op.g.P * op.g.P * op.g.P : (size_t)op.g.P * op.g.P;
  reduce_body(op.s.scratch, (size_t)op.s.n_coils * PD, t.y, t.z,
              (size_t)blockIdx.x * 256 + threadIdx.x);
}

// Batched lane kernels (kernel width 3 in every plan): coil blocks of
// kBatchLaneCB, one sample's coils per quad.
template <int DIM>
__global__ void __launch_bounds__(kBatchThreads) __attribute__((amdgpu_waves_per_eu(6)))
    spread_batch_lane_kernel(const BatchOp* __restrict__ ops, BatchList l) {
  uint4 it;
  if (!batch_item(l, blockIdx.x, &it)) return;
  const BatchOp& op = ops[it.x];
  spread_lane_body<DIM, 3, kBatchLaneCB, kBatchThreads>(op.g, op.s, it.y, (int)it.z);
}

template <int DIM>
__global__ void __launch_bounds__(kBatchThreads) __attribute__((amdgpu_waves_per_eu(6)))
    interp_batch_lane_kernel(const BatchOp* __restrict__ ops, BatchList l) {
  uint4 it;
  if (!batch_item(l, blockIdx.x, &it)) return;
  const BatchOp& op = ops[it.x];
  interp_lane_body<DIM, 3, kBatchLaneCB, kBatchThreads>(op.g, op.i, it.y, (int)it.z);
}

template <int DIM>
__global__ void __launch_bounds__(kBatchSmallThreads) __attribute__((amdgpu_waves_per_eu(6)))
    spread_batch_lane_small_kernel(const BatchOp* __restrict__ ops, BatchList l) {
  uint4 it;
  if (!batch_item(l, blockIdx.x, &it)) return;
  const BatchOp& op = ops[it.x];
  spread_lane_body<DIM, 3, kBatchLaneCB, kBatchSmallThreads, true>(op.g, op.s, it.y, (int)it.z);
}

template <int DIM>
__global__ void __launch_bounds__(kBatchSmallThreads) __attribute__((amdgpu_waves_per_eu(6)))
    interp_batch_lane_small_kernel(const BatchOp* __restrict__ ops, BatchList l) {
  uint4 it;
  if (!batch_item(l, blockIdx.x, &it)) return;
  const BatchOp& op = ops[it.x];
  interp_lane_body<DIM, 3, kBatchLaneCB, kBatchSmallThreads, true>(op.g, op.i, it.y, (int)it.z);
}

hipError_t launch_interp_batch_small(int dim, const BatchOp* ops, const BatchList& l, size_t lds,
                                     hipStream_t s) {
  return dim == 3 ? launch_items(interp_batch_lane_small_kernel<3>, ops, l, kBatchSmallThreads, lds, s)
                  : launch_items(interp_batch_lane_small_kernel<2>, ops, l, kBatchSmallThreads, lds, s);
}

hipError_t launch_spread_batch_small(int dim, const BatchOp* ops, const BatchList& l, size_t lds,
                                     hipStream_t s) {
  return dim == 3 ? launch_items(spread_batch_lane_small_kernel<3>, ops, l, kBatchSmallThreads, lds, s)
                  : launch_items(spread_batch_lane_small_kernel<2>, ops, l, kBatchSmallThreads, lds, s);
}

bool lane_batch_ok(const Geom& g) {
  return lane_spread_ok(g, 2 * kBatchThreads) && lane_interp_ok(g, 2 * kBatchThreads);
}

size_t lane_lds_bytes(const Geom& g, int cb, bool lut) { return LaneLayout(g, 3, cb, lut).bytes; }

hipError_t launch_spread_batch(int dim, int nx, const BatchOp* ops, const BatchList& l,
                               size_t lds, hipStream_t s) {
  if (nx == 3)
    return dim == 3 ? launch_items(spread_batch_lane_kernel<3>, ops, l, kBatchThreads, lds, s)
                    : launch_items(spread_batch_lane_kernel<2>, ops, l, kBatchThreads, lds, s);
  if (dim == 3)
    return nx == 4 ? launch_items(spread_batch_kernel<3, 4>, ops, l, kBatchThreads, lds, s)
                   : launch_items(spread_batch_kernel<3, 8>, ops, l, kBatchThreads, lds, s);
  return nx == 4 ? launch_items(spread_batch_kernel<2, 4>, ops, l, kBatchThreads, lds, s)
                 : launch_items(spread_batch_kernel<2, 8>, ops, l, kBatchThreads, lds, s);
}

hipError_t launch_reduce_batch(const BatchOp* ops, const BatchList& l, uint32_t max_kstride,
                               hipStream_t s) {
  if (l.n == 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_batch_kernel, dim3((max_kstride + 255) / 256, l.n), dim3(256), 0, s,
                     ops, l.items);
  return hipGetLastError();
}

hipError_t launch_interp_batch(int dim, int nx, const BatchOp* ops, const BatchList& l,
                               size_t lds, hipStream_t s) {
  if (nx == 3)
    return dim == 3 ? launch_items(interp_batch_lane_kernel<3>, ops, l, kBatchThreads, lds, s)
                    : launch_items(interp_batch_lane_kernel<2>, ops, l, kBatchThreads, lds, s);
  if (dim == 3)
    return nx == 4 ? launch_items(interp_batch_kernel<3, 4>, ops, l, kBatchThreads, lds, s)
                   : launch_items(interp_batch_kernel<3, 8>, ops, l, kBatchThreads, lds, s);
  return nx == 4 ? launch_items(interp_batch_kernel<2, 4>, ops, l, kBatchThreads, lds, s)
                 : launch_items(interp_batch_kernel<2, 8>, ops, l, kBatchThreads, lds, s);
}

hipError_t launch_fold_batch(int dim, const BatchOp* ops, const BatchList& l, int threads,
                             bool generic, hipStream_t s) {
  if (generic)
    return dim == 3 ? launch_items(fold_batch_kernel<3, true>, ops, l, threads, 0, s)
                    : launch_items(fold_batch_kernel<2, true>, ops, l, threads, 0, s);
  return dim == 3 ? launch_items(fold_batch_kernel<3, false>, ops, l, threads, 0, s)
                  : launch_items(fold_batch_kernel<2, false>, ops, l, threads, 0, s);
}


}  // namespace gnufft
