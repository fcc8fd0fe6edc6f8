// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of gpuNUFFT's sector-based NUFFT gridding semantics, used as
// the checker for the HIP product path. Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load this library; the product
// (gpunufft_amd) never links, loads or calls it.
//
// Every function cites the reference file:line it restates (paths relative to
// the reference tree). Arithmetic follows the reference exactly: DType is
// float, mixed float/double expressions keep the reference's promotions, and
// the library is compiled with -ffp-contract=off so no FMA changes a rounding
// decision (ceil/floor/round of grid positions and LUT indices).
//
// Pinning status of each function (how a plausible mistake is caught) is
// listed in oracle/README.md and checked by tests/test_oracle_*.py.

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <utility>
#include <vector>

typedef float DType;
typedef unsigned int IndType;

extern "C" {

// ---------------------------------------------------------------------------
// Kaiser-Bessel lookup table.
// ---------------------------------------------------------------------------

// Modified Bessel I0, Abramowitz-Stegun 9.8.1/9.8.2 polynomials, float
// arithmetic. CUDA/src/gpuNUFFT_utils.cpp:3-34
float o_i0(float x) {
  DType ax = std::fabs(x);
  DType ans;
  DType y;
  if (ax < (DType)3.75f) {
    y = x / 3.75f, y = y * y;
    ans = 1.0f +
          y * (3.5156229f +
               y * (3.0899424f +
                    y * (1.2067492f +
                         y * (0.2659732f + y * (0.360768e-1f + y * 0.45813e-2f)))));
  } else {
    y = 3.75f / ax;
    // std::exp / std::sqrt on a float argument resolve to the float overloads,
    // as in the reference (DType ax).
    ans = (std::exp(ax) / std::sqrt(ax)) *
          (0.39894228f +
           y * (0.1328592e-1f +
                y * (0.225319e-2f +
                     y * (-0.157565e-2f +
                          y * (0.916281e-2f +
                               y * (-0.2057706e-1f +
                                    y * (0.2635537e-1f +
                                         y * (-0.1647633e-1f + y * 0.392377e-2f))))))));
  }
  return ans;
}

// beta = pi*sqrt((kw/osr*(osr-0.5))^2-0.8). CUDA/inc/gpuNUFFT_utils.hpp:60-63
float o_beta(unsigned kw, float osr) {
  return (DType)M_PI *
         std::sqrt(std::pow((DType)kw / osr * (osr - (DType)0.5), (DType)2.0) - (DType)0.8);
}

// KB(r) = I0(beta*sqrt(1-r^2))/I0(beta). CUDA/inc/gpuNUFFT_utils.hpp:72-75
float o_kernel_value(float radius, unsigned kw, float osr) {
  return o_i0(o_beta(kw, osr) * std::sqrt((DType)1.0 - std::pow(radius, (DType)2.0))) /
         o_i0(o_beta(kw, osr));
}

// Nearest-entry LUT size. CUDA/src/gpuNUFFT_utils.cpp:47-52
long o_kernel_size_nearest(float osr, unsigned kw) {
  long kernel_osf = (long)(std::floor((DType)0.91 / (osr * 0.001f)));
  return (long)(kernel_osf * (DType)kw * 0.5);
}

// Linear-interpolation (texture) LUT size. CUDA/src/gpuNUFFT_utils.cpp:54-60
long o_kernel_size_linear(float osr, unsigned kw) {
  long kernel_osf = (long)(std::floor(std::sqrt(0.37 / ((osr * osr) * 0.0001f))));
  return (long)(kernel_osf * (DType)kw * 0.5);
}

// 1-D table over r^2. CUDA/src/gpuNUFFT_utils.cpp:74-107
void o_load1d_kernel(float* tab, long K, int kw, float osr) {
  for (long i = 1; i < K - 1; i++) {
    if (kw == 1) {
      tab[i] = 1.0f;
    } else {
      double rsqr = (double)std::sqrt(i / (double)(K - 1));
      tab[i] = (DType)o_kernel_value((DType)rsqr, kw, osr);
    }
  }
  tab[0] = 1.0f;
  tab[K - 1] = 0.0f;
}

// Separable 2-D table T2[i + j*K] = T[j]*T[i]. CUDA/src/gpuNUFFT_utils.cpp:109-124
void o_load2d_kernel(float* tab, long K, int kw, float osr) {
  o_load1d_kernel(tab, K, kw, osr);
  for (long j = 0; j < K; j++)
    for (long i = 0; i < K; i++) tab[i + j * K] = tab[j] * tab[i];
}

// ---------------------------------------------------------------------------
// Derived gridding constants (GpuNUFFTInfo).
// ---------------------------------------------------------------------------

typedef struct {
  int dim;               // 2 or 3
  unsigned img[3];       // image dims (z = 0 in 2-D)
  unsigned grid[3];      // oversampled grid dims (z = 0 in 2-D)
  unsigned grid_count;   // Gx*Gy*max(Gz,1)
  unsigned img_count;
  unsigned sector_width;
  unsigned sectors[3];   // sectors per dim (z = 1 in 2-D)
  unsigned sector_count;
  int kernel_width;
  long kernel_count;     // entries in the table the kernels index
  float kernel_radius;
  float radius_sqr;
  float radius_sqr_inv;
  float dist_multiplier;
  float aniso[3];
  int sector_pad_width;
  int sector_pad_max;
  int sector_dim;
  int sector_offset;
  float osf;
} o_info;

// CUDA/src/gpuNUFFT_operator.cpp:53-146 (initGpuNUFFTInfo) with the grid
// dims of gpuNUFFT_types.hpp:109-117 and sector counts of
// gpuNUFFT_operator_factory.cpp:23-27.
void o_init_info(o_info* gi, const unsigned img[3], float osf, unsigned kw, unsigned sw,
                 long kernel_count) {
  memset(gi, 0, sizeof(*gi));
  gi->dim = img[2] == 0 ? 2 : 3;
  for (int a = 0; a < 3; a++) {
    gi->img[a] = img[a];
    gi->grid[a] = (IndType)(img[a] * (DType)osf);
  }
  gi->grid_count = gi->grid[0] * gi->grid[1] * (gi->grid[2] == 0 ? 1 : gi->grid[2]);
  gi->img_count = img[0] * img[1] * (img[2] == 0 ? 1 : img[2]);
  gi->sector_width = sw;
  for (int a = 0; a < 3; a++) {
    IndType s = (IndType)std::ceil(static_cast<DType>(gi->grid[a]) / sw);
    gi->sectors[a] = s == 0 ? 1 : s;  // Dimensions::count treats 0 as 1
  }
  gi->sector_count = gi->sectors[0] * gi->sectors[1] * gi->sectors[2];
  gi->kernel_width = (int)kw;
  gi->kernel_count = kernel_count;
  unsigned max_grid = std::max(std::max(gi->grid[0], gi->grid[1]), gi->grid[2]);
  double kernel_radius = static_cast<double>(kw) / 2.0;
  double radius = kernel_radius / static_cast<double>(max_grid);
  double radius_sqr = radius * radius;
  double inv = 1.0 / radius_sqr;
  gi->dist_multiplier = (DType)((kernel_count - 1) * inv);
  gi->kernel_radius = (DType)kernel_radius;
  gi->radius_sqr = (DType)radius_sqr;
  gi->radius_sqr_inv = (DType)inv;
  for (int a = 0; a < 3; a++) gi->aniso[a] = (DType)gi->grid[a] / (DType)max_grid;
  gi->sector_pad_width = (int)sw + 2 * (int)std::floor((DType)kw / (DType)2.0);
  gi->sector_pad_max = gi->sector_pad_width - 1;
  gi->sector_dim = gi->sector_pad_width * gi->sector_pad_width *
                   (gi->dim == 3 ? gi->sector_pad_width : 1);
  gi->sector_offset = (int)std::floor((DType)gi->sector_pad_width / (DType)2.0);
  gi->osf = osf;
}

// ---------------------------------------------------------------------------
// Sector precompute.
// ---------------------------------------------------------------------------

// Sample -> sector id along one axis. CUDA/inc/precomp_utils.hpp:34-48.
// float product, double add, round half away from zero, double divide, floor,
// clamp to ceil((float)G/sw)-1 computed in float.
static IndType sector_mapping(DType coord, IndType gridDim, DType sectorWidth) {
  int sector = (int)std::floor(
      std::round(coord * (DType)(gridDim) + 0.5 * (DType)(gridDim)) / sectorWidth);
  if (sector >= (int)std::ceil((DType)(gridDim) / sectorWidth))
    sector = (int)(std::ceil((DType)(gridDim) / sectorWidth) - 1);
  if (sector < 0) sector = 0;
  return sector;
}

// Linear sector id per sample, trajectory SoA [x..., y..., (z...)].
// CUDA/src/gpuNUFFT_operator_factory.cpp:128-181 (CPU branch) and
// CUDA/src/gpu/precomp_kernels.cu:8-45; linearisation precomp_utils.hpp:142-153.
void o_assign_sectors(const o_info* gi, const float* traj, unsigned M, unsigned* out) {
  for (unsigned i = 0; i < M; i++) {
    IndType sx = sector_mapping(traj[i], gi->grid[0], (DType)gi->sector_width);
    IndType sy = sector_mapping(traj[i + M], gi->grid[1], (DType)gi->sector_width);
    IndType sz = 0;
    if (gi->dim == 3) sz = sector_mapping(traj[i + 2 * M], gi->grid[2], (DType)gi->sector_width);
    out[i] = sx + gi->sectors[0] * (sy + gi->sectors[1] * sz);
  }
}

// The reference's own ordering: std::sort of (index, sector) pairs compared by
// sector only (unstable). CUDA/inc/gpuNUFFT_types.hpp:33-49,
// CUDA/src/gpuNUFFT_operator_factory.cpp:56-72. libstdc++'s introsort is
// deterministic, so this reproduces the reference permutation exactly.
struct OPair : std::pair<IndType, IndType> {
  OPair(IndType a, IndType b) : std::pair<IndType, IndType>(a, b) {}
  bool operator<(const OPair& o) const { return this->second < o.second; }
  bool operator>(const OPair& o) const { return this->second > o.second; }
};

void o_sort_std(const unsigned* sectors, unsigned M, unsigned* idx_out, unsigned* sec_out) {
  std::vector<OPair> v;
  v.reserve(M);
  for (unsigned i = 0; i < M; i++) v.push_back(OPair(i, sectors[i]));
  std::sort(v.begin(), v.end());
  for (unsigned i = 0; i < M; i++) {
    idx_out[i] = v[i].first;
    sec_out[i] = v[i].second;
  }
}

// [0, n0, n0+n1, ...], length S+1. CUDA/src/gpuNUFFT_operator_factory.cpp:183-205
void o_sector_data_count(const unsigned* sorted_sec, unsigned M, unsigned S, unsigned* out) {
  unsigned cnt = 0;
  out[0] = 0;
  for (unsigned i = 0; i < S; i++) {
    while (cnt < M && i == sorted_sec[cnt]) cnt++;
    out[i + 1] = cnt;
  }
}

// Centers c = s*sw + floor(sw/2), interleaved per sector.
// CUDA/src/gpuNUFFT_operator_factory.cpp:207-270
void o_sector_centers(const o_info* gi, unsigned* out) {
  IndType sw = gi->sector_width;
  IndType half = (IndType)(int)std::floor(static_cast<DType>(sw) / (DType)2.0);
  if (gi->dim == 2) {
    for (IndType y = 0; y < gi->sectors[1]; y++)
      for (IndType x = 0; x < gi->sectors[0]; x++) {
        int index = (int)x + (int)gi->sectors[0] * (int)y;
        out[2 * index] = x * sw + half;
        out[2 * index + 1] = y * sw + half;
      }
  } else {
    for (IndType z = 0; z < gi->sectors[2]; z++)
      for (IndType y = 0; y < gi->sectors[1]; y++)
        for (IndType x = 0; x < gi->sectors[0]; x++) {
          int index = (int)x + (int)gi->sectors[0] * ((int)y + (int)gi->sectors[1] * (int)z);
          out[3 * index] = x * sw + half;
          out[3 * index + 1] = y * sw + half;
          out[3 * index + 2] = z * sw + half;
        }
  }
}

// Balanced processing order: sectors by count descending (std::sort with
// greater<>), empty sectors dropped, each sector split into 256-sample
// payloads. Returns the number of (sector, offset) pairs written to out
// (interleaved). CUDA/src/gpuNUFFT_operator_factory.cpp:74-126
unsigned o_processing_order(const unsigned* sdc, unsigned S, unsigned* out) {
  std::vector<OPair> cps;
  for (unsigned i = 0; i < S; i++) cps.push_back(OPair(i, sdc[i + 1] - sdc[i]));
  std::sort(cps.begin(), cps.end(), std::greater<OPair>());
  unsigned n = 0;
  for (unsigned i = 0; i < cps.size(); i++) {
    if (cps[i].second > 0) {
      out[2 * n] = cps[i].first;
      out[2 * n + 1] = 0;
      n++;
      if (cps[i].second > 256) {
        int remaining = (int)cps[i].second;
        IndType offset = 1;
        while ((remaining - 256) > 0) {
          remaining -= 256;
          out[2 * n] = cps[i].first;
          out[2 * n + 1] = (offset++) * 256;
          n++;
        }
      }
    } else {
      break;
    }
  }
  return n;
}

// ---------------------------------------------------------------------------
// Gridding helpers. CUDA/inc/gpuNUFFT_utils.hpp:89-104, 194-209, 268-278.
// ---------------------------------------------------------------------------

static void set_minmax(DType* x, int* mn, int* mx, int maximum, DType radius) {
  *mn = (int)std::ceil(*x - radius);
  *mx = (int)std::floor(*x + radius);
  if (*mn < 0) *mn = 0;
  if (*mx < 0) *mx = 0;
  if (*mx >= maximum) *mx = maximum;
  if (*mn >= (DType)maximum) *mn = (int)(maximum - 2 * radius);
}

static DType map_kspace_to_grid(DType pos, IndType gridDim, IndType center, int offset) {
  return (pos * (DType)gridDim) + ((DType)0.5 * ((DType)gridDim)) - (DType)center + (DType)offset;
}

static DType map_grid_to_kspace(int gridPos, IndType gridDim, IndType center, int offset) {
  return static_cast<DType>((DType)gridPos + (DType)center - (DType)offset) /
             static_cast<DType>((DType)gridDim) -
         (DType)0.5;
}

// Periodic wrap of a padded-sector coordinate into the grid. The reference
// wraps once (+-G, gpuNUFFT_utils.hpp:268-278); a modulo is identical for every
// in-range access and stays defined for grids smaller than the padded sector.
static int wrap_index(int coord, int center, int width, int offset) {
  int p = center - offset + coord;
  p %= width;
  if (p < 0) p += width;
  return p;
}

// Texture-mode 1-D linear lookup emulating a normalized-coordinate,
// clamp-addressed linear-filter fetch of the separable 2-D table: texel
// coordinate u*K-0.5, fractional weight in 9-bit fixed point with 8 fractional
// bits (NVIDIA CUDA Programming Guide, "Linear Filtering"), rounded to
// nearest. CUDA/inc/cuda_utils.cuh:27-36, CUDA/src/gpu/std_gpuNUFFT_kernels.cu:49-60.
static DType lerp_lookup(const float* tab, long K, DType u) {
  DType xb = u * (DType)K - 0.5f;
  DType fl = std::floor(xb);
  int i0 = (int)fl;
  DType a = xb - fl;
  a = std::nearbyint(a * 256.0f) / 256.0f;
  int i1 = i0 + 1;
  i0 = i0 < 0 ? 0 : (i0 > K - 1 ? (int)K - 1 : i0);
  i1 = i1 < 0 ? 0 : (i1 > K - 1 ? (int)K - 1 : i1);
  return (1.0f - a) * tab[i0] + a * tab[i1];
}

// Lookup modes: 0 = nearest entry of the 1-D table (CONST_LOOKUP),
// 1 = emulated bilinear texture lookup (TEXTURE2D_LOOKUP).
#define O_LOOKUP_NEAREST 0
#define O_LOOKUP_LINEAR 1

// Per-axis tap weight; returns 0 and sets *skip when the reference skips the
// tap (nearest mode radius test). Nearest: CUDA/src/gpu/atomic/
// atomic_gpuNUFFT_kernels.cu:343-368; linear: texture_gpuNUFFT_kernels.cu:63-82.
static DType axis_weight(const o_info* gi, int lookup, const float* lut, DType dsqr, int* skip) {
  if (lookup == O_LOOKUP_NEAREST) {
    if (!(dsqr < gi->radius_sqr)) {
      *skip = 1;
      return 0.0f;
    }
    *skip = 0;
    return lut[(int)std::round(dsqr * gi->dist_multiplier)];
  }
  *skip = 0;
  return lerp_lookup(lut, gi->kernel_count, dsqr * gi->radius_sqr_inv);
}

// ---------------------------------------------------------------------------
// Adjoint spread (E6), GPU semantics with periodic wrap.
// Samples are in sector order: crds SoA [d][M], data complex [C][M].
// grid complex [C][Ng], accumulated into (caller zero-fills).
// 3-D: CUDA/src/gpu/atomic/atomic_gpuNUFFT_kernels.cu:302-407 (nearest),
//      texture_gpuNUFFT_kernels.cu:22-133 (linear).
// 2-D: atomic_gpuNUFFT_kernels.cu:469-572, texture_gpuNUFFT_kernels.cu:209-311.
// Per sector: taps accumulated into a padded sub-grid in sample order
// (z, y, x loops), then the sub-grid is added into the grid with wrap.
// For the linear lookup mode, kernel_count must be the 1-D table length.
// ---------------------------------------------------------------------------
// One sector's samples accumulated in the padded sub-grid `sub` (zeroed
// here) and flushed into `grid` with periodic wrap.
static void spread_sector(const o_info* gi, int lookup, const float* lut, const float* crds,
                          const float* data, unsigned M, unsigned C, const unsigned* sdc,
                          const unsigned* centers, unsigned sec, float* sub, float* grid) {
  const int P = gi->sector_pad_width;
  const int off = gi->sector_offset;
  const int d3 = gi->dim == 3;
  std::fill(sub, sub + (size_t)gi->sector_dim * 2 * C, 0.0f);
  {
    IndType cx = centers[(d3 ? 3 : 2) * sec];
    IndType cy = centers[(d3 ? 3 : 2) * sec + 1];
    IndType cz = d3 ? centers[3 * sec + 2] : 0;
    for (unsigned s = sdc[sec]; s < sdc[sec + 1]; s++) {
      DType px = crds[s], py = crds[s + M], pz = d3 ? crds[s + 2 * M] : 0.0f;
      int imin, imax, jmin, jmax, kmin = 0, kmax = 0;
      DType ix = map_kspace_to_grid(px, gi->grid[0], cx, off);
      set_minmax(&ix, &imin, &imax, gi->sector_pad_max, gi->kernel_radius);
      DType jy = map_kspace_to_grid(py, gi->grid[1], cy, off);
      set_minmax(&jy, &jmin, &jmax, gi->sector_pad_max, gi->kernel_radius);
      if (d3) {
        DType kz = map_kspace_to_grid(pz, gi->grid[2], cz, off);
        set_minmax(&kz, &kmin, &kmax, gi->sector_pad_max, gi->kernel_radius);
      }
      for (int k = kmin; k <= kmax; k++) {
        DType wz = 1.0f;
        if (d3) {
          DType kz = map_grid_to_kspace(k, gi->grid[2], cz, off);
          DType dz = (kz - pz) * gi->aniso[2];
          dz *= dz;
          int skip;
          wz = axis_weight(gi, lookup, lut, dz, &skip);
          if (skip) continue;
        }
        for (int j = jmin; j <= jmax; j++) {
          DType jy2 = map_grid_to_kspace(j, gi->grid[1], cy, off);
          DType dy = (jy2 - py) * gi->aniso[1];
          dy *= dy;
          int skip;
          DType wy = axis_weight(gi, lookup, lut, dy, &skip);
          if (skip) continue;
          for (int i = imin; i <= imax; i++) {
            DType ix2 = map_grid_to_kspace(i, gi->grid[0], cx, off);
            DType dx = (ix2 - px) * gi->aniso[0];
            dx *= dx;
            DType wx = axis_weight(gi, lookup, lut, dx, &skip);
            if (skip) continue;
            DType val = d3 ? wz * wy * wx : wy * wx;
            int ind = i + P * (j + P * k);
            for (unsigned c = 0; c < C; c++) {
              const float* v = data + 2 * ((size_t)c * M + s);
              float* t = &sub[2 * ((size_t)c * gi->sector_dim + ind)];
              t[0] += val * v[0];
              t[1] += val * v[1];
            }
          }
        }
      }
    }
    // Flush with periodic wrap (atomic_gpuNUFFT_kernels.cu:384-406).
    for (int s_ind = 0; s_ind < gi->sector_dim; s_ind++) {
      int x = s_ind % P;
      int z = d3 ? s_ind / (P * P) : 0;
      int y = (s_ind - z * P * P) / P;
      int gx = wrap_index(x, (int)cx, (int)gi->grid[0], off);
      int gy = wrap_index(y, (int)cy, (int)gi->grid[1], off);
      int gz = d3 ? wrap_index(z, (int)cz, (int)gi->grid[2], off) : 0;
      size_t g = (size_t)gx + (size_t)gi->grid[0] * ((size_t)gy + (size_t)gi->grid[1] * gz);
      for (unsigned c = 0; c < C; c++) {
        float* t = &sub[2 * ((size_t)c * gi->sector_dim + s_ind)];
        float* o = grid + 2 * ((size_t)c * gi->grid_count + g);
        o[0] += t[0];
        o[1] += t[1];
      }
    }
  }
}

void o_spread(const o_info* gi, int lookup, const float* lut, const float* crds,
              const float* data, unsigned M, unsigned C, const unsigned* sdc,
              const unsigned* centers, float* grid) {
  std::vector<float> sub((size_t)gi->sector_dim * 2 * C);
  for (unsigned sec = 0; sec < gi->sector_count; sec++)
    spread_sector(gi, lookup, lut, crds, data, M, C, sdc, centers, sec, sub.data(), grid);
}

// The same spread on `threads` OpenMP threads (CPU baseline only): sectors
// are shared out dynamically, each thread flushes into a private grid of
// all coils, and the private grids are summed at the end (in thread order).
void o_spread_mt(const o_info* gi, int lookup, const float* lut, const float* crds,
                 const float* data, unsigned M, unsigned C, const unsigned* sdc,
                 const unsigned* centers, float* grid, int threads) {
  if (threads < 1) threads = 1;
  const size_t n = (size_t)gi->grid_count * 2 * C;
  std::vector<std::vector<float>> priv((size_t)threads);
#pragma omp parallel num_threads(threads)
  {
    const int t = omp_get_thread_num();
    std::vector<float>& g = priv[(size_t)t];
    g.assign(n, 0.0f);
    std::vector<float> sub((size_t)gi->sector_dim * 2 * C);
#pragma omp for schedule(dynamic, 4)
    for (long sec = 0; sec < (long)gi->sector_count; sec++)
      spread_sector(gi, lookup, lut, crds, data, M, C, sdc, centers, (unsigned)sec, sub.data(),
                    g.data());
#pragma omp for schedule(static)
    for (long i = 0; i < (long)n; i++) {
      float acc = grid[i];
      for (int k = 0; k < threads; k++) acc += priv[(size_t)k][(size_t)i];
      grid[i] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// Forward interpolation (E7): sample = sum over taps of weight*grid value,
// taps in (z, y, x) order, grid read with periodic wrap.
// CUDA/src/gpu/atomic/atomic_gpuNUFFT_kernels.cu:966-1066 (3-D),
// :1234-1328 (2-D); texture mode texture_gpuNUFFT_kernels.cu:467-570, 636-739.
// out complex [C][M] in sector order (overwritten).
// ---------------------------------------------------------------------------
// One sector's samples interpolated from `grid` into `out`.
static void interp_sector(const o_info* gi, int lookup, const float* lut, const float* crds,
                          const float* grid, unsigned M, unsigned C, const unsigned* sdc,
                          const unsigned* centers, unsigned sec, float* out) {
  const int off = gi->sector_offset;
  const int d3 = gi->dim == 3;
  {
    IndType cx = centers[(d3 ? 3 : 2) * sec];
    IndType cy = centers[(d3 ? 3 : 2) * sec + 1];
    IndType cz = d3 ? centers[3 * sec + 2] : 0;
    for (unsigned s = sdc[sec]; s < sdc[sec + 1]; s++) {
      DType px = crds[s], py = crds[s + M], pz = d3 ? crds[s + 2 * M] : 0.0f;
      int imin, imax, jmin, jmax, kmin = 0, kmax = 0;
      DType ix = map_kspace_to_grid(px, gi->grid[0], cx, off);
      set_minmax(&ix, &imin, &imax, gi->sector_pad_max, gi->kernel_radius);
      DType jy = map_kspace_to_grid(py, gi->grid[1], cy, off);
      set_minmax(&jy, &jmin, &jmax, gi->sector_pad_max, gi->kernel_radius);
      if (d3) {
        DType kz = map_kspace_to_grid(pz, gi->grid[2], cz, off);
        set_minmax(&kz, &kmin, &kmax, gi->sector_pad_max, gi->kernel_radius);
      }
      for (unsigned c = 0; c < C; c++) {
        float acc_r = 0.0f, acc_i = 0.0f;
        for (int k = kmin; k <= kmax; k++) {
          DType wz = 1.0f;
          if (d3) {
            DType kz = map_grid_to_kspace(k, gi->grid[2], cz, off);
            DType dz = (kz - pz) * gi->aniso[2];
            dz *= dz;
            int skip;
            wz = axis_weight(gi, lookup, lut, dz, &skip);
            if (skip) continue;
          }
          for (int j = jmin; j <= jmax; j++) {
            DType jy2 = map_grid_to_kspace(j, gi->grid[1], cy, off);
            DType dy = (jy2 - py) * gi->aniso[1];
            dy *= dy;
            int skip;
            DType wy = axis_weight(gi, lookup, lut, dy, &skip);
            if (skip) continue;
            for (int i = imin; i <= imax; i++) {
              DType ix2 = map_grid_to_kspace(i, gi->grid[0], cx, off);
              DType dx = (ix2 - px) * gi->aniso[0];
              dx *= dx;
              DType wx = axis_weight(gi, lookup, lut, dx, &skip);
              if (skip) continue;
              DType val = d3 ? wz * wy * wx : wy * wx;
              int gx = wrap_index(i, (int)cx, (int)gi->grid[0], off);
              int gy = wrap_index(j, (int)cy, (int)gi->grid[1], off);
              int gz = d3 ? wrap_index(k, (int)cz, (int)gi->grid[2], off) : 0;
              size_t g = (size_t)gx + (size_t)gi->grid[0] * ((size_t)gy + (size_t)gi->grid[1] * gz);
              const float* v = grid + 2 * ((size_t)c * gi->grid_count + g);
              acc_r += v[0] * val;
              acc_i += v[1] * val;
            }
          }
        }
        out[2 * ((size_t)c * M + s)] = acc_r;
        out[2 * ((size_t)c * M + s) + 1] = acc_i;
      }
    }
  }
}

void o_interp(const o_info* gi, int lookup, const float* lut, const float* crds,
              const float* grid, unsigned M, unsigned C, const unsigned* sdc,
              const unsigned* centers, float* out) {
  for (unsigned sec = 0; sec < gi->sector_count; sec++)
    interp_sector(gi, lookup, lut, crds, grid, M, C, sdc, centers, sec, out);
}

// The same interpolation on `threads` OpenMP threads (CPU baseline only):
// every sample is written by one sector, so sectors run independently.
void o_interp_mt(const o_info* gi, int lookup, const float* lut, const float* crds,
                 const float* grid, unsigned M, unsigned C, const unsigned* sdc,
                 const unsigned* centers, float* out, int threads) {
  if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4)
  for (long sec = 0; sec < (long)gi->sector_count; sec++)
    interp_sector(gi, lookup, lut, crds, grid, M, C, sdc, centers, (unsigned)sec, out);
}

// ---------------------------------------------------------------------------
// The reference CPU gridding, restated: 3-D, isotropic grid of `width`,
// interleaved coordinates (x,y,z per sample), interleaved complex data,
// grid points outside the grid DROPPED (not wrapped).
// CUDA/src/cpu/gpuNUFFT_cpu.cpp:3-159
// ---------------------------------------------------------------------------
void o_cpu_gridding(const float* data, const float* crds, float* gdata, const float* kernel,
                    const int* sectors, int sector_count, const int* sector_centers,
                    int sector_width, int kernel_width, int kernel_count, int width) {
  int imin, imax, jmin, jmax, kmin, kmax, i, j, k, ind;
  DType x, y, z, ix, jy, kz;
  DType dx_sqr, dy_sqr, dz_sqr, val;
  int center_x, center_y, center_z, max_x, max_y, max_z;

  DType kernel_radius = static_cast<DType>(kernel_width) / 2.0f;
  DType radius = kernel_radius / static_cast<DType>(width);
  DType radiusSquared = radius * radius;
  DType kernelRadius_invSqr = 1 / radiusSquared;
  DType dist_multiplier = (kernel_count - 1) * kernelRadius_invSqr;

  int sector_pad_width = sector_width + 2 * (int)std::floor(kernel_width / 2.0f);
  int sector_dim = sector_pad_width * sector_pad_width * sector_pad_width;
  int sector_offset = (int)std::floor(sector_pad_width / 2.0f);

  std::vector<std::vector<DType>> sdata(sector_count);
  for (int sec = 0; sec < sector_count; sec++) {
    sdata[sec].assign((size_t)sector_dim * 2, 0.0f);
    center_x = sector_centers[sec * 3];
    center_y = sector_centers[sec * 3 + 1];
    center_z = sector_centers[sec * 3 + 2];
    for (int data_cnt = sectors[sec]; data_cnt < sectors[sec + 1]; data_cnt++) {
      x = crds[3 * data_cnt];
      y = crds[3 * data_cnt + 1];
      z = crds[3 * data_cnt + 2];
      max_x = sector_pad_width - 1;
      max_y = sector_pad_width - 1;
      max_z = sector_pad_width - 1;
      ix = (x + 0.5f) * (width)-center_x + sector_offset;
      set_minmax(&ix, &imin, &imax, max_x, kernel_radius);
      jy = (y + 0.5f) * (width)-center_y + sector_offset;
      set_minmax(&jy, &jmin, &jmax, max_y, kernel_radius);
      kz = (z + 0.5f) * (width)-center_z + sector_offset;
      set_minmax(&kz, &kmin, &kmax, max_z, kernel_radius);
      for (k = kmin; k <= kmax; k++) {
        kz = static_cast<DType>((k + center_z - sector_offset)) / static_cast<DType>((width)) - 0.5f;
        dz_sqr = kz - z;
        dz_sqr *= dz_sqr;
        if (dz_sqr < radiusSquared) {
          for (j = jmin; j <= jmax; j++) {
            jy = static_cast<DType>(j + center_y - sector_offset) / static_cast<DType>((width)) - 0.5f;
            dy_sqr = jy - y;
            dy_sqr *= dy_sqr;
            if (dy_sqr < radiusSquared) {
              for (i = imin; i <= imax; i++) {
                ix = static_cast<DType>(i + center_x - sector_offset) / static_cast<DType>((width)) - 0.5f;
                dx_sqr = ix - x;
                dx_sqr *= dx_sqr;
                if (dx_sqr < radiusSquared) {
                  val = kernel[(int)std::round(dz_sqr * dist_multiplier)] *
                        kernel[(int)std::round(dy_sqr * dist_multiplier)] *
                        kernel[(int)std::round(dx_sqr * dist_multiplier)];
                  ind = i + sector_pad_width * (j + sector_pad_width * k);
                  sdata[sec][2 * ind] += val * data[2 * data_cnt];
                  sdata[sec][2 * ind + 1] += val * data[2 * data_cnt + 1];
                }
              }
            }
          }
        }
      }
    }
  }
  for (int sec = 0; sec < sector_count; sec++) {
    center_x = sector_centers[sec * 3];
    center_y = sector_centers[sec * 3 + 1];
    center_z = sector_centers[sec * 3 + 2];
    int sector_ind_offset = (center_x - sector_offset) +
                            width * ((center_y - sector_offset) + width * (center_z - sector_offset));
    for (int zz = 0; zz < sector_pad_width; zz++)
      for (int yy = 0; yy < sector_pad_width; yy++)
        for (int xx = 0; xx < sector_pad_width; xx++) {
          int s_ind = 2 * (xx + sector_pad_width * (yy + sector_pad_width * zz));
          ind = 2 * (sector_ind_offset + xx + width * (yy + width * zz));
          if ((center_x - sector_offset + xx) >= width || (center_x - sector_offset + xx) < 0 ||
              (center_y - sector_offset + yy) >= width || (center_y - sector_offset + yy) < 0 ||
              (center_z - sector_offset + zz) >= width || (center_z - sector_offset + zz) < 0)
            continue;
          gdata[ind] += sdata[sec][s_ind];
          gdata[ind + 1] += sdata[sec][s_ind + 1];
        }
  }
}

}  // extern "C"
