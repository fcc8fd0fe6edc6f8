// C ABI of libgpunufft_hip.so (declarations: include/gpunufft_hip.h).
//
// A plan is one operator: geometry, lookup table, deapodization array, the
// device-resident sector tables of one trajectory, optional coil
// sensitivities, a coil-batched work grid and rocFFT plans. All operator
// state is per plan (kernel arguments, no process-global symbols), so plans
// on different streams run concurrently.

#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "../../../include/gpunufft_hip.h"
#include "gnufft_internal.hpp"

using namespace gnufft;

#ifndef GNUFFT_NOALLOC_SCRATCH_BYTES
#define GNUFFT_NOALLOC_SCRATCH_BYTES (64ull << 20)
#endif

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

}  // namespace

void gnufft::set_last_error(const std::string& msg) { g_last_error = msg; }

namespace {

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail(GNUFFT_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));  \
  } while (0)

#define ROCFFT_TRY(expr)                                                                   \
  do {                                                                                     \
    rocfft_status s_ = (expr);                                                             \
    if (s_ != rocfft_status_success)                                                       \
      return fail(GNUFFT_ERR_FFT, std::string(#expr) + " failed with rocfft status " +     \
                                      std::to_string((int)s_));                            \
  } while (0)

std::once_flag g_rocfft_once;

// ---------------------------------------------------------------------------
// Kaiser-Bessel lookup table, float arithmetic as the reference
// (CUDA/src/gpuNUFFT_utils.cpp:3-107, CUDA/inc/gpuNUFFT_utils.hpp:60-75).
// ---------------------------------------------------------------------------
float bessel_i0(float x) {
  float ax = std::fabs(x);
  float ans, y;
  if (ax < 3.75f) {
    y = x / 3.75f, y = y * y;
    ans = 1.0f + y * (3.5156229f + y * (3.0899424f + y * (1.2067492f +
                  y * (0.2659732f + y * (0.360768e-1f + y * 0.45813e-2f)))));
  } else {
    y = 3.75f / ax;
    ans = (std::exp(ax) / std::sqrt(ax)) *
          (0.39894228f + y * (0.1328592e-1f + y * (0.225319e-2f + y * (-0.157565e-2f +
           y * (0.916281e-2f + y * (-0.2057706e-1f + y * (0.2635537e-1f +
           y * (-0.1647633e-1f + y * 0.392377e-2f))))))));
  }
  return ans;
}

float kb_beta(unsigned kw, float osf) {
  return (float)M_PI * std::sqrt(std::pow((float)kw / osf * (osf - 0.5f), 2.0f) - 0.8f);
}

float kb_value(float radius, unsigned kw, float osf) {
  return bessel_i0(kb_beta(kw, osf) * std::sqrt(1.0f - std::pow(radius, 2.0f))) /
         bessel_i0(kb_beta(kw, osf));
}

long lut_size(float osf, unsigned kw, int lookup) {
  if (lookup == GNUFFT_LOOKUP_NEAREST) {
    long kosf = (long)(std::floor(0.91f / (osf * 0.001f)));
    return (long)(kosf * (float)kw * 0.5);
  }
  long kosf = (long)(std::floor(std::sqrt(0.37 / ((osf * osf) * 0.0001f))));
  return (long)(kosf * (float)kw * 0.5);
}

std::vector<float> make_lut(long K, unsigned kw, float osf) {
  std::vector<float> t((size_t)K, 0.0f);
  for (long i = 1; i < K - 1; i++) {
    if (kw == 1) {
      t[i] = 1.0f;
    } else {
      double r = (double)std::sqrt(i / (double)(K - 1));
      t[i] = kb_value((float)r, kw, osf);
    }
  }
  t[0] = 1.0f;
  t[K - 1] = 0.0f;
  return t;
}

// Sector index along one axis, host twin of the device assignment
// (CUDA/inc/precomp_utils.hpp:34-48).
#pragma clang fp contract(off)
int host_sector_of(float coord, int G, int sw) {
  double v = (double)(coord * (float)G) + 0.5 * (double)(float)G;
  int sector = (int)std::floor(std::round(v) / (double)(float)sw);
  int smax = (int)std::ceil((float)G / (float)sw);
  if (sector >= smax) sector = smax - 1;
  if (sector < 0) sector = 0;
  return sector;
}
#pragma clang fp contract(on)

// Geometry of an in-place C2C rocFFT plan pair on one device.
struct FftKey {
  int device = 0, dims = 0;
  uint32_t G0 = 0, G1 = 0, G2 = 0, batch = 0;
  bool operator<(const FftKey& o) const {
    return std::tie(device, dims, G0, G1, G2, batch) <
           std::tie(o.device, o.dims, o.G0, o.G1, o.G2, o.batch);
  }
};

struct FftPlans {
  rocfft_plan fwd = nullptr, inv = nullptr;
  rocfft_execution_info info = nullptr;
  void* work = nullptr;
  size_t work_bytes = 0;
  FftKey key;
};

// Idle rocFFT plan pairs, kept when their owner (a plan or a batch group)
// lets them go, so that a later owner of the same geometry on the same
// device skips rocFFT's plan construction (milliseconds of host time). A
// pair has one owner at a time; the pool keeps at most kFftPoolMax pairs
// (with their twiddle tables in device memory) until the process ends.
constexpr size_t kFftPoolMax = 32;
std::mutex g_fft_pool_mu;
std::multimap<FftKey, std::pair<rocfft_plan, rocfft_plan>> g_fft_pool;

bool take_pooled_fft(const FftKey& key, rocfft_plan* fwd, rocfft_plan* inv) {
  std::lock_guard<std::mutex> lk(g_fft_pool_mu);
  auto it = g_fft_pool.find(key);
  if (it == g_fft_pool.end()) return false;
  *fwd = it->second.first;
  *inv = it->second.second;
  g_fft_pool.erase(it);
  return true;
}

// Returns f's plan pair to the pool (or destroys it when the pool is full).
void release_fft_plans(FftPlans& f) {
  if (f.fwd && f.inv) {
    std::lock_guard<std::mutex> lk(g_fft_pool_mu);
    if (g_fft_pool.size() < kFftPoolMax) {
      g_fft_pool.emplace(f.key, std::make_pair(f.fwd, f.inv));
      f.fwd = f.inv = nullptr;
    }
  }
  if (f.fwd) rocfft_plan_destroy(f.fwd);
  if (f.inv) rocfft_plan_destroy(f.inv);
  f.fwd = f.inv = nullptr;
}

// HIP-event stage timer of one plan.
struct StageTimer {
  bool on = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[GNUFFT_T_COUNT];
  double ms[GNUFFT_T_COUNT] = {0, 0, 0, 0, 0};
  uint32_t calls[GNUFFT_T_COUNT] = {0, 0, 0, 0, 0};

  hipEvent_t begin(int cat, hipStream_t s) {
    if (!on) return nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess) return nullptr;
    if (hipEventCreate(&b) != hipSuccess) {
      (void)hipEventDestroy(a);
      return nullptr;
    }
    (void)hipEventRecord(a, s);
    ev[cat].push_back({a, b});
    return b;
  }
  void end(hipEvent_t b, hipStream_t s) {
    if (b) (void)hipEventRecord(b, s);
  }
  void collect() {
    for (int c = 0; c < GNUFFT_T_COUNT; c++) {
      for (auto& e : ev[c]) {
        float t = 0.0f;
        if (hipEventSynchronize(e.second) == hipSuccess &&
            hipEventElapsedTime(&t, e.first, e.second) == hipSuccess) {
          ms[c] += t;
          calls[c]++;
        }
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
      }
      ev[c].clear();
    }
  }
};

}  // namespace

struct gnufft_plan_s {
  int device = 0;
  Geom g{};
  uint32_t img[3] = {0, 0, 0}, grid[3] = {0, 0, 0}, sectors[3] = {1, 1, 1};
  uint32_t S = 0;
  int lookup = 0;
  unsigned kw = 3, sw = 8;
  float osf = 2.0f;
  std::vector<float> lut;
  float* lut_d = nullptr;
  float* deapo_d = nullptr;
  // trajectory tables
  uint32_t M = 0;
  float* traj_sorted_d = nullptr;
  float* dens_sorted_d = nullptr;
  uint32_t* idx_d = nullptr;
  uint32_t* assigned_d = nullptr;  // sector id per sample, caller order
  uint32_t* sdc_d = nullptr;
  uint32_t* centers_d = nullptr;
  Chunk* chunks_d = nullptr;
  uint32_t n_chunks = 0;
  uint32_t chunk = 1024;
  // samples per LDS round of the gridding kernels: the mean chunk occupancy
  // rounded up to 64 (sparse trajectories get small footprints, hence more
  // workgroups per CU), at most `chunk`
  uint32_t round_target = 1024;
  WorkList work{};
  // the chunks of each sector merged in runs of up to kInterpRun (the
  // round-based interp stages a sector once per run); empty when no sector
  // has more than one chunk
  Chunk* runs_d = nullptr;
  WorkList work_runs{};
  int lane_per_cu = 0;  // gnufft_plan_set_workgroups_per_cu (0: default)
  uint32_t* sector_chunks_d = nullptr;  // [S+1] first chunk of each sector
  uint2* reduce_d = nullptr;            // (first chunk, count) of multi-chunk sectors
  uint32_t n_reduce = 0;
  // per-axis padded-sector cover lists of the grid (fold)
  uint32_t* cover_d = nullptr;
  Cover cover{};
  // spread sub-grids, [chunk][padded point][coil], capacity in coils
  float2* scratch_d = nullptr;
  uint32_t scratch_coils = 0;
  // sensitivities
  float2* sens_d = nullptr;
  uint32_t n_sens = 0;
  // work grid, capacity in coils
  float2* grid_d = nullptr;
  uint32_t grid_coils = 0;
  float2* img_tmp_d = nullptr;
  uint32_t img_tmp_coils = 0;
  std::map<uint32_t, FftPlans> ffts;
  // fused FFT (GNUFFT_FFT_FUSED): twiddles [axis][sign][G_a] and stage tables
  bool fused_ok = false;
  bool fused = false;
  float2* fft_tw_d = nullptr;
  size_t fft_tw_off[3][3] = {};
  StageTimer timer;
  // bumped whenever device memory that a captured graph or a batch may
  // hold a pointer to is freed or replaced (trajectory tables, work grid,
  // scratch, sensitivities, image buffer) or the FFT kind changes
  uint64_t generation = 0;
  // the part of that a batch depends on: trajectory tables and sensitivities
  uint64_t table_generation = 0;
};

// Runs `expr` (a hipError_t launch) bracketed by the plan's stage timer.
#define TIMED(plan, cat, stream, expr)                         \
  do {                                                         \
    hipEvent_t t_end_ = (plan)->timer.begin((cat), (stream));  \
    hipError_t t_e_ = (expr);                                  \
    (plan)->timer.end(t_end_, (stream));                       \
    HIP_TRY(t_e_);                                             \
  } while (0)

uint64_t gnufft::plan_generation(const gnufft_plan_s* p) { return p ? p->generation : 0; }

namespace {

void dfree(void* p) {
  if (p) (void)hipFree(p);
}

void free_traj(gnufft_plan p) {
  p->generation++;
  p->table_generation++;
  dfree(p->traj_sorted_d);
  dfree(p->dens_sorted_d);
  dfree(p->idx_d);
  dfree(p->assigned_d);
  dfree(p->sdc_d);
  dfree(p->centers_d);
  dfree(p->chunks_d);
  dfree(p->runs_d);
  dfree(p->sector_chunks_d);
  dfree(p->reduce_d);
  dfree(p->scratch_d);
  p->traj_sorted_d = p->dens_sorted_d = nullptr;
  p->idx_d = p->assigned_d = p->sdc_d = p->centers_d = nullptr;
  p->chunks_d = nullptr;
  p->runs_d = nullptr;
  p->work_runs = WorkList{};
  p->sector_chunks_d = nullptr;
  p->reduce_d = nullptr;
  p->n_reduce = 0;
  p->scratch_d = nullptr;
  p->scratch_coils = 0;
  p->M = 0;
  p->n_chunks = 0;
  p->work = WorkList{};
  p->g.M = 0;
}

void free_ffts(gnufft_plan p) {
  for (auto& kv : p->ffts) {
    release_fft_plans(kv.second);
    if (kv.second.info) rocfft_execution_info_destroy(kv.second.info);
    dfree(kv.second.work);
  }
  p->ffts.clear();
}

// Deapodization, following the reference's numeric procedure
// (CUDA/src/gpuNUFFT_operator_factory.cpp:359-448): one unit sample at k=0
// spread with sector width 8 and the operator's lookup mode, INVERSE shift,
// unnormalised inverse DFT, INVERSE shift, crop, 1/|.|.
// The spread of one sample is its tap tensor of float products
// (wz*wy)*wx exactly as the spread kernel forms them; its DFT is evaluated
// axis by axis in double (the phases are separable), which reproduces the
// reference's values also where |K| is tiny and the rounding of the tap
// products is amplified (osf near 1). deapo = sqrt(Ng/Ni)/|K(p)|: the
// compensation the reference's observed outputs correspond to. The x and y
// sums run on the host (tap count x image row / plane), the z sum and the
// modulus per pixel on the device.
int compute_deapo(gnufft_plan p, hipStream_t s) {
  Geom d = p->g;  // same table, radius and anisotropy; sector width 8
  d.sw = 8;
  d.P = 8 + 2 * (int)std::floor((float)p->kw / 2.0f);
  d.Pmax = d.P - 1;
  d.off = (int)std::floor((float)d.P / 2.0f);
  const int D = d.dim;
  typedef std::complex<double> cd;
  // per axis: tap weights and the phase matrix E[t][x]
  std::vector<float> w[3];
  std::vector<cd> E[3];
  int T[3] = {1, 1, 1};
  for (int a = 0; a < 3; a++) {
    const int N = d.N[a];
    if (a >= D) {
      w[a].assign(1, 1.0f);
      E[a].assign((size_t)N, cd(1.0, 0.0));
      continue;
    }
    const int G = d.G[a];
    const int sec = host_sector_of(0.0f, G, 8);
    const int center = sec * 8 + 4;
    const float u = map_kspace_to_grid(0.0f, G, center, d.off);
    int lo, hi;
    tap_range(u, d.Pmax, d.kradius, &lo, &hi);
    std::vector<int> pos;
    for (int i = lo; i <= hi; i++) {
      float wt;
      if (!axis_weight(d, p->lut.data(), a, i, center, 0.0f, &wt)) continue;
      w[a].push_back(wt);
      pos.push_back(wrap(wrap(center - d.off + i, G) - d.shift_inv[a], G));
    }
    T[a] = (int)w[a].size();
    E[a].resize((size_t)T[a] * N);
    for (int t = 0; t < T[a]; t++)
      for (int x = 0; x < N; x++) {
        const int m = wrap(x + d.crop[a] + d.shift_inv[a], G);
        const double ph = 2.0 * M_PI * (double)(((long long)pos[t] * m) % G) / (double)G;
        E[a][(size_t)t * N + x] = cd(std::cos(ph), std::sin(ph));
      }
  }
  const int Nx = d.N[0], Ny = d.N[1], Nz = d.N[2];
  // A[tz][ty][x] = sum_tx W[tz][ty][tx] Ex[tx][x]
  std::vector<cd> A((size_t)T[2] * T[1] * Nx), B((size_t)T[2] * Ny * Nx);
  for (int tz = 0; tz < T[2]; tz++)
    for (int ty = 0; ty < T[1]; ty++)
      for (int x = 0; x < Nx; x++) {
        cd acc(0.0, 0.0);
        for (int tx = 0; tx < T[0]; tx++) {
          const float wt = D == 3 ? (w[2][tz] * w[1][ty]) * w[0][tx] : w[1][ty] * w[0][tx];
          acc += (double)wt * E[0][(size_t)tx * Nx + x];
        }
        A[((size_t)tz * T[1] + ty) * Nx + x] = acc;
      }
  // B[tz][y][x] = sum_ty A[tz][ty][x] Ey[ty][y]
  for (int tz = 0; tz < T[2]; tz++)
    for (int y = 0; y < Ny; y++)
      for (int x = 0; x < Nx; x++) {
        cd acc(0.0, 0.0);
        for (int ty = 0; ty < T[1]; ty++)
          acc += A[((size_t)tz * T[1] + ty) * Nx + x] * E[1][(size_t)ty * Ny + y];
        B[((size_t)tz * Ny + y) * Nx + x] = acc;
      }
  // the z axis on the device: Ni points (3.5 M at 152^3) would take tens
  // of milliseconds of host time per plan
  const double norm = std::sqrt((double)d.Ng / (double)d.Ni);
  static_assert(sizeof(cd) == sizeof(double2), "complex<double> is two doubles");
  double2* tab = nullptr;
  const size_t nb = B.size(), ne = E[2].size();
  HIP_TRY(hipMalloc(&tab, (nb + ne) * sizeof(double2)));
  hipError_t e = hipMemcpyAsync(tab, B.data(), nb * sizeof(double2), hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(tab + nb, E[2].data(), ne * sizeof(double2), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = launch_deapo_combine(tab, tab + nb, T[2], Nx, Ny, Nz, norm, p->deapo_d, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);  // host tables B, E go out of scope
  (void)hipFree(tab);
  HIP_TRY(e);
  return GNUFFT_OK;
}

int ensure_grid(gnufft_plan p, uint32_t coils) {
  if (p->grid_coils >= coils) return GNUFFT_OK;
  p->generation++;
  dfree(p->grid_d);
  p->grid_d = nullptr;
  p->grid_coils = 0;
  HIP_TRY(hipMalloc(&p->grid_d, (size_t)coils * p->g.Ng * sizeof(float2)));
  p->grid_coils = coils;
  return GNUFFT_OK;
}

// Coils processed together: all of them unless the grid would take more
// than half of the free device memory (coil chunking,
// CUDA/src/gpuNUFFT_operator.cpp:327-368).
size_t padded_points(const Geom& g) {
  return g.dim == 3 ? (size_t)g.P * g.P * g.P : (size_t)g.P * g.P;
}

// spread sub-grids [chunk][padded point] per coil; the same memory is the
// fused FFT's intermediate array (used after the fold has consumed it)
size_t scratch_per_coil(gnufft_plan p) {
  const size_t sub = (size_t)p->n_chunks * padded_points(p->g);
  const size_t fft = p->fused_ok ? fused_fft_tmp_elems(p->g) : 0;
  return std::max(sub, fft) * sizeof(float2);
}

int ensure_scratch(gnufft_plan p, uint32_t coils) {
  if (p->scratch_coils >= coils) return GNUFFT_OK;
  p->generation++;
  dfree(p->scratch_d);
  p->scratch_d = nullptr;
  p->scratch_coils = 0;
  const size_t bytes = (size_t)coils * scratch_per_coil(p);
  if (bytes == 0) return GNUFFT_OK;
  if (hipMalloc(&p->scratch_d, bytes) != hipSuccess)
    return fail(GNUFFT_ERR_OUT_OF_MEMORY, "spread: sub-grid scratch allocation");
  p->scratch_coils = coils;
  return GNUFFT_OK;
}

uint32_t coil_batch(gnufft_plan p, uint32_t n_coils) {
  // buffers already hold every coil (steady state, e.g. under graph capture)
  if (p->grid_coils >= n_coils && p->scratch_coils >= n_coils) return n_coils;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 1;
  // grid + FFT work + spread sub-grids
  size_t per = (size_t)p->g.Ng * sizeof(float2) * 2 + scratch_per_coil(p);
  size_t avail = fr / 2 + (size_t)p->grid_coils * p->g.Ng * sizeof(float2) +
                 (size_t)p->scratch_coils * scratch_per_coil(p);
  size_t cc = per ? avail / per : n_coils;
  if (cc < 1) cc = 1;
  return (uint32_t)std::min<size_t>(cc, n_coils);
}

// In-place C2C rocFFT plans (both directions) of `batch` grids of g's size.
// On any failure everything made so far is released (plans, whether new or
// taken from the pool, execution info and work buffer).
int create_fft(const Geom& g, uint32_t batch, FftPlans* out) {
  FftPlans f;
  size_t lengths[3] = {(size_t)g.G[0], (size_t)g.G[1], (size_t)g.G[2]};
  const size_t dims = (size_t)g.dim;
  HIP_TRY(hipGetDevice(&f.key.device));
  f.key.dims = g.dim;
  f.key.G0 = g.G[0];
  f.key.G1 = g.G[1];
  f.key.G2 = g.G[2];
  f.key.batch = batch;
  auto bail = [&](int code, const std::string& msg) {
    if (f.fwd) rocfft_plan_destroy(f.fwd);
    if (f.inv) rocfft_plan_destroy(f.inv);
    if (f.info) rocfft_execution_info_destroy(f.info);
    if (f.work) (void)hipFree(f.work);
    f = FftPlans();
    return fail(code, msg);
  };
  if (!take_pooled_fft(f.key, &f.fwd, &f.inv)) {
    if (rocfft_plan_create(&f.fwd, rocfft_placement_inplace, rocfft_transform_type_complex_forward,
                           rocfft_precision_single, dims, lengths, batch, nullptr) != rocfft_status_success) {
      f.fwd = nullptr;
      return bail(GNUFFT_ERR_FFT, "rocfft_plan_create (forward) failed");
    }
    if (rocfft_plan_create(&f.inv, rocfft_placement_inplace, rocfft_transform_type_complex_inverse,
                           rocfft_precision_single, dims, lengths, batch, nullptr) != rocfft_status_success) {
      f.inv = nullptr;
      return bail(GNUFFT_ERR_FFT, "rocfft_plan_create (inverse) failed");
    }
  }
  size_t w1 = 0, w2 = 0;
  if (rocfft_plan_get_work_buffer_size(f.fwd, &w1) != rocfft_status_success ||
      rocfft_plan_get_work_buffer_size(f.inv, &w2) != rocfft_status_success)
    return bail(GNUFFT_ERR_FFT, "rocfft_plan_get_work_buffer_size failed");
  f.work_bytes = std::max(w1, w2);
  if (f.work_bytes && hipMalloc(&f.work, f.work_bytes) != hipSuccess) {
    f.work = nullptr;
    return bail(GNUFFT_ERR_OUT_OF_MEMORY, "rocFFT work buffer allocation");
  }
  if (rocfft_execution_info_create(&f.info) != rocfft_status_success) {
    f.info = nullptr;
    return bail(GNUFFT_ERR_FFT, "rocfft_execution_info_create failed");
  }
  if (f.work_bytes &&
      rocfft_execution_info_set_work_buffer(f.info, f.work, f.work_bytes) != rocfft_status_success)
    return bail(GNUFFT_ERR_FFT, "rocfft_execution_info_set_work_buffer failed");
  *out = f;
  return GNUFFT_OK;
}

void destroy_fft(FftPlans& f) {
  release_fft_plans(f);
  if (f.info) rocfft_execution_info_destroy(f.info);
  if (f.work) (void)hipFree(f.work);
  f = FftPlans();
}

int get_fft(gnufft_plan p, uint32_t batch, FftPlans** out) {
  auto it = p->ffts.find(batch);
  if (it != p->ffts.end()) {
    *out = &it->second;
    return GNUFFT_OK;
  }
  FftPlans f;
  int rc = create_fft(p->g, batch, &f);
  if (rc) return rc;
  auto ins = p->ffts.emplace(batch, f);
  *out = &ins.first->second;
  return GNUFFT_OK;
}

int exec_fft(FftPlans& f, bool inverse, void* buf, hipStream_t s) {
  ROCFFT_TRY(rocfft_execution_info_set_stream(f.info, s));
  void* bufs[1] = {buf};
  ROCFFT_TRY(rocfft_execute(inverse ? f.inv : f.fwd, bufs, nullptr, f.info));
  return GNUFFT_OK;
}

int run_fft(gnufft_plan p, uint32_t batch, bool inverse, void* buf, hipStream_t s) {
  FftPlans* f = nullptr;
  int rc = get_fft(p, batch, &f);
  if (rc) return rc;
  ROCFFT_TRY(rocfft_execution_info_set_stream(f->info, s));
  void* bufs[1] = {buf};
  hipEvent_t t_end = p->timer.begin(GNUFFT_T_FFT, s);
  rocfft_status st = rocfft_execute(inverse ? f->inv : f->fwd, bufs, nullptr, f->info);
  p->timer.end(t_end, s);
  ROCFFT_TRY(st);
  return GNUFFT_OK;
}

int copy_in(void** dst, const void* src, size_t bytes, int location, hipStream_t s) {
  HIP_TRY(hipMalloc(dst, bytes ? bytes : 4));
  if (bytes)
    HIP_TRY(hipMemcpyAsync(*dst, src, bytes,
                           location == GNUFFT_DEVICE ? hipMemcpyDeviceToDevice
                                                     : hipMemcpyHostToDevice,
                           s));
  return GNUFFT_OK;
}

// Gridding work list from the sector offsets.
int build_chunks(gnufft_plan p, hipStream_t s) {
  const uint32_t S = p->S;
  uint32_t* off_d = nullptr;
  uint32_t* cnt_d = nullptr;
  void* temp = nullptr;
  size_t temp_bytes = 0;
  HIP_TRY(hipMalloc(&off_d, (S + 1) * sizeof(uint32_t)));
  p->sector_chunks_d = off_d;  // kept: first chunk of each sector (fold)
  HIP_TRY(hipMalloc(&cnt_d, (S + 1) * sizeof(uint32_t)));
  hipError_t e = count_chunks(p->sdc_d, S, p->chunk, cnt_d, off_d, nullptr, &temp_bytes, s);
  if (e == hipSuccess) e = hipMalloc(&temp, temp_bytes ? temp_bytes : 4);
  if (e == hipSuccess) e = count_chunks(p->sdc_d, S, p->chunk, cnt_d, off_d, temp, &temp_bytes, s);
  uint32_t total = 0;
  if (e == hipSuccess)
    e = hipMemcpyAsync(&total, off_d + S, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = hipMalloc(&p->chunks_d, (size_t)(total ? total : 1) * sizeof(Chunk));
  if (e == hipSuccess) e = launch_fill_chunks(p->sdc_d, off_d, S, p->chunk, p->chunks_d, s);
  std::vector<Chunk> h(total);
  if (e == hipSuccess && total)
    e = hipMemcpyAsync(h.data(), p->chunks_d, (size_t)total * sizeof(Chunk), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(cnt_d);
  dfree(temp);
  if (e != hipSuccess) return fail(GNUFFT_ERR_DEVICE, std::string("chunk list: ") + hipGetErrorString(e));
  p->n_chunks = total;
  std::vector<uint32_t> counts(total);
  for (uint32_t i = 0; i < total; i++) counts[i] = h[i].end - h[i].begin;
  p->work.chunks = p->chunks_d;
  partition_work(counts.data(), total, &p->work);
  {
    const uint64_t mean = total ? ((uint64_t)p->M + total - 1) / total : 64;
    p->round_target = (uint32_t)std::min<uint64_t>(p->chunk, std::max<uint64_t>(64, (mean + 63) / 64 * 64));
  }
  // sectors split over several chunks: their sub-grids are summed before the
  // fold (one workgroup per 256 words of each such sector, so the dense
  // k-space centre spreads over the whole GPU)
  std::vector<uint2> tasks;
  for (uint32_t i = 0; i < total;) {
    uint32_t j = i + 1;
    while (j < total && h[j].sector == h[i].sector) j++;
    if (j - i > 1) tasks.push_back(make_uint2(i, j - i));
    i = j;
  }
  p->n_reduce = (uint32_t)tasks.size();
  if (!tasks.empty()) {
    int rc = copy_in((void**)&p->reduce_d, tasks.data(), tasks.size() * sizeof(uint2), GNUFFT_HOST, s);
    if (rc) return rc;
    std::vector<Chunk> runs;
    std::vector<uint32_t> run_counts;
    for (uint32_t i = 0; i < total;) {
      uint32_t j = i + 1;
      while (j < total && j - i < (uint32_t)kInterpRun && h[j].sector == h[i].sector) j++;
      runs.push_back(Chunk{h[i].sector, h[i].begin, h[j - 1].end, 0});
      run_counts.push_back(h[j - 1].end - h[i].begin);
      i = j;
    }
    rc = copy_in((void**)&p->runs_d, runs.data(), runs.size() * sizeof(Chunk), GNUFFT_HOST, s);
    if (rc) {
      (void)hipStreamSynchronize(s);  // the task upload above reads `tasks`
      return rc;
    }
    p->work_runs.chunks = p->runs_d;
    partition_work(run_counts.data(), (uint32_t)runs.size(), &p->work_runs);
    HIP_TRY(hipStreamSynchronize(s));
  }
  return GNUFFT_OK;
}

bool valid_stop(int stop) { return stop >= 0 && stop <= 3; }

// Spread arguments shared by every entry point: samples in sector order
// (no gather, no density) unless the caller sets them.
SpreadArgs spread_args(gnufft_plan p, const float2* data, float2* grid, uint32_t n_coils) {
  SpreadArgs a{};
  a.data = data;
  a.coil_stride = p->M;
  a.crds = p->traj_sorted_d;
  a.work = p->work;
  a.lut = p->lut_d;
  a.scratch = p->scratch_d;
  a.sector_chunks = p->sector_chunks_d;
  a.reduce_tasks = p->reduce_d;
  a.n_reduce = p->n_reduce;
  a.cover = p->cover;
  a.grid = grid;
  a.n_coils = (int)n_coils;
  a.lane_per_cu = p->lane_per_cu;
  return a;
}


}  // namespace

extern "C" {

const char* gnufft_last_error(void) { return g_last_error.c_str(); }

int gnufft_abi_version(void) { return GNUFFT_ABI_VERSION; }

int gnufft_malloc(void** ptr, size_t bytes) {
  if (!ptr) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "gnufft_malloc: null ptr");
  HIP_TRY(hipMalloc(ptr, bytes ? bytes : 4));
  return GNUFFT_OK;
}

int gnufft_free(void* ptr) {
  if (ptr) HIP_TRY(hipFree(ptr));
  return GNUFFT_OK;
}

int gnufft_memcpy(void* dst, const void* src, size_t bytes, int dst_loc, int src_loc,
                  void* stream) {
  if (bytes == 0) return GNUFFT_OK;
  if (!dst || !src) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "gnufft_memcpy: null pointer");
  hipMemcpyKind k = dst_loc == GNUFFT_DEVICE
                        ? (src_loc == GNUFFT_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice)
                        : (src_loc == GNUFFT_DEVICE ? hipMemcpyDeviceToHost : hipMemcpyHostToHost);
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, k, (hipStream_t)stream));
  if (dst_loc == GNUFFT_HOST || src_loc == GNUFFT_HOST)
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  return GNUFFT_OK;
}

int gnufft_zero(void* ptr, size_t bytes, void* stream) {
  if (!ptr && bytes) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "gnufft_zero: null pointer");
  HIP_TRY(launch_zero(ptr, bytes, (hipStream_t)stream));
  return GNUFFT_OK;
}

int gnufft_stream_synchronize(void* stream) {
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  return GNUFFT_OK;
}

int gnufft_mem_info(size_t* free_bytes, size_t* total_bytes) {
  size_t f = 0, t = 0;
  HIP_TRY(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return GNUFFT_OK;
}

int gnufft_plan_create(gnufft_plan* plan, const uint32_t img_dims[3], uint32_t kernel_width,
                       uint32_t sector_width, float osf, int lookup, void* stream) {
  if (!plan || !img_dims) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "plan_create: null argument");
  *plan = nullptr;
  if (img_dims[0] == 0 || img_dims[1] == 0)
    return fail(GNUFFT_ERR_INVALID_ARGUMENT, "plan_create: image width and height must be > 0");
  if (kernel_width < 1 || kernel_width > 7)
    return fail(GNUFFT_ERR_UNSUPPORTED, "plan_create: kernel width must be in [1, 7]");
  if (sector_width < 1) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "plan_create: sector width 0");
  if (!(osf >= 1.0f)) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "plan_create: osf must be >= 1");
  if (lookup != GNUFFT_LOOKUP_NEAREST && lookup != GNUFFT_LOOKUP_LINEAR)
    return fail(GNUFFT_ERR_INVALID_ARGUMENT, "plan_create: unknown lookup mode");
  hipStream_t s = (hipStream_t)stream;
  std::call_once(g_rocfft_once, []() { rocfft_setup(); });

  gnufft_plan p = new (std::nothrow) gnufft_plan_s();
  if (!p) return fail(GNUFFT_ERR_OUT_OF_MEMORY, "plan_create: host allocation");
  (void)hipGetDevice(&p->device);
  p->kw = kernel_width;
  p->sw = sector_width;
  p->osf = osf;
  p->lookup = lookup;
  Geom& g = p->g;
  g.dim = img_dims[2] == 0 ? 2 : 3;
  // grid dims: (IndType)(N * osf), CUDA/inc/gpuNUFFT_types.hpp:109-117
  for (int a = 0; a < 3; a++) {
    p->img[a] = img_dims[a];
    p->grid[a] = (uint32_t)(img_dims[a] * osf);
  }
  if (p->grid[0] == 0 || p->grid[1] == 0 || (g.dim == 3 && p->grid[2] == 0)) {
    delete p;
    return fail(GNUFFT_ERR_INVALID_ARGUMENT, "plan_create: empty grid");
  }
  const uint32_t max_grid = std::max(std::max(p->grid[0], p->grid[1]), p->grid[2]);
  for (int a = 0; a < 3; a++) {
    const bool used = a < g.dim;
    g.G[a] = used ? (int)p->grid[a] : 1;
    g.N[a] = used ? (int)p->img[a] : 1;
    p->sectors[a] = used ? (uint32_t)std::ceil((float)p->grid[a] / (float)sector_width) : 1;
    g.S[a] = (int)p->sectors[a];
    g.shift_inv[a] = used ? g.G[a] / 2 : 0;
    g.shift_fwd[a] = used ? (g.G[a] + 1) / 2 : 0;
    g.crop[a] = used ? (int)(uint32_t)((float)p->img[a] * (osf - 1.0f) / 2.0f) : 0;
    g.aniso[a] = (float)p->grid[a] / (float)max_grid;
  }
  p->S = p->sectors[0] * p->sectors[1] * p->sectors[2];
  g.sw = (int)sector_width;
  g.kw = (int)kernel_width;
  g.P = (int)sector_width + 2 * (int)std::floor((float)kernel_width / 2.0f);
  g.Pmax = g.P - 1;
  g.off = (int)std::floor((float)g.P / 2.0f);
  if (g.P > 255) {
    delete p;
    return fail(GNUFFT_ERR_UNSUPPORTED, "plan_create: padded sector width above 255");
  }
  g.mP = udiv_magic((uint32_t)g.P);
  g.mP2 = udiv_magic((uint32_t)(g.P * g.P));
  g.mPD = udiv_magic((uint32_t)(g.dim == 3 ? g.P * g.P * g.P : g.P * g.P));
  g.lookup = lookup;
  const long K = lut_size(osf, kernel_width, lookup);
  if (K < 2 || K > 100000) {
    delete p;
    return fail(GNUFFT_ERR_INVALID_ARGUMENT, "plan_create: lookup table size out of range");
  }
  g.K = (int)K;
  // radius constants, CUDA/src/gpuNUFFT_operator.cpp:92-140
  const double kernel_radius = (double)kernel_width / 2.0;
  const double radius = kernel_radius / (double)max_grid;
  const double rsq = radius * radius;
  const double inv = 1.0 / rsq;
  g.kradius = (float)kernel_radius;
  g.rsq = (float)rsq;
  g.rsq_inv = (float)inv;
  g.dist_mult = (float)((double)(K - 1) * inv);
  g.Ng = (long long)g.G[0] * g.G[1] * g.G[2];
  g.Ni = (long long)g.N[0] * g.N[1] * g.N[2];
  if (g.Ng >= (1ll << 31)) {
    delete p;
    return fail(GNUFFT_ERR_UNSUPPORTED, "plan_create: grid above 2^31 points per coil");
  }
  g.fft_scale = 1.0f / (float)std::sqrt((float)g.Ng);
  g.M = 0;
  p->chunk = 1024;

  p->lut = make_lut(K, kernel_width, osf);
  int rc = copy_in((void**)&p->lut_d, p->lut.data(), p->lut.size() * sizeof(float), GNUFFT_HOST, s);
  if (rc == GNUFFT_OK) {
    hipError_t e = hipMalloc(&p->deapo_d, (size_t)g.Ni * sizeof(float));
    if (e != hipSuccess) rc = fail(GNUFFT_ERR_OUT_OF_MEMORY, "plan_create: deapo allocation");
  }
  if (rc == GNUFFT_OK) rc = compute_deapo(p, s);
  if (rc == GNUFFT_OK && fused_fft_supported(g)) {
    const std::vector<float2> tw = fused_fft_twiddles(g, p->fft_tw_off);
    rc = copy_in((void**)&p->fft_tw_d, tw.data(), tw.size() * sizeof(float2), GNUFFT_HOST, s);
    p->fused_ok = rc == GNUFFT_OK;
    p->fused = p->fused_ok && fused_fft_preferred(g);
  }
  if (rc == GNUFFT_OK) {
    std::vector<uint32_t> words;
    size_t base[6];
    build_cover(g, &words, base, p->cover.max_len, &p->cover.max_cov);
    rc = copy_in((void**)&p->cover_d, words.data(), words.size() * sizeof(uint32_t), GNUFFT_HOST, s);
    if (rc == GNUFFT_OK) {
      for (int a = 0; a < 3; a++) {
        p->cover.off[a] = p->cover_d + base[a];
        p->cover.ent[a] = p->cover_d + base[3 + a];
      }
      if (hipStreamSynchronize(s) != hipSuccess) rc = fail(GNUFFT_ERR_DEVICE, "plan_create: cover lists");
    }
  }
  if (rc != GNUFFT_OK) {
    gnufft_plan_destroy(p);
    return rc;
  }
  *plan = p;
  return GNUFFT_OK;
}

int gnufft_plan_set_fft(gnufft_plan p, int kind) {
  if (!p) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "set_fft: null plan");
  p->generation++;
  if (kind == GNUFFT_FFT_ROCFFT) {
    p->fused = false;
    return GNUFFT_OK;
  }
  if (kind != GNUFFT_FFT_FUSED) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "set_fft: unknown kind");
  if (!p->fused_ok)
    return fail(GNUFFT_ERR_UNSUPPORTED,
                "set_fft: fused FFT needs grid lengths with prime factors 2, 3, 5, 7, 11, 13, 17, 19 "
                "up to 2048");
  p->fused = true;
  return GNUFFT_OK;
}


int gnufft_plan_set_workgroups_per_cu(gnufft_plan p, int n) {
  if (!p) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "set_workgroups_per_cu: null plan");
  if (n < 0 || n > 8) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "set_workgroups_per_cu: n must be 0..8");
  p->lane_per_cu = n;
  return GNUFFT_OK;
}

int gnufft_plan_get_fft(gnufft_plan p, int* kind) {
  if (!p || !kind) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "get_fft: null argument");
  *kind = p->fused ? GNUFFT_FFT_FUSED : GNUFFT_FFT_ROCFFT;
  return GNUFFT_OK;
}

int gnufft_plan_set_timing(gnufft_plan p, int enable) {
  if (!p) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "set_timing: null plan");
  p->timer.on = enable != 0;
  return GNUFFT_OK;
}

int gnufft_plan_get_timing(gnufft_plan p, gnufft_timing* out, int reset) {
  if (!p || !out) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "get_timing: null argument");
  p->timer.collect();
  for (int c = 0; c < GNUFFT_T_COUNT; c++) {
    out->ms[c] = p->timer.ms[c];
    out->calls[c] = p->timer.calls[c];
    if (reset) {
      p->timer.ms[c] = 0.0;
      p->timer.calls[c] = 0;
    }
  }
  return GNUFFT_OK;
}

int gnufft_plan_destroy(gnufft_plan p) {
  if (!p) return GNUFFT_OK;
  (void)hipDeviceSynchronize();
  p->timer.collect();
  free_traj(p);
  free_ffts(p);
  dfree(p->lut_d);
  dfree(p->deapo_d);
  dfree(p->cover_d);
  dfree(p->sens_d);
  dfree(p->grid_d);
  dfree(p->img_tmp_d);
  dfree(p->fft_tw_d);
  delete p;
  return GNUFFT_OK;
}

int gnufft_plan_get_info(gnufft_plan p, gnufft_plan_info* info) {
  if (!p || !info) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "plan_get_info: null argument");
  memset(info, 0, sizeof(*info));
  info->dim = p->g.dim;
  for (int a = 0; a < 3; a++) {
    info->img[a] = p->img[a];
    info->grid[a] = p->grid[a];
    info->sectors[a] = p->sectors[a];
  }
  info->sector_count = p->S;
  info->sector_width = p->sw;
  info->kernel_width = p->kw;
  info->kernel_count = (uint32_t)p->g.K;
  info->sector_pad_width = (uint32_t)p->g.P;
  info->sector_offset = (uint32_t)p->g.off;
  info->n_samples = p->M;
  info->n_sens_coils = p->n_sens;
  info->n_chunks = p->n_chunks;
  info->lookup = p->lookup;
  info->has_dens = p->dens_sorted_d != nullptr;
  info->osf = p->osf;
  info->radius_sqr = p->g.rsq;
  info->dist_multiplier = p->g.dist_mult;
  info->grid_count = (uint64_t)p->g.Ng;
  info->img_count = (uint64_t)p->g.Ni;
  return GNUFFT_OK;
}

int gnufft_plan_set_trajectory(gnufft_plan p, const float* traj, uint32_t n_samples,
                               const float* dens, int location, void* stream) {
  if (!p || (!traj && n_samples)) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "set_trajectory: null");
  hipStream_t s = (hipStream_t)stream;
  free_traj(p);
  const uint32_t M = n_samples;
  const int D = p->g.dim;
  float* traj_in = nullptr;
  float* dens_in = nullptr;
  uint32_t *keys = nullptr, *vals = nullptr, *keys_sorted = nullptr;
  void* temp = nullptr;
  int rc = GNUFFT_OK;
  hipError_t e = hipSuccess;
  size_t temp_bytes = 0;
  int key_bits = 1;
  while ((1ull << key_bits) < (unsigned long long)p->S) key_bits++;

  rc = copy_in((void**)&traj_in, traj, (size_t)D * M * sizeof(float), location, s);
  if (rc == GNUFFT_OK && dens) rc = copy_in((void**)&dens_in, dens, (size_t)M * 4, location, s);
  if (rc != GNUFFT_OK) goto done;
  e = hipMalloc(&keys, (size_t)(M ? M : 1) * 4);
  if (e == hipSuccess) e = hipMalloc(&vals, (size_t)(M ? M : 1) * 4);
  if (e == hipSuccess) e = hipMalloc(&keys_sorted, (size_t)(M ? M : 1) * 4);
  if (e == hipSuccess) e = hipMalloc(&p->idx_d, (size_t)(M ? M : 1) * 4);
  if (e == hipSuccess) e = hipMalloc(&p->traj_sorted_d, (size_t)(M ? M : 1) * D * 4);
  if (e == hipSuccess && dens) e = hipMalloc(&p->dens_sorted_d, (size_t)(M ? M : 1) * 4);
  if (e == hipSuccess) e = hipMalloc(&p->sdc_d, (size_t)(p->S + 1) * 4);
  if (e == hipSuccess) e = hipMalloc(&p->centers_d, (size_t)p->S * D * 4);
  if (e != hipSuccess) {
    rc = fail(GNUFFT_ERR_OUT_OF_MEMORY, std::string("set_trajectory: ") + hipGetErrorString(e));
    goto done;
  }
  e = launch_assign_sectors(p->g, traj_in, M, keys, vals, s);
  if (e == hipSuccess && M) e = sort_pairs(nullptr, &temp_bytes, keys, keys_sorted, vals, p->idx_d, M, key_bits, s);
  if (e == hipSuccess) e = hipMalloc(&temp, temp_bytes ? temp_bytes : 4);
  if (e == hipSuccess && M) e = sort_pairs(temp, &temp_bytes, keys, keys_sorted, vals, p->idx_d, M, key_bits, s);
  if (e == hipSuccess) e = launch_sector_data_count(keys_sorted, M, p->S, p->sdc_d, s);
  if (e == hipSuccess) e = launch_sector_centers(p->g, p->centers_d, s);
  if (e == hipSuccess) e = launch_gather_traj(traj_in, dens_in, p->idx_d, M, D, p->traj_sorted_d, p->dens_sorted_d, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    rc = fail(GNUFFT_ERR_DEVICE, std::string("set_trajectory: ") + hipGetErrorString(e));
    goto done;
  }
  p->assigned_d = keys;
  keys = nullptr;
  p->M = M;
  p->g.M = (int)M;
  rc = build_chunks(p, s);
done:
  dfree(traj_in);
  dfree(dens_in);
  dfree(keys);
  dfree(vals);
  dfree(keys_sorted);
  dfree(temp);
  if (rc != GNUFFT_OK) free_traj(p);
  return rc;
}

int gnufft_check_precomputed(const uint32_t img_dims[3], uint32_t sector_width, float osf,
                             const float* traj_sorted, const uint32_t* data_indices,
                             const uint32_t* sector_data_count, const uint32_t* sector_centers,
                             uint32_t n_samples) {
  if (!img_dims || !sector_data_count || !sector_centers || ((!traj_sorted || !data_indices) && n_samples))
    return fail(GNUFFT_ERR_INVALID_ARGUMENT, "check_precomputed: null argument");
  if (sector_width == 0 || !(osf >= 1.0f) || img_dims[0] == 0 || img_dims[1] == 0)
    return fail(GNUFFT_ERR_INVALID_ARGUMENT, "check_precomputed: bad geometry");
  const int D = img_dims[2] == 0 ? 2 : 3;
  const uint32_t M = n_samples;
  int G[3] = {1, 1, 1};
  uint32_t S[3] = {1, 1, 1};
  for (int a = 0; a < D; a++) {
    G[a] = (int)(uint32_t)(img_dims[a] * osf);  // gpuNUFFT_types.hpp:109-117
    if (G[a] <= 0) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "check_precomputed: empty grid");
    S[a] = (uint32_t)std::ceil((float)G[a] / (float)sector_width);
  }
  const uint64_t nS = (uint64_t)S[0] * S[1] * S[2];
  // sector offsets: length S+1, starting at 0, non-decreasing, ending at M
  // (computeSectorDataCount, gpuNUFFT_operator_factory.cpp:183-205)
  if (sector_data_count[0] != 0)
    return fail(GNUFFT_ERR_INVALID_ARGUMENT, "check_precomputed: sector_data_count[0] != 0");
  for (uint64_t s = 0; s < nS; s++)
    if (sector_data_count[s + 1] < sector_data_count[s])
      return fail(GNUFFT_ERR_INVALID_ARGUMENT,
                  "check_precomputed: sector_data_count decreases at sector " + std::to_string(s));
  if (sector_data_count[nS] != M)
    return fail(GNUFFT_ERR_INVALID_ARGUMENT, "check_precomputed: sector_data_count[S] != sample count");
  // data indices: a permutation of [0, M)
  {
    std::vector<uint8_t> seen(M, 0);
    for (uint32_t i = 0; i < M; i++) {
      const uint32_t v = data_indices[i];
      if (v >= M || seen[v])
        return fail(GNUFFT_ERR_INVALID_ARGUMENT,
                    "check_precomputed: data_indices is not a permutation of [0, M) (entry " +
                        std::to_string(i) + ")");
      seen[v] = 1;
    }
  }
  // centers in the standard layout (gpuNUFFT_operator_factory.cpp:207-270):
  // the kernels derive padded-sector positions from it
  for (uint64_t sct = 0; sct < nS; sct++) {
    const uint32_t idx[3] = {(uint32_t)(sct % S[0]), (uint32_t)((sct / S[0]) % S[1]),
                             (uint32_t)(sct / ((uint64_t)S[0] * S[1]))};
    for (int a = 0; a < D; a++)
      if (sector_centers[sct * D + a] != idx[a] * sector_width + sector_width / 2)
        return fail(GNUFFT_ERR_INVALID_ARGUMENT,
                    "check_precomputed: sector centers differ from the plan's sector layout");
  }
  // every sorted sample lies in the sector its range names (E3 assignment,
  // CUDA/inc/precomp_utils.hpp:34-48, :142-153)
  for (uint64_t sct = 0; sct < nS; sct++)
    for (uint32_t i = sector_data_count[sct]; i < sector_data_count[sct + 1]; i++) {
      uint64_t lin = 0, mul = 1;
      for (int a = 0; a < D; a++) {
        lin += mul * (uint64_t)host_sector_of(traj_sorted[(size_t)a * M + i], G[a], (int)sector_width);
        mul *= S[a];
      }
      if (lin != sct)
        return fail(GNUFFT_ERR_INVALID_ARGUMENT,
                    "check_precomputed: sorted sample " + std::to_string(i) + " lies outside sector " +
                        std::to_string(sct));
    }
  return GNUFFT_OK;
}

int gnufft_plan_load_precomputed(gnufft_plan p, const float* traj_sorted,
                                 const uint32_t* data_indices, const uint32_t* sector_data_count,
                                 const uint32_t* sector_centers, const float* dens_sorted,
                                 const float* deapo, uint32_t n_samples, int location,
                                 void* stream) {
  if (!p || !sector_data_count || !sector_centers || ((!traj_sorted || !data_indices) && n_samples))
    return fail(GNUFFT_ERR_INVALID_ARGUMENT, "load_precomputed: null argument");
  hipStream_t s = (hipStream_t)stream;
  const uint32_t M = n_samples;
  const int D = p->g.dim;
  {
    // validate on the host before anything reaches the kernels: a bad index
    // or offset table would otherwise become out-of-range device accesses
    std::vector<float> tr;
    std::vector<uint32_t> di, sdc, cen;
    const float* tr_h = traj_sorted;
    const uint32_t *di_h = data_indices, *sdc_h = sector_data_count, *cen_h = sector_centers;
    if (location == GNUFFT_DEVICE) {
      tr.resize((size_t)D * M);
      di.resize(M);
      sdc.resize((size_t)p->S + 1);
      cen.resize((size_t)p->S * D);
      if (M) {
        HIP_TRY(hipMemcpy(tr.data(), traj_sorted, tr.size() * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(di.data(), data_indices, di.size() * 4, hipMemcpyDeviceToHost));
      }
      HIP_TRY(hipMemcpy(sdc.data(), sector_data_count, sdc.size() * 4, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(cen.data(), sector_centers, cen.size() * 4, hipMemcpyDeviceToHost));
      tr_h = tr.data();
      di_h = di.data();
      sdc_h = sdc.data();
      cen_h = cen.data();
    }
    const uint32_t dims[3] = {p->img[0], p->img[1], D == 3 ? p->img[2] : 0u};
    int rc = gnufft_check_precomputed(dims, p->sw, p->osf, tr_h, di_h, sdc_h, cen_h, M);
    if (rc) return rc;
  }
  free_traj(p);
  int rc = copy_in((void**)&p->traj_sorted_d, traj_sorted, (size_t)D * M * 4, location, s);
  if (!rc) rc = copy_in((void**)&p->idx_d, data_indices, (size_t)M * 4, location, s);
  if (!rc) rc = copy_in((void**)&p->sdc_d, sector_data_count, (size_t)(p->S + 1) * 4, location, s);
  if (!rc) rc = copy_in((void**)&p->centers_d, sector_centers, (size_t)p->S * D * 4, location, s);
  if (!rc && dens_sorted) rc = copy_in((void**)&p->dens_sorted_d, dens_sorted, (size_t)M * 4, location, s);
  if (!rc && deapo)
    HIP_TRY(hipMemcpyAsync(p->deapo_d, deapo, (size_t)p->g.Ni * 4,
                           location == GNUFFT_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                           s));
  if (!rc && hipStreamSynchronize(s) != hipSuccess)
    rc = fail(GNUFFT_ERR_DEVICE, "load_precomputed: table upload");
  if (!rc) {
    p->M = M;
    p->g.M = (int)M;
    rc = build_chunks(p, s);
  }
  if (rc) free_traj(p);
  return rc;
}

int gnufft_plan_set_sens(gnufft_plan p, const void* sens, uint32_t n_coils, int location,
                         void* stream) {
  if (!p) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "set_sens: null plan");
  p->generation++;
  p->table_generation++;
  dfree(p->sens_d);
  p->sens_d = nullptr;
  p->n_sens = 0;
  if (!sens || n_coils == 0) return GNUFFT_OK;
  int rc = copy_in((void**)&p->sens_d, sens, (size_t)n_coils * p->g.Ni * sizeof(float2), location,
                   (hipStream_t)stream);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  p->n_sens = n_coils;
  return GNUFFT_OK;
}

int gnufft_plan_get_tables(gnufft_plan p, uint32_t* data_indices, uint32_t* sector_data_count,
                           uint32_t* sector_centers, float* traj_sorted, float* dens_sorted,
                           float* deapo, float* lut) {
  if (!p) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "get_tables: null plan");
  const int D = p->g.dim;
  if ((data_indices || sector_data_count || sector_centers || traj_sorted || dens_sorted) &&
      !p->sdc_d)
    return fail(GNUFFT_ERR_STATE, "get_tables: no trajectory set");
  if (dens_sorted && !p->dens_sorted_d) return fail(GNUFFT_ERR_STATE, "get_tables: no density compensation");
  const uint32_t M = p->M;
  if (sector_data_count) HIP_TRY(hipMemcpy(sector_data_count, p->sdc_d, (size_t)(p->S + 1) * 4, hipMemcpyDeviceToHost));
  if (sector_centers) HIP_TRY(hipMemcpy(sector_centers, p->centers_d, (size_t)p->S * D * 4, hipMemcpyDeviceToHost));
  if ((data_indices || traj_sorted || dens_sorted) && M) {
    // The kernels run on the device sort's order (stable within a sector).
    // A computed plan reports the reference's order instead: std::sort of
    // (index, sector) pairs compared by sector only, which libstdc++'s
    // introsort leaves in an implementation-defined order within sectors
    // (CUDA/src/gpuNUFFT_operator_factory.cpp:56-72, comparator
    // CUDA/inc/gpuNUFFT_types.hpp:33-49; tables gathered as at :499-522).
    // Loaded plans report the tables they were given.
    std::vector<uint32_t> idx(M);
    HIP_TRY(hipMemcpy(idx.data(), p->idx_d, (size_t)M * 4, hipMemcpyDeviceToHost));
    std::vector<uint32_t> ref;  // reference position -> device-order position
    if (p->assigned_d) {
      std::vector<uint32_t> sec(M);
      HIP_TRY(hipMemcpy(sec.data(), p->assigned_d, (size_t)M * 4, hipMemcpyDeviceToHost));
      std::vector<std::pair<uint32_t, uint32_t>> v;
      v.reserve(M);
      for (uint32_t i = 0; i < M; i++) v.emplace_back(i, sec[i]);
      std::sort(v.begin(), v.end(), [](const std::pair<uint32_t, uint32_t>& a,
                                       const std::pair<uint32_t, uint32_t>& b) {
        return a.second < b.second;
      });
      std::vector<uint32_t> pos(M);  // caller index -> device-order position
      for (uint32_t i = 0; i < M; i++) pos[idx[i]] = i;
      ref.resize(M);
      for (uint32_t i = 0; i < M; i++) ref[i] = pos[v[i].first];
    }
    auto reorder = [&](const float* dev, uint32_t rows, float* out) -> int {
      std::vector<float> h((size_t)rows * M);
      HIP_TRY(hipMemcpy(h.data(), dev, h.size() * 4, hipMemcpyDeviceToHost));
      for (uint32_t r = 0; r < rows; r++)
        for (uint32_t i = 0; i < M; i++)
          out[(size_t)r * M + i] = h[(size_t)r * M + (ref.empty() ? i : ref[i])];
      return GNUFFT_OK;
    };
    if (data_indices)
      for (uint32_t i = 0; i < M; i++) data_indices[i] = idx[ref.empty() ? i : ref[i]];
    if (traj_sorted) {
      int rc = reorder(p->traj_sorted_d, (uint32_t)D, traj_sorted);
      if (rc) return rc;
    }
    if (dens_sorted) {
      int rc = reorder(p->dens_sorted_d, 1, dens_sorted);
      if (rc) return rc;
    }
  }
  if (deapo) HIP_TRY(hipMemcpy(deapo, p->deapo_d, (size_t)p->g.Ni * 4, hipMemcpyDeviceToHost));
  if (lut) memcpy(lut, p->lut.data(), p->lut.size() * sizeof(float));
  return GNUFFT_OK;
}

int gnufft_plan_get_assigned_sectors(gnufft_plan p, uint32_t* sectors_host) {
  if (!p || !sectors_host) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "get_assigned_sectors: null");
  if (!p->assigned_d) return fail(GNUFFT_ERR_STATE, "get_assigned_sectors: tables were loaded, not computed");
  if (p->M) HIP_TRY(hipMemcpy(sectors_host, p->assigned_d, (size_t)p->M * 4, hipMemcpyDeviceToHost));
  return GNUFFT_OK;
}

int gnufft_spread_sorted(gnufft_plan p, const void* data_sorted, void* grid, uint32_t n_coils,
                         void* stream) {
  if (!p || !grid || (!data_sorted && p->M)) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "spread_sorted: null");
  if (!p->sdc_d) return fail(GNUFFT_ERR_STATE, "spread_sorted: no trajectory set");
  hipStream_t s = (hipStream_t)stream;
  const uint32_t cc = coil_batch(p, n_coils);
  int rc = ensure_scratch(p, cc);
  if (rc) return rc;
  for (uint32_t c0 = 0; c0 < n_coils; c0 += cc) {
    const uint32_t nc = std::min(cc, n_coils - c0);
    SpreadArgs a = spread_args(p, (const float2*)data_sorted + (size_t)c0 * p->M,
                               (float2*)grid + (size_t)c0 * p->g.Ng, nc);
    a.fold_shift = false;
    TIMED(p, GNUFFT_T_SPREAD, s, launch_spread(p->g, a, (int)p->round_target, s));
  }
  return GNUFFT_OK;
}

int gnufft_interp_sorted(gnufft_plan p, const void* grid, void* data_sorted, uint32_t n_coils,
                         void* stream) {
  if (!p || !grid || (!data_sorted && p->M)) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "interp_sorted: null");
  if (!p->sdc_d) return fail(GNUFFT_ERR_STATE, "interp_sorted: no trajectory set");
  InterpArgs a{};
  a.grid = (const float2*)grid;
  a.crds = p->traj_sorted_d;
  a.work = p->work;
  a.work_runs = p->work_runs;
  a.lane_per_cu = p->lane_per_cu;
  a.lut = p->lut_d;
  a.out = (float2*)data_sorted;
  a.coil_stride = p->M;
  a.scale = 1.0f;
  a.n_coils = (int)n_coils;
  TIMED(p, GNUFFT_T_INTERP, (hipStream_t)stream, launch_interp(p->g, a, (int)p->round_target, (hipStream_t)stream));
  return GNUFFT_OK;
}

int gnufft_adjoint(gnufft_plan p, const void* kdata, void* out, uint32_t n_coils, int stop,
                   void* stream) {
  if (!p || !out || (!kdata && p->M)) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "adjoint: null argument");
  if (!valid_stop(stop)) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "adjoint: unknown stop stage");
  if (!p->sdc_d) return fail(GNUFFT_ERR_STATE, "adjoint: no trajectory set");
  if (n_coils == 0) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "adjoint: zero coils");
  const bool use_sens = p->n_sens > 0 && stop == GNUFFT_STOP_DEAPODIZATION;
  if (p->n_sens > 0 && p->n_sens != n_coils)
    return fail(GNUFFT_ERR_INVALID_ARGUMENT, "adjoint: coil count differs from sensitivity coils");
  hipStream_t s = (hipStream_t)stream;
  const Geom& g = p->g;
  const bool grid_out = stop == GNUFFT_STOP_CONVOLUTION || stop == GNUFFT_STOP_DENSITY_ESTIMATION;
  const uint32_t cc = coil_batch(p, n_coils);
  if (!grid_out) {
    int rc = ensure_grid(p, cc);
    if (rc) return rc;
  }
  {
    int rc = ensure_scratch(p, cc);
    if (rc) return rc;
  }
  for (uint32_t c0 = 0; c0 < n_coils; c0 += cc) {
    const uint32_t nc = std::min(cc, n_coils - c0);
    float2* grid = grid_out ? (float2*)out + (size_t)c0 * g.Ng : p->grid_d;
    {
      SpreadArgs a = spread_args(p, (const float2*)kdata + (size_t)c0 * p->M, grid, nc);
      a.gather = p->idx_d;
      a.dens = p->dens_sorted_d;
      a.fold_shift = !grid_out;
      TIMED(p, GNUFFT_T_SPREAD, s, launch_spread(g, a, (int)p->round_target, s));
    }
    if (grid_out) continue;
    if (p->fused) {
      float2* dst = use_sens ? (float2*)out : (float2*)out + (size_t)c0 * g.Ni;
      TIMED(p, GNUFFT_T_FFT, s,
            fused_fft_adjoint(g, grid, p->scratch_d, (int)nc, p->deapo_d,
                              use_sens ? p->sens_d + (size_t)c0 * g.Ni : nullptr, dst, stop,
                              use_sens && c0 > 0, p->fft_tw_d, p->fft_tw_off, s));
      continue;
    }
    int rc = run_fft(p, nc, true, grid, s);
    if (rc) return rc;
    if (use_sens) {
      TIMED(p, GNUFFT_T_PREPOST, s, launch_adjoint_post(g, grid, (int)nc, p->deapo_d, p->sens_d + (size_t)c0 * g.Ni,
                                  (float2*)out, c0 > 0, stop, s));
    } else {
      TIMED(p, GNUFFT_T_PREPOST, s, launch_adjoint_post(g, grid, (int)nc, p->deapo_d, nullptr,
                                  (float2*)out + (size_t)c0 * g.Ni, false, stop, s));
    }
  }
  return GNUFFT_OK;
}

int gnufft_forward(gnufft_plan p, const void* img, void* out, uint32_t n_coils, int stop,
                   void* stream) {
  if (!p || !img || (!out && p->M)) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "forward: null argument");
  if (!valid_stop(stop)) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "forward: unknown stop stage");
  if (!p->sdc_d) return fail(GNUFFT_ERR_STATE, "forward: no trajectory set");
  if (n_coils == 0) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "forward: zero coils");
  const bool use_sens = p->n_sens > 0;
  if (use_sens && p->n_sens != n_coils)
    return fail(GNUFFT_ERR_INVALID_ARGUMENT, "forward: coil count differs from sensitivity coils");
  hipStream_t s = (hipStream_t)stream;
  const Geom& g = p->g;
  if (stop == GNUFFT_STOP_DENSITY_ESTIMATION) {
    // interpolation from the deapodized image used as the grid
    // (gpuNUFFT_operator.cpp:1178-1187): defined only for grid == image size
    if (g.Ng != g.Ni)
      return fail(GNUFFT_ERR_UNSUPPORTED, "forward DENSITY_ESTIMATION needs grid == image size (osf 1)");
    if (p->img_tmp_coils < n_coils) {
      p->generation++;
      dfree(p->img_tmp_d);
      p->img_tmp_d = nullptr;
      p->img_tmp_coils = 0;
      HIP_TRY(hipMalloc(&p->img_tmp_d, (size_t)n_coils * g.Ni * sizeof(float2)));
      p->img_tmp_coils = n_coils;
    }
    TIMED(p, GNUFFT_T_PREPOST, s, launch_image_prep(g, (const float2*)img, !use_sens, p->sens_d, p->deapo_d,
                              p->img_tmp_d, (int)n_coils, s));
    InterpArgs a{};
    a.grid = p->img_tmp_d;
    a.crds = p->traj_sorted_d;
    a.work = p->work;
    a.work_runs = p->work_runs;
    a.lane_per_cu = p->lane_per_cu;
    a.lut = p->lut_d;
    a.out = (float2*)out;
    a.coil_stride = p->M;
    a.scatter = p->idx_d;
    a.scale = 1.0f;
    a.n_coils = (int)n_coils;
    TIMED(p, GNUFFT_T_INTERP, s, launch_interp(g, a, (int)p->round_target, s));
    return GNUFFT_OK;
  }
  const uint32_t cc = coil_batch(p, n_coils);
  int rc = ensure_grid(p, cc);
  if (rc) return rc;
  if (p->fused) {
    rc = ensure_scratch(p, cc);
    if (rc) return rc;
  }
  for (uint32_t c0 = 0; c0 < n_coils; c0 += cc) {
    const uint32_t nc = std::min(cc, n_coils - c0);
    const float2* src = use_sens ? (const float2*)img : (const float2*)img + (size_t)c0 * g.Ni;
    const float2* sens = use_sens ? p->sens_d + (size_t)c0 * g.Ni : nullptr;
    if (p->fused) {
      TIMED(p, GNUFFT_T_FFT, s,
            fused_fft_forward(g, src, !use_sens, sens, p->deapo_d, p->grid_d, p->scratch_d,
                              (int)nc, p->fft_tw_d, p->fft_tw_off, s));
    } else {
      TIMED(p, GNUFFT_T_PREPOST, s, launch_forward_pre(g, src, !use_sens, sens, p->deapo_d,
                                                       p->grid_d, (int)nc, s));
      rc = run_fft(p, nc, false, p->grid_d, s);
      if (rc) return rc;
    }
    InterpArgs a{};
    a.grid = p->grid_d;
    a.crds = p->traj_sorted_d;
    a.work = p->work;
    a.work_runs = p->work_runs;
    a.lane_per_cu = p->lane_per_cu;
    a.lut = p->lut_d;
    a.out = (float2*)out + (size_t)c0 * p->M;
    a.coil_stride = p->M;
    a.scatter = p->idx_d;
    a.dens = p->dens_sorted_d;
    a.scale = g.fft_scale;
    a.n_coils = (int)nc;
    a.fold_shift = true;
    TIMED(p, GNUFFT_T_INTERP, s, launch_interp(g, a, (int)p->round_target, s));
  }
  return GNUFFT_OK;
}

// ---------------------------------------------------------------------------
// Batches of independent plans
// ---------------------------------------------------------------------------
}  // extern "C"

struct gnufft_batch_s {
  int dim = 2, nx = 4;
  std::vector<gnufft_plan> plans;  // in batch order (grouped by grid size)
  std::vector<uint32_t> coils;
  std::vector<uint32_t> caller_index;  // batch position -> caller's plan index
  std::vector<uint64_t> table_gen;     // each plan's table generation at creation
  std::vector<BatchOp> ops_h;
  BatchOp* ops_d = nullptr;
  uint4* items_d = nullptr;
  BatchList spread{}, interp{}, fold{}, reduce{}, post{}, pre{};
  // lane kernels: chunks of at most 2 * kBatchSmallThreads samples, run in
  // small workgroups (spread and interp share the lists)
  BatchList spread_small{}, interp_small{};
  size_t spread_lds = 0, interp_lds = 0, spread_small_lds = 0, interp_small_lds = 0;
  int fold_threads = 64;
  bool fold_generic = false;  // some plan needs the fold's generic body
  uint32_t max_kstride = 0;
  float2* grid_d = nullptr;
  float2* scratch_d = nullptr;
  struct Group {
    uint32_t first, count;  // plans [first, first + count) in batch order
    size_t offset;          // grid elements before the group
    FftPlans fft;
  };
  std::vector<Group> groups;
  // plans on GNUFFT_FFT_FUSED: one descriptor per plan and one item list
  // (plan, block) per pass; the other plans: rocFFT per grid size with the
  // pre / post kernels
  bool fused = false;            // any plan on the fused FFT
  bool fft_prime = false;        // a fused plan needs the prime stages 7 .. 19
  std::vector<char> op_fused;
  std::vector<char> fft_desc_h;
  void* fft_desc_d = nullptr;
  uint2* fft_items_d = nullptr;
  struct FftPass {
    const uint2* items = nullptr;
    uint32_t n = 0;
    size_t lds = 0;
  };
  FftPass fft_adj[3], fft_fwd[3];          // passes on the standard tiles
  FftPass fft_adj_big[3], fft_fwd_big[3];  // y / z passes of long lines (large tiles)
  hipEvent_t uploaded = nullptr;  // last descriptor upload (ops_h may be reused after it)
  hipEvent_t done = nullptr;      // end of the last call's device work, on its stream
};

namespace {

void free_batch(gnufft_batch b) {
  if (b->uploaded) (void)hipEventDestroy(b->uploaded);
  if (b->done) (void)hipEventDestroy(b->done);
  dfree(b->ops_d);
  dfree(b->items_d);
  dfree(b->grid_d);
  dfree(b->scratch_d);
  dfree(b->fft_desc_d);
  dfree(b->fft_items_d);
  for (auto& gr : b->groups) destroy_fft(gr.fft);
  delete b;
}

// The batch holds its plans' trajectory tables and sensitivities by
// pointer: a plan whose tables were replaced since batch creation makes
// the batch unusable.
int batch_check_plans(gnufft_batch b) {
  for (size_t k = 0; k < b->plans.size(); k++)
    if (b->plans[k]->table_generation != b->table_gen[k])
      return fail(GNUFFT_ERR_STATE,
                  "batch: plan " + std::to_string(b->caller_index[k]) +
                      " changed its trajectory or sensitivities after batch_create; create the batch again");
  return GNUFFT_OK;
}

// Wait until the previous descriptor upload has read ops_h, and order this
// call's device work after the previous call's (its descriptors, work grid
// and scratch are shared), whichever stream that one ran on.
int batch_wait_upload(gnufft_batch b, hipStream_t s) {
  if (b->uploaded) HIP_TRY(hipEventSynchronize(b->uploaded));
  if (b->done) HIP_TRY(hipStreamWaitEvent(s, b->done, 0));
  return GNUFFT_OK;
}

int batch_mark_done(gnufft_batch b, hipStream_t s) {
  if (!b->done) HIP_TRY(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(b->done, s));
  return GNUFFT_OK;
}

int batch_upload_ops(gnufft_batch b, hipStream_t s) {
  HIP_TRY(hipMemcpyAsync(b->ops_d, b->ops_h.data(), b->ops_h.size() * sizeof(BatchOp),
                         hipMemcpyHostToDevice, s));
  if (b->fused)
    HIP_TRY(hipMemcpyAsync(b->fft_desc_d, b->fft_desc_h.data(), b->fft_desc_h.size(),
                           hipMemcpyHostToDevice, s));
  if (!b->uploaded) HIP_TRY(hipEventCreateWithFlags(&b->uploaded, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(b->uploaded, s));
  return GNUFFT_OK;
}

}  // namespace

extern "C" {

int gnufft_batch_create(gnufft_batch* batch, const gnufft_plan* plans, uint32_t n_plans,
                        const uint32_t* n_coils, void* stream) {
  if (!batch || !plans || !n_coils || n_plans == 0)
    return fail(GNUFFT_ERR_INVALID_ARGUMENT, "batch_create: null argument or empty batch");
  *batch = nullptr;
  hipStream_t s = (hipStream_t)stream;
  const Geom& g0 = plans[0]->g;
  const int nx = g0.kw + 1 <= 4 ? 4 : 8;
  for (uint32_t i = 0; i < n_plans; i++) {
    const gnufft_plan p = plans[i];
    if (!p || !p->sdc_d) return fail(GNUFFT_ERR_STATE, "batch_create: plan without trajectory");
    if (p->g.dim != g0.dim || (p->g.kw + 1 <= 4 ? 4 : 8) != nx)
      return fail(GNUFFT_ERR_UNSUPPORTED,
                  "batch_create: plans must share the dimension and the kernel-width class");
    if (n_coils[i] == 0) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "batch_create: zero coils");
    if (p->n_sens > 0 && p->n_sens != n_coils[i])
      return fail(GNUFFT_ERR_INVALID_ARGUMENT, "batch_create: coils differ from sensitivity coils");
  }
  gnufft_batch b = new (std::nothrow) gnufft_batch_s();
  if (!b) return fail(GNUFFT_ERR_OUT_OF_MEMORY, "batch_create: host allocation");
  // plans on the fused FFT run it inside the batch (with the kernels'
  // prime-stage variants when some plan needs them); the others take rocFFT
  // and the separate pre / post kernels here
  auto batch_fused = [](const gnufft_plan q) { return q->fused && fused_fft_supported(q->g); };
  b->dim = g0.dim;
  b->nx = nx;
  // kernel width 3 in every plan: the lane kernels (tap records in
  // registers), flagged as nx = 3 for the batched launchers
  bool lane = true;
  for (uint32_t i = 0; i < n_plans; i++) lane = lane && lane_batch_ok(plans[i]->g);
  if (lane) b->nx = 3;
  const uint32_t grid_cb = lane ? (uint32_t)kBatchLaneCB : (uint32_t)kBatchCB;
  b->fused = false;
  b->fft_prime = false;
  for (uint32_t i = 0; i < n_plans; i++) {
    b->fused = b->fused || batch_fused(plans[i]);
    b->fft_prime = b->fft_prime || (batch_fused(plans[i]) && fused_fft_prime(plans[i]->g));
  }
  // batch order: grouped by grid size (one rocFFT plan per group)
  std::vector<uint32_t> order(n_plans);
  for (uint32_t i = 0; i < n_plans; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) {
    const Geom &ga = plans[a]->g, &gc = plans[c]->g;
    if (ga.G[2] != gc.G[2]) return ga.G[2] < gc.G[2];
    if (ga.G[1] != gc.G[1]) return ga.G[1] < gc.G[1];
    return ga.G[0] < gc.G[0];
  });
  // One launch per stage uses the largest footprint of any plan, so each
  // plan's round size is the largest that still lets kBatchPerCU workgroups
  // share a CU (then fewer): one dense plan must not cut the occupancy of
  // the whole batch.
  auto batch_round = [](int (*round)(const Geom&, int, int, size_t), const Geom& g, int target,
                        size_t reserve) {
    const int need = std::min(64, target);
    for (int k = kBatchPerCU; k >= 1; k--) {
      const int r = round(g, kBatchCB, target, lds_budget() / k - 1024 - reserve);
      if (r >= need || (k == 1 && r > 0)) return r;
    }
    return 0;
  };
  size_t grid_total = 0, scratch_total = 0;
  std::vector<size_t> grid_off(n_plans), scratch_off(n_plans);
  for (uint32_t k = 0; k < n_plans; k++) {
    const gnufft_plan p = plans[order[k]];
    const uint32_t C = n_coils[order[k]];
    b->plans.push_back(p);
    b->coils.push_back(C);
    b->caller_index.push_back(order[k]);
    b->table_gen.push_back(p->table_generation);
    b->op_fused.push_back(batch_fused(p) ? 1 : 0);
    grid_off[k] = grid_total;
    grid_total += (size_t)C * p->g.Ng;
    scratch_off[k] = scratch_total;
    // spread sub-grids, then (fused FFT) the transform's intermediate array
    scratch_total += std::max((size_t)p->n_chunks * C * padded_points(p->g),
                              batch_fused(p) ? (size_t)C * fused_fft_tmp_elems(p->g) : (size_t)0);
    // rocFFT groups: runs of consecutive non-fused plans of one grid size
    if (!batch_fused(p)) {
      const gnufft_plan q = k ? plans[order[k - 1]] : nullptr;
      if (!q || batch_fused(q) || p->g.G[0] != q->g.G[0] || p->g.G[1] != q->g.G[1] ||
          p->g.G[2] != q->g.G[2])
        b->groups.push_back({k, 0, grid_off[k], FftPlans()});
      b->groups.back().count++;
    }
  }
  int rc = GNUFFT_OK;
  if (hipMalloc(&b->grid_d, std::max<size_t>(grid_total, 1) * sizeof(float2)) != hipSuccess ||
      hipMalloc(&b->scratch_d, std::max<size_t>(scratch_total, 1) * sizeof(float2)) != hipSuccess) {
    free_batch(b);
    return fail(GNUFFT_ERR_OUT_OF_MEMORY, "batch_create: grid / scratch allocation");
  }
  // descriptors and item lists
  // (chunk, coil block) items of large (sp) and small (ss) chunks
  std::vector<uint4> sp, ss, rd, fo, po, pr;
  std::vector<uint32_t> sp_cost, ss_cost, fo_cost;
  for (uint32_t k = 0; k < n_plans && !rc; k++) {
    const gnufft_plan p = b->plans[k];
    const uint32_t C = b->coils[k];
    const uint32_t cbs = (C + grid_cb - 1) / grid_cb;  // spread / interp coil blocks
    BatchOp op{};
    op.g = p->g;
    op.s = spread_args(p, nullptr, b->grid_d + grid_off[k], C);
    op.s.scratch = b->scratch_d + scratch_off[k];
    op.s.gather = p->idx_d;
    op.s.dens = p->dens_sorted_d;
    op.s.fold_shift = true;
    op.s.n_reduce = 0;
    op.s.sub_cb = lane ? kBatchLaneCB : kBatchCB;  // the batched spread's coil block
    op.i.grid = b->grid_d + grid_off[k];
    op.i.crds = p->traj_sorted_d;
    op.i.work = p->work;
    op.i.lut = p->lut_d;
    op.i.coil_stride = p->M;
    op.i.scatter = p->idx_d;
    op.i.dens = p->dens_sorted_d;
    op.i.scale = p->g.fft_scale;
    op.i.n_coils = (int)C;
    op.i.fold_shift = true;
    op.post = PostArgs{b->grid_d + grid_off[k], (int)C, p->deapo_d,
                       p->n_sens ? p->sens_d : nullptr, nullptr, GNUFFT_STOP_DEAPODIZATION};
    op.pre = PreArgs{nullptr, p->n_sens ? 0 : 1, p->n_sens ? p->sens_d : nullptr, p->deapo_d,
                     b->grid_d + grid_off[k], (int)C};
    op.R_spread = batch_round(spread_round, p->g, (int)p->round_target, 0);
    op.R_interp = batch_round(interp_round, p->g, (int)p->round_target, 3072);
    if (op.R_spread <= 0 || op.R_interp <= 0) {
      rc = fail(GNUFFT_ERR_UNSUPPORTED, "batch_create: padded sector does not fit in LDS");
      break;
    }
    if (lane) {
      b->spread_lds = std::max(b->spread_lds, lane_lds_bytes(p->g, kBatchLaneCB));
      b->spread_small_lds = std::max(b->spread_small_lds, lane_lds_bytes(p->g, kBatchLaneCB, false));
      b->interp_small_lds = b->spread_small_lds;
      b->interp_lds = std::max(b->interp_lds, lane_lds_bytes(p->g, kBatchLaneCB));
    } else {
      b->spread_lds = std::max(b->spread_lds, spread_lds_bytes(p->g, kBatchCB, op.R_spread));
      b->interp_lds = std::max(b->interp_lds, interp_lds_bytes(p->g, kBatchCB, op.R_interp));
    }
    const int ft = fold_threads(p->g, p->cover);
    if (ft == 0) {
      rc = fail(GNUFFT_ERR_UNSUPPORTED, "batch_create: too many fold candidates");
      break;
    }
    b->fold_threads = std::max(b->fold_threads, ft);
    b->fold_generic = b->fold_generic || fold_generic(p->cover);
    b->max_kstride = std::max<uint32_t>(b->max_kstride, C * (uint32_t)padded_points(p->g));
    b->ops_h.push_back(op);
    if (b->fused) b->fft_desc_h.resize((k + 1) * fused_fft_desc_bytes());
    if (batch_fused(p)) {
      const size_t db = fused_fft_desc_bytes();
      if (!fused_fft_desc_build(b->fft_desc_h.data() + k * db, p->g, b->grid_d + grid_off[k],
                                b->scratch_d + scratch_off[k], (int)C, p->deapo_d,
                                p->n_sens ? p->sens_d : nullptr, p->fft_tw_d, p->fft_tw_off)) {
        rc = fail(GNUFFT_ERR_UNSUPPORTED, "batch_create: fused FFT descriptor");
        break;
      }
    }
    // chunk list (host copy) for the gridding items and the reduce tasks
    std::vector<Chunk> ch(p->n_chunks);
    if (p->n_chunks &&
        hipMemcpy(ch.data(), p->chunks_d, ch.size() * sizeof(Chunk), hipMemcpyDeviceToHost) != hipSuccess) {
      rc = fail(GNUFFT_ERR_DEVICE, "batch_create: chunk list copy");
      break;
    }
    for (uint32_t ci = 0; ci < p->n_chunks; ci++) {
      const uint32_t n = ch[ci].end - ch[ci].begin;
      const bool small = lane && n <= 2u * (uint32_t)kBatchSmallThreads;
      for (uint32_t cb = 0; cb < cbs; cb++) {
        (small ? ss : sp).push_back(make_uint4(k, ci, cb, 0));
        (small ? ss_cost : sp_cost).push_back(n + (small ? 100u : 300u));
      }
    }
    for (uint32_t i = 0; i < p->n_chunks;) {
      uint32_t j = i + 1;
      while (j < p->n_chunks && ch[j].sector == ch[i].sector) j++;
      if (j - i > 1) rd.push_back(make_uint4(k, i, j - i, 0));
      i = j;
    }
    const uint32_t fcbs = (C + kBatchCB - 1) / kBatchCB;  // fold coil blocks
    for (uint32_t sct = 0; sct < p->S; sct++)
      for (uint32_t cb = 0; cb < fcbs; cb++) {
        fo.push_back(make_uint4(k, sct, cb, 0));
        fo_cost.push_back(1u);
      }
    if (!batch_fused(p)) {  // pre / post run beside rocFFT; the fused FFT has them inside
      for (uint64_t blk = 0; blk * 256 < (uint64_t)p->g.Ni; blk++) po.push_back(make_uint4(k, (uint32_t)blk, 0, 0));
      for (uint64_t blk = 0; blk * 256 < (uint64_t)p->g.Ng; blk++) pr.push_back(make_uint4(k, (uint32_t)blk, 0, 0));
    }
  }
  if (!rc) {
    std::vector<uint4> all;
    all.insert(all.end(), sp.begin(), sp.end());
    all.insert(all.end(), ss.begin(), ss.end());
    all.insert(all.end(), rd.begin(), rd.end());
    all.insert(all.end(), fo.begin(), fo.end());
    all.insert(all.end(), po.begin(), po.end());
    all.insert(all.end(), pr.begin(), pr.end());
    if (hipMalloc(&b->ops_d, b->ops_h.size() * sizeof(BatchOp)) != hipSuccess ||
        hipMalloc(&b->items_d, std::max<size_t>(all.size(), 1) * sizeof(uint4)) != hipSuccess) {
      rc = fail(GNUFFT_ERR_OUT_OF_MEMORY, "batch_create: descriptor allocation");
    } else if (hipMemcpyAsync(b->items_d, all.data(), all.size() * sizeof(uint4),
                              hipMemcpyHostToDevice, s) != hipSuccess) {
      rc = fail(GNUFFT_ERR_DEVICE, "batch_create: item upload");
    }
  }
  if (!rc) {
    size_t at = 0;
    auto list = [&](BatchList& l, const std::vector<uint4>& v, const std::vector<uint32_t>* cost) {
      l.items = b->items_d + at;
      if (cost) {
        partition_items(cost->data(), (uint32_t)v.size(), &l);
      } else {
        l.n = (uint32_t)v.size();
      }
      at += v.size();
    };
    list(b->spread, sp, &sp_cost);
    b->interp = b->spread;
    list(b->spread_small, ss, &ss_cost);
    b->interp_small = b->spread_small;
    list(b->reduce, rd, nullptr);
    list(b->fold, fo, &fo_cost);
    list(b->post, po, nullptr);
    list(b->pre, pr, nullptr);
    if (b->fused) {
      // items (plan, block) of every pass of both directions, one array
      const size_t db = fused_fft_desc_bytes();
      const int np = fused_fft_pass_count(b->dim);
      std::vector<uint2> fi;
      std::vector<size_t> first;
      for (int dir = 0; dir < 2; dir++)
        for (int ps = 0; ps < np; ps++)
          for (int big = 0; big < 2; big++) {
            gnufft_batch_s::FftPass& fp = dir == 0 ? (big ? b->fft_adj_big[ps] : b->fft_adj[ps])
                                                   : (big ? b->fft_fwd_big[ps] : b->fft_fwd[ps]);
            first.push_back(fi.size());
            for (uint32_t k = 0; k < n_plans; k++) {
              if (!b->op_fused[k]) continue;
              const void* d = b->fft_desc_h.data() + k * db;
              if (fused_fft_pass_big(d, dir == 0, ps) != big) continue;
              const uint32_t nb = fused_fft_pass_blocks(d, dir == 0, ps);
              for (uint32_t blk = 0; blk < nb; blk++) fi.push_back(make_uint2(k, blk));
              fp.lds = std::max(fp.lds, fused_fft_pass_lds(d, dir == 0, ps));
            }
            fp.n = (uint32_t)(fi.size() - first.back());
          }
      if (hipMalloc(&b->fft_desc_d, std::max<size_t>(b->fft_desc_h.size(), 1)) != hipSuccess ||
          hipMalloc(&b->fft_items_d, std::max<size_t>(fi.size(), 1) * sizeof(uint2)) != hipSuccess) {
        rc = fail(GNUFFT_ERR_OUT_OF_MEMORY, "batch_create: FFT descriptor allocation");
      } else if (hipMemcpyAsync(b->fft_items_d, fi.data(), fi.size() * sizeof(uint2),
                                hipMemcpyHostToDevice, s) != hipSuccess) {
        rc = fail(GNUFFT_ERR_DEVICE, "batch_create: FFT item upload");
      } else {
        size_t j = 0;
        for (int dir = 0; dir < 2; dir++)
          for (int ps = 0; ps < np; ps++)
            for (int big = 0; big < 2; big++)
              (dir == 0 ? (big ? b->fft_adj_big[ps] : b->fft_adj[ps])
                        : (big ? b->fft_fwd_big[ps] : b->fft_fwd[ps]))
                  .items = b->fft_items_d + first[j++];
      }
    }
    for (auto& gr : b->groups) {
      uint32_t coils = 0;
      for (uint32_t k = gr.first; k < gr.first + gr.count; k++) coils += b->coils[k];
      rc = create_fft(b->plans[gr.first]->g, coils, &gr.fft);
      if (rc) break;
    }
  }
  if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = fail(GNUFFT_ERR_DEVICE, "batch_create: sync");
  if (rc) {
    free_batch(b);
    return rc;
  }
  *batch = b;
  return GNUFFT_OK;
}

int gnufft_batch_destroy(gnufft_batch batch) {
  if (!batch) return GNUFFT_OK;
  (void)hipDeviceSynchronize();
  free_batch(batch);
  return GNUFFT_OK;
}

int gnufft_batch_adjoint(gnufft_batch b, const void* const* kdata, void* const* out, void* stream) {
  if (!b || !kdata || !out) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "batch_adjoint: null argument");
  hipStream_t s = (hipStream_t)stream;
  int rc0 = batch_check_plans(b);
  if (!rc0) rc0 = batch_wait_upload(b, s);
  if (rc0) return rc0;
  for (size_t k = 0; k < b->ops_h.size(); k++) {
    const uint32_t ci = b->caller_index[k];
    if (!kdata[ci] || !out[ci]) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "batch_adjoint: null buffer");
    b->ops_h[k].s.data = (const float2*)kdata[ci];
    b->ops_h[k].post.out = (float2*)out[ci];
    if (b->op_fused[k])
      fused_fft_desc_set_io(b->fft_desc_h.data() + k * fused_fft_desc_bytes(), (float2*)out[ci],
                            nullptr);
  }
  int rc = batch_upload_ops(b, s);
  if (rc) return rc;
  HIP_TRY(launch_spread_batch(b->dim, b->nx, b->ops_d, b->spread, b->spread_lds, s));
  HIP_TRY(launch_spread_batch_small(b->dim, b->ops_d, b->spread_small, b->spread_small_lds, s));
  HIP_TRY(launch_reduce_batch(b->ops_d, b->reduce, b->max_kstride, s));
  HIP_TRY(launch_fold_batch(b->dim, b->ops_d, b->fold, b->fold_threads, b->fold_generic, s));
  if (b->fused) {
    for (int ps = 0; ps < fused_fft_pass_count(b->dim); ps++) {
      HIP_TRY(launch_fused_fft_batch(b->dim, true, ps, false, b->fft_prime, b->fft_desc_d,
                                     b->fft_adj[ps].items, b->fft_adj[ps].n, b->fft_adj[ps].lds, s));
      HIP_TRY(launch_fused_fft_batch(b->dim, true, ps, true, b->fft_prime, b->fft_desc_d,
                                     b->fft_adj_big[ps].items, b->fft_adj_big[ps].n,
                                     b->fft_adj_big[ps].lds, s));
    }
  }
  for (auto& gr : b->groups) {
    rc = exec_fft(gr.fft, true, b->grid_d + gr.offset, s);
    if (rc) return rc;
  }
  HIP_TRY(launch_post_batch(b->ops_d, b->post, s));
  return batch_mark_done(b, s);
}

int gnufft_batch_forward(gnufft_batch b, const void* const* img, void* const* out, void* stream) {
  if (!b || !img || !out) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "batch_forward: null argument");
  hipStream_t s = (hipStream_t)stream;
  int rc0 = batch_check_plans(b);
  if (!rc0) rc0 = batch_wait_upload(b, s);
  if (rc0) return rc0;
  for (size_t k = 0; k < b->ops_h.size(); k++) {
    const uint32_t ci = b->caller_index[k];
    if (!img[ci] || !out[ci]) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "batch_forward: null buffer");
    b->ops_h[k].pre.img = (const float2*)img[ci];
    b->ops_h[k].i.out = (float2*)out[ci];
    if (b->op_fused[k])
      fused_fft_desc_set_io(b->fft_desc_h.data() + k * fused_fft_desc_bytes(), nullptr,
                            (const float2*)img[ci]);
  }
  int rc = batch_upload_ops(b, s);
  if (rc) return rc;
  HIP_TRY(launch_pre_batch(b->ops_d, b->pre, s));
  for (auto& gr : b->groups) {
    rc = exec_fft(gr.fft, false, b->grid_d + gr.offset, s);
    if (rc) return rc;
  }
  if (b->fused) {
    for (int ps = 0; ps < fused_fft_pass_count(b->dim); ps++) {
      HIP_TRY(launch_fused_fft_batch(b->dim, false, ps, false, b->fft_prime, b->fft_desc_d,
                                     b->fft_fwd[ps].items, b->fft_fwd[ps].n, b->fft_fwd[ps].lds, s));
      HIP_TRY(launch_fused_fft_batch(b->dim, false, ps, true, b->fft_prime, b->fft_desc_d,
                                     b->fft_fwd_big[ps].items, b->fft_fwd_big[ps].n,
                                     b->fft_fwd_big[ps].lds, s));
    }
  }
  HIP_TRY(launch_interp_batch(b->dim, b->nx, b->ops_d, b->interp, b->interp_lds, s));
  HIP_TRY(launch_interp_batch_small(b->dim, b->ops_d, b->interp_small, b->interp_small_lds, s));
  return batch_mark_done(b, s);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Adjoint without allocation (no counterpart in the reference): the work
// grid and the spread sub-grids live in a fixed per-device area, so the call
// makes no hipMalloc / hipFree and no allocating library call. The area is
// a file-scope device array; host code uses only the address that
// hipGetSymbolAddress returns for the current device (the array's host-side
// name is no device address).
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kNoAllocBytes = (size_t)GNUFFT_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_noalloc_mu;
std::map<int, unsigned char*> g_noalloc_area;  // device -> address of its `scratch`

// Address of this device's area, looked up once per device (with that
// device current, as the caller's calls are).
int noalloc_area(unsigned char** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  auto it = g_noalloc_area.find(dev);
  if (it != g_noalloc_area.end()) {
    *out = it->second;
    return GNUFFT_OK;
  }
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(GNUFFT_ERR_DEVICE, std::string("adjoint_noalloc: work area address: ") +
                                       hipGetErrorString(e));
  g_noalloc_area[dev] = static_cast<unsigned char*>(ptr);
  *out = static_cast<unsigned char*>(ptr);
  return GNUFFT_OK;
}

size_t round_up_256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

size_t gnufft_noalloc_capacity(void) { return kNoAllocBytes; }

int gnufft_adjoint_noalloc(gnufft_plan p, const void* kdata, void* out, uint32_t n_coils, int stop,
                           void* stream) {
  if (!p || !out || (!kdata && p->M))
    return fail(GNUFFT_ERR_INVALID_ARGUMENT, "adjoint_noalloc: null argument");
  if (!valid_stop(stop)) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "adjoint_noalloc: unknown stop stage");
  if (!p->sdc_d) return fail(GNUFFT_ERR_STATE, "adjoint_noalloc: no trajectory set");
  if (n_coils == 0) return fail(GNUFFT_ERR_INVALID_ARGUMENT, "adjoint_noalloc: zero coils");
  if (p->n_sens > 0 && p->n_sens != n_coils)
    return fail(GNUFFT_ERR_INVALID_ARGUMENT, "adjoint_noalloc: coil count differs from sensitivity coils");
  const bool grid_out = stop == GNUFFT_STOP_CONVOLUTION || stop == GNUFFT_STOP_DENSITY_ESTIMATION;
  if (!grid_out && !p->fused)
    return fail(GNUFFT_ERR_UNSUPPORTED,
                "adjoint_noalloc: needs the fused FFT (rocFFT plans allocate work memory)");
  const Geom& g = p->g;
  // per coil: work grid (unless the caller's buffer is the grid) and the
  // spread sub-grids / fused FFT intermediate
  const size_t grid_b = grid_out ? 0 : (size_t)g.Ng * sizeof(float2);
  const size_t scr_b = scratch_per_coil(p);
  const size_t per = round_up_256(grid_b) + round_up_256(scr_b);
  if (per == 0 || per > kNoAllocBytes)
    return fail(GNUFFT_ERR_UNSUPPORTED,
                "adjoint_noalloc: one coil needs " + std::to_string(per) +
                    " bytes of work memory, more than the no-allocation area (" +
                    std::to_string(kNoAllocBytes) + " bytes)");
  unsigned char* area = nullptr;
  int rc = noalloc_area(&area);
  if (rc) return rc;
  const uint32_t cc = (uint32_t)std::min<size_t>(n_coils, kNoAllocBytes / per);
  const bool use_sens = p->n_sens > 0 && stop == GNUFFT_STOP_DEAPODIZATION;
  hipStream_t s = (hipStream_t)stream;
  for (uint32_t c0 = 0; c0 < n_coils; c0 += cc) {
    const uint32_t nc = std::min(cc, n_coils - c0);
    float2* grid = grid_out ? (float2*)out + (size_t)c0 * g.Ng : (float2*)area;
    float2* scr = (float2*)(area + round_up_256(grid_b * nc));
    SpreadArgs a = spread_args(p, (const float2*)kdata + (size_t)c0 * p->M, grid, nc);
    a.scratch = scr;
    a.gather = p->idx_d;
    a.dens = p->dens_sorted_d;
    a.fold_shift = !grid_out;
    HIP_TRY(launch_spread(g, a, (int)p->round_target, s));
    if (grid_out) continue;
    float2* dst = use_sens ? (float2*)out : (float2*)out + (size_t)c0 * g.Ni;
    HIP_TRY(fused_fft_adjoint(g, grid, scr, (int)nc, p->deapo_d,
                              use_sens ? p->sens_d + (size_t)c0 * g.Ni : nullptr, dst, stop,
                              use_sens && c0 > 0, p->fft_tw_d, p->fft_tw_off, s));
  }
  return GNUFFT_OK;
}

}  // extern "C"
