// Internal declarations shared by the HIP translation units of
// libgpunufft_hip.so. Not part of the public ABI (include/gpunufft_hip.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace gnufft {

// Geometry and kernel constants of one operator, passed to kernels by value.
// Field meanings follow GpuNUFFTInfo (CUDA/inc/gpuNUFFT_types.hpp:230-311,
// derived in CUDA/src/gpuNUFFT_operator.cpp:53-146).
struct Geom {
  int dim;          // 2 or 3
  int G[3];         // grid dims; G[2] = 1 in 2-D
  int N[3];         // image dims; N[2] = 1 in 2-D
  int S[3];         // sectors per dim; S[2] = 1 in 2-D
  int sw;           // sector width
  int P;            // padded sector width sw + 2*floor(kw/2)
  int Pmax;         // P - 1
  int off;          // floor(P/2)
  int kw;           // kernel width
  int K;            // lookup table entries
  int lookup;       // GNUFFT_LOOKUP_*
  float kradius;    // kw/2
  float rsq;        // squared kernel radius relative to the largest grid dim
  float rsq_inv;
  float dist_mult;  // (K-1)/rsq
  float aniso[3];   // G_a / max(G)
  int shift_inv[3]; // floor(G/2): INVERSE fftshift offset
  int shift_fwd[3]; // ceil(G/2): FORWARD fftshift offset
  int crop[3];      // (int)(N*(osf-1)/2)
  long long Ng;     // grid points per coil
  long long Ni;     // image points per coil
  int M;            // samples
  float fft_scale;  // 1/sqrt(Ng)
  // multiply-high reciprocals: q = umulhi(n, mP) == n / P for n < 2^20
  uint32_t mP, mP2, mPD;  // for P, P*P and the padded sector point count
};

// ceil(2^32 / d): exact quotient via __umulhi(n, m) for n*d < 2^32.
__host__ inline uint32_t udiv_magic(uint32_t d) {
  return (uint32_t)((0x100000000ull + d - 1) / d);
}

__device__ inline uint32_t udiv(uint32_t n, uint32_t m) { return __umulhi(n, m); }

// One gridding work item: samples [begin, end) of a sector.
struct Chunk {
  uint32_t sector;
  uint32_t begin;
  uint32_t end;
  uint32_t pad;
};

// ---------------------------------------------------------------------------
// Exact per-sample index / weight arithmetic. Contraction is disabled so the
// float rounding of every ceil/floor/round input matches the reference's
// separate multiply and add (CUDA/inc/gpuNUFFT_utils.hpp:89-104, 194-209).
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)

__device__ __host__ inline float map_kspace_to_grid(float pos, int G, int center, int off) {
  return (pos * (float)G) + (0.5f * (float)G) - (float)center + (float)off;
}

__device__ __host__ inline float map_grid_to_kspace(int i, int G, int center, int off) {
  return ((float)i + (float)center - (float)off) / (float)G - 0.5f;
}

__device__ __host__ inline void tap_range(float u, int Pmax, float radius, int* lo, int* hi) {
  int mn = (int)ceilf(u - radius);
  int mx = (int)floorf(u + radius);
  if (mn < 0) mn = 0;
  if (mx < 0) mx = 0;
  if (mx >= Pmax) mx = Pmax;
  if (mn >= (float)Pmax) mn = (int)((float)Pmax - 2.0f * radius);
  if (mn < 0) mn = 0;  // only reachable when P < kw + 1 (reference reads out of bounds)
  *lo = mn;
  *hi = mx;
}

// Emulated clamp-addressed linear texture fetch of the 1-D table at the
// normalised coordinate u: texel coordinate u*K-0.5, fractional weight with
// 8 fractional bits (round to nearest).
__device__ __host__ inline float lerp_fetch(const float* tab, int K, float u) {
  float xb = u * (float)K - 0.5f;
  float fl = floorf(xb);
  int i0 = (int)fl;
  float a = xb - fl;
  a = rintf(a * 256.0f) * (1.0f / 256.0f);
  int i1 = i0 + 1;
  i0 = i0 < 0 ? 0 : (i0 > K - 1 ? K - 1 : i0);
  i1 = i1 < 0 ? 0 : (i1 > K - 1 ? K - 1 : i1);
  return (1.0f - a) * tab[i0] + a * tab[i1];
}

// Weight of one tap along one axis; returns false when the nearest-mode
// radius test rejects the tap (atomic_gpuNUFFT_kernels.cu:343-368); the
// linear mode has no radius test (texture_gpuNUFFT_kernels.cu:63-82).
__device__ __host__ inline bool axis_weight(const Geom& g, const float* lut, int a, int i,
                                            int center, float pos, float* w) {
  float kp = map_grid_to_kspace(i, g.G[a], center, g.off);
  float d = (kp - pos) * g.aniso[a];
  d *= d;
  if (g.lookup == 0) {
    if (!(d < g.rsq)) return false;
    int li = (int)roundf(d * g.dist_mult);
    *w = lut[li < g.K - 1 ? li : g.K - 1];
    return true;
  }
  *w = lerp_fetch(lut, g.K, d * g.rsq_inv);
  return true;
}

// Same weight with the tap's k-space position precomputed
// (kp = map_grid_to_kspace(i, G, center, off), identical float value).
__device__ __host__ inline bool axis_weight_kp(const Geom& g, const float* lut, int a, float kp,
                                               float pos, float* w) {
  float d = (kp - pos) * g.aniso[a];
  d *= d;
  if (g.lookup == 0) {
    if (!(d < g.rsq)) return false;
    int li = (int)roundf(d * g.dist_mult);
    *w = lut[li < g.K - 1 ? li : g.K - 1];
    return true;
  }
  *w = lerp_fetch(lut, g.K, d * g.rsq_inv);
  return true;
}

// Sector index along one axis (CUDA/inc/precomp_utils.hpp:34-48): float
// product, double sum, round half away from zero in double, double
// quotient, floor, clamped to ceil((float)G/sw)-1 evaluated in float.
__device__ inline int sector_of(float coord, int G, int sw) {
  double v = (double)(coord * (float)G) + 0.5 * (double)(float)G;
  int sector = (int)floor(round(v) / (double)(float)sw);
  int smax = (int)ceilf((float)G / (float)sw);
  if (sector >= smax) sector = smax - 1;
  if (sector < 0) sector = 0;
  return sector;
}

#pragma clang fp contract(on)

// Periodic wrap into [0, G). Callers pass coordinates within a few grid
// lengths of the grid, so a short loop of +-G replaces a division.
__device__ __host__ inline int wrap(int p, int G) {
  while (p >= G) p -= G;
  while (p < 0) p += G;
  return p;
}

// Message returned by gnufft_last_error() on this thread (capi.hip).
void set_last_error(const std::string& msg);

}  // namespace gnufft

struct gnufft_plan_s;

namespace gnufft {

// Generation of a plan's device buffers (capi.hip): changes whenever memory
// that a captured graph or a batch may reference is freed or replaced.
uint64_t plan_generation(const gnufft_plan_s* p);

// ---------------------------------------------------------------------------
// Launchers implemented in the .hip translation units (all asynchronous on
// `stream`; return hipError_t).
// ---------------------------------------------------------------------------

// precompute.hip
hipError_t launch_assign_sectors(const Geom& g, const float* traj, uint32_t M, uint32_t* keys,
                                 uint32_t* vals, hipStream_t s);
hipError_t sort_pairs(void* temp, size_t* temp_bytes, const uint32_t* keys_in,
                      uint32_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out,
                      uint32_t M, int key_bits, hipStream_t s);
hipError_t launch_sector_data_count(const uint32_t* sorted_keys, uint32_t M, uint32_t S,
                                    uint32_t* sdc, hipStream_t s);
hipError_t launch_sector_centers(const Geom& g, uint32_t* centers, hipStream_t s);
hipError_t launch_gather_traj(const float* traj, const float* dens, const uint32_t* idx,
                              uint32_t M, int dim, float* traj_sorted, float* dens_sorted,
                              hipStream_t s);
hipError_t count_chunks(const uint32_t* sdc, uint32_t S, uint32_t chunk, uint32_t* cnt,
                        uint32_t* chunk_off, void* temp, size_t* temp_bytes, hipStream_t s);
hipError_t launch_fill_chunks(const uint32_t* sdc, const uint32_t* chunk_off, uint32_t S,
                              uint32_t chunk, Chunk* chunks, hipStream_t s);

// gridding.hip

// Gridding work list: chunks in sector order, split into 8 contiguous
// ranges of about equal cost; range x is processed on XCD x (workgroups
// are dispatched round-robin over the XCDs), so neighbouring sectors share
// an L2 and no XCD is left with the dense k-space centre alone.
struct WorkList {
  const Chunk* chunks;
  uint32_t n_chunks;
  uint32_t xcd_begin[9];  // range x = [xcd_begin[x], xcd_begin[x+1])
  uint32_t max_range;     // longest range: workgroups per XCD
};

// Split `n` chunks of the given sample counts into the 8 ranges of `w`.
void partition_work(const uint32_t* counts, uint32_t n, WorkList* w);

// Padded-sector coverage, per axis a: the home interval
// [s*sw, min(s*sw + sw, G)) of sector index s along a meets (with periodic
// wrap) the padded regions of the sector indices
// ent[a][off[a][s] .. off[a][s+1]).
struct Cover {
  const uint32_t* off[3];
  const uint32_t* ent[3];
  int max_len[3];  // longest list per axis
  // most (candidate, periodic image) pairs covering one grid coordinate of
  // one axis; the fold indexes them per coordinate when at most kCovMax
  int max_cov;
};

// Host: build the per-axis lists of `g` as one array (offsets of axis a at
// word base[a], entries at word base[3 + a]); max_len[a] = longest list,
// max_cov as in Cover.
void build_cover(const Geom& g, std::vector<uint32_t>* words, size_t base[6], int max_len[3],
                 int* max_cov);

struct SpreadArgs {
  const float2* data;      // [coil][...] samples
  long long coil_stride;   // elements between coils of `data`
  const uint32_t* gather;  // NULL: data already in sector order
  const float* dens;       // NULL: no density weighting (sorted order)
  const float* crds;       // sorted SoA [dim][M]
  WorkList work;
  const float* lut;
  float2* scratch;         // sub-grids [chunk][coil block][padded point][block width]
  int sub_cb;              // coil block of that layout (the spread's; set by its launcher)
  const uint32_t* sector_chunks;  // [S+1] first chunk of each sector
  const uint2* reduce_tasks;      // (first chunk, count) of each multi-chunk sector
  uint32_t n_reduce;
  Cover cover;
  float2* grid;            // [coil][Ng], every point written (no prior zeroing)
  int n_coils;
  bool fold_shift;         // store at INVERSE-fftshifted positions
  int lane_per_cu;         // lane kernels: workgroups per CU their coil block is sized for (0: default)
};
hipError_t launch_spread(const Geom& g, const SpreadArgs& a, int chunk, hipStream_t s);

struct InterpArgs {
  const float2* grid;      // [coil][Ng]
  const float* crds;
  WorkList work;
  const float* lut;
  float2* out;             // [coil][...]
  long long coil_stride;
  const uint32_t* scatter; // NULL: write in sector order
  const float* dens;       // NULL: no density weighting
  float scale;             // 1 (kernel only) or 1/sqrt(Ng)
  int n_coils;
  bool fold_shift;         // read grid at FORWARD-fftshifted positions
  int lane_per_cu;         // as in SpreadArgs
  // runs of up to kInterpRun consecutive chunks of one sector, for the
  // round-based interp (it stages a sector once per run; n_chunks 0: none)
  WorkList work_runs;
};
// Chunks per interp run: 4 stages the dense k-space centre's sectors fewer
// times at the 128^3 koosh workload (interp 0.62 -> 0.60 ms); 8 and whole
// sectors leave a long tail of busy workgroups (0.67 and 1.62 ms).
constexpr int kInterpRun = 4;
hipError_t launch_interp(const Geom& g, const InterpArgs& a, int chunk, hipStream_t s);


// prepost.hip
hipError_t launch_zero(void* p, size_t bytes, hipStream_t s);
// out[z][y][x] = norm / |sum_t B[t][y][x] E[t][z]| (double), t < T.
hipError_t launch_deapo_combine(const double2* B, const double2* E, int T, int Nx, int Ny, int Nz,
                                double norm, float* out, hipStream_t s);
hipError_t launch_adjoint_post(const Geom& g, const float2* grid, int n_coils,
                               const float* deapo, const float2* sens, float2* out,
                               bool accumulate, int stop, hipStream_t s);
hipError_t launch_forward_pre(const Geom& g, const float2* img, bool img_per_coil,
                              const float2* sens, const float* deapo, float2* grid, int n_coils,
                              hipStream_t s);
hipError_t launch_image_prep(const Geom& g, const float2* img, bool img_per_coil,
                             const float2* sens, const float* deapo, float2* out, int n_coils,
                             hipStream_t s);

// ---------------------------------------------------------------------------
// Batches of independent operators: every stage of every operator runs in
// one launch per stage. Workgroups take items (operator, index, coil block)
// from a flat list split into 8 XCD ranges of equal cost, and read their
// operator's geometry and arguments from a descriptor in device memory.
// ---------------------------------------------------------------------------
struct PostArgs {  // adjoint tail (crop, scale, deapo, coil combine)
  const float2* grid;
  int n_coils;
  const float* deapo;
  const float2* sens;  // NULL: per-coil output
  float2* out;
  int stop;
};

struct PreArgs {  // forward head (sens, deapo, pad, shift)
  const float2* img;
  int img_per_coil;
  const float2* sens;
  const float* deapo;
  float2* grid;
  int n_coils;
};

struct BatchOp {
  Geom g;
  SpreadArgs s;
  InterpArgs i;
  PostArgs post;
  PreArgs pre;
  int R_spread;
  int R_interp;
};

struct BatchList {
  const uint4* items;  // (operator, index, coil block, unused)
  uint32_t n;
  uint32_t xcd_begin[9];
  uint32_t max_range;
};

// Split items of the given costs into the 8 XCD ranges of `l` (items stay in order).
void partition_items(const uint32_t* costs, uint32_t n, BatchList* l);

// gridding.hip (CB: coils per workgroup, the same for every operator).
// Batched spread / interp workgroups have kBatchThreads threads and their
// rounds are sized so that kBatchPerCU of them share a CU. (Smaller
// workgroups do not raise that count: the 8-coil sub-grid dominates the LDS
// footprint, 20.7 KB at 2-D sw 16.)
constexpr int kBatchCB = 8;
constexpr int kBatchLaneCB = 4;  // coil block of the batched lane (kernel width 3) kernels
constexpr int kBatchThreads = 512;
// Batched lane kernels: chunks of at most 2 * kBatchSmallThreads samples run
// in workgroups of kBatchSmallThreads (sparse 2-D plans are mostly such
// chunks, and a 512-thread workgroup would leave most of its waves idle).
constexpr int kBatchSmallThreads = 128;
constexpr int kBatchPerCU = 4;
hipError_t launch_spread_batch(int dim, int nx, const BatchOp* ops, const BatchList& l,
                               size_t lds, hipStream_t s);
// generic: some plan's cover exceeds the fold's per-coordinate tables
// (fold_generic), so the kernel variant with the generic body runs
hipError_t launch_fold_batch(int dim, const BatchOp* ops, const BatchList& l, int threads,
                             bool generic, hipStream_t s);
bool fold_generic(const Cover& cover);
hipError_t launch_interp_batch(int dim, int nx, const BatchOp* ops, const BatchList& l,
                               size_t lds, hipStream_t s);
// lane kernels in kBatchSmallThreads workgroups, for chunks of at most
// 2 * kBatchSmallThreads samples
hipError_t launch_spread_batch_small(int dim, const BatchOp* ops, const BatchList& l, size_t lds,
                                     hipStream_t s);
hipError_t launch_interp_batch_small(int dim, const BatchOp* ops, const BatchList& l, size_t lds,
                                     hipStream_t s);
size_t spread_lds_bytes(const Geom& g, int cb, int R);
// Batched lane kernels (launch_*_batch with nx == 3): usable for plan g;
// their dynamic LDS for a coil block of cb.
bool lane_batch_ok(const Geom& g);
hipError_t launch_reduce_batch(const BatchOp* ops, const BatchList& l, uint32_t max_kstride,
                               hipStream_t s);
// lut: with the kernel table copied to LDS (the small-workgroup spread reads it from global memory)
size_t lane_lds_bytes(const Geom& g, int cb, bool lut = true);
size_t interp_lds_bytes(const Geom& g, int cb, int R);
int spread_round(const Geom& g, int cb, int chunk, size_t limit);
int interp_round(const Geom& g, int cb, int chunk, size_t limit);
size_t lds_budget();
int fold_threads(const Geom& g, const Cover& cover);

// prepost.hip: items (operator, block of 256 pixels / grid points, 0, 0)
hipError_t launch_post_batch(const BatchOp* ops, const BatchList& l, hipStream_t s);
hipError_t launch_pre_batch(const BatchOp* ops, const BatchList& l, hipStream_t s);

// Fused oversampled-grid FFT (fft.hip).
bool fused_fft_supported(const Geom& g);
bool fused_fft_preferred(const Geom& g);
// Some axis of g needs a prime stage (7 .. 19).
bool fused_fft_prime(const Geom& g);
size_t fused_fft_tmp_elems(const Geom& g);
std::vector<float2> fused_fft_twiddles(const Geom& g, size_t offsets[3][3]);
hipError_t fused_fft_adjoint(const Geom& g, float2* grid, float2* tmp, int coils,
                             const float* deapo, const float2* sens, float2* out, int stop,
                             bool accumulate, const float2* tw, const size_t tw_off[3][3],
                             hipStream_t s);
hipError_t fused_fft_forward(const Geom& g, const float2* img, bool img_per_coil,
                             const float2* sens, const float* deapo, float2* grid, float2* tmp,
                             int coils, const float2* tw, const size_t tw_off[3][3],
                             hipStream_t s);
// Batches: one opaque descriptor (fused_fft_desc_bytes) per operator in
// device memory; each pass of every operator is one launch over items
// (operator, block). Passes: adjoint ax[0 .. d-2] then x, forward x then
// ax[0 .. d-2] (fused_fft_pass_count(d) of each).
size_t fused_fft_desc_bytes();
bool fused_fft_desc_build(void* desc, const Geom& g, float2* grid, float2* tmp, int coils,
                          const float* deapo, const float2* sens, const float2* tw,
                          const size_t tw_off[3][3]);
void fused_fft_desc_set_io(void* desc, float2* adjoint_out, const float2* forward_img);
int fused_fft_pass_count(int dim);
uint32_t fused_fft_pass_blocks(const void* desc, bool adjoint, int pass);
size_t fused_fft_pass_lds(const void* desc, bool adjoint, int pass);
int fused_fft_pass_big(const void* desc, bool adjoint, int pass);
// prime: some operator's lengths need the prime stages 7 .. 19
hipError_t launch_fused_fft_batch(int dim, bool adjoint, int pass, bool big, bool prime,
                                  const void* descs_d, const uint2* items, uint32_t n_items,
                                  size_t lds, hipStream_t s);

}  // namespace gnufft
