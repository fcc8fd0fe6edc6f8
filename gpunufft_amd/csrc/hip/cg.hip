// Device-resident CG-SENSE solver on one plan (gnufft_cg_*).
//
// Solves (A^H A + alpha I) x = A^H y with the plain conjugate-gradient loop
// of the reference's demo reconstruction (matlab/demo/utils/cg_sense_2d.m:
// right-hand side :46-47 with useMulticoil, system matrix :51 and :99-102,
// iterations :63-82), where A is the plan's forward operator including
// coil sensitivities and density compensation. Vectors and scalars live in
// device memory; the scalar reductions are two-stage in double precision in
// a fixed order (deterministic); no iteration needs the host. One iteration
// (forward, adjoint, CG updates) is captured into a HIP graph after a warm-up
// and replayed.

#include <hip/hip_runtime.h>

#include <new>
#include <string>

#include "../../../include/gpunufft_hip.h"
#include "gnufft_internal.hpp"

namespace {

constexpr int kCgThreads = 256;
constexpr int kCgBlocks = 512;  // partial sums per reduction

int cg_fail(int code, const std::string& msg) {
  gnufft::set_last_error(msg);
  return code;
}

#define CG_HIP(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return cg_fail(GNUFFT_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CG_ABI(expr)                  \
  do {                                \
    int rc_ = (expr);                 \
    if (rc_ != GNUFFT_OK) return rc_; \
  } while (0)

__device__ inline double block_sum(double v) {
  __shared__ double red[kCgThreads / 64];
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kCgThreads / 64; w++) t += red[w];
  return t;  // valid in thread 0
}

// Ap += alpha p; part[block] = sum Re(conj(p) Ap)
__global__ void __launch_bounds__(kCgThreads) cg_pap_kernel(const float2* __restrict__ p,
                                                            float2* __restrict__ Ap, float alpha,
                                                            size_t n, double* __restrict__ part) {
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (size_t)gridDim.x * kCgThreads) {
    const float2 a = p[i];
    float2 b = Ap[i];
    b.x += alpha * a.x;
    b.y += alpha * a.y;
    Ap[i] = b;
    acc += (double)a.x * b.x + (double)a.y * b.y;
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// s[idx] = sum of part[0 .. nb) in index order (one block)
__global__ void __launch_bounds__(kCgThreads) cg_reduce_kernel(const double* __restrict__ part,
                                                               int nb, double* __restrict__ s,
                                                               int idx) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += kCgThreads) acc += part[i];
  acc = block_sum(acc);
  if (threadIdx.x == 0) s[idx] = acc;
}

// a = rr / pAp; x += a p; r -= a Ap; part[block] = sum |r|^2
__global__ void __launch_bounds__(kCgThreads) cg_step_kernel(float2* __restrict__ x,
                                                             float2* __restrict__ r,
                                                             const float2* __restrict__ p,
                                                             const float2* __restrict__ Ap,
                                                             const double* __restrict__ s, size_t n,
                                                             double* __restrict__ part) {
  const float a = s[1] != 0.0 ? (float)(s[0] / s[1]) : 0.0f;
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (size_t)gridDim.x * kCgThreads) {
    const float2 pi = p[i], api = Ap[i];
    float2 xi = x[i], ri = r[i];
    xi.x += a * pi.x;
    xi.y += a * pi.y;
    ri.x -= a * api.x;
    ri.y -= a * api.y;
    x[i] = xi;
    r[i] = ri;
    acc += (double)ri.x * ri.x + (double)ri.y * ri.y;
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// p = r + (rr_new / rr) p
__global__ void __launch_bounds__(kCgThreads) cg_dir_kernel(float2* __restrict__ p,
                                                            const float2* __restrict__ r,
                                                            const double* __restrict__ s,
                                                            size_t n) {
  const float b = s[0] != 0.0 ? (float)(s[2] / s[0]) : 0.0f;
  for (size_t i = (size_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (size_t)gridDim.x * kCgThreads) {
    const float2 ri = r[i];
    float2 pi = p[i];
    pi.x = ri.x + b * pi.x;
    pi.y = ri.y + b * pi.y;
    p[i] = pi;
  }
}

// rr = rr_new
__global__ void cg_shift_kernel(double* s) {
  if (threadIdx.x == 0) s[0] = s[2];
}

// x = 0; r = p = y; part[block] = sum |y|^2
__global__ void __launch_bounds__(kCgThreads) cg_init_kernel(float2* __restrict__ x,
                                                             float2* __restrict__ r,
                                                             float2* __restrict__ p,
                                                             const float2* __restrict__ y, size_t n,
                                                             double* __restrict__ part) {
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (size_t)gridDim.x * kCgThreads) {
    const float2 v = y[i];
    x[i] = make_float2(0.0f, 0.0f);
    r[i] = v;
    p[i] = v;
    acc += (double)v.x * v.x + (double)v.y * v.y;
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

}  // namespace

struct gnufft_cg_s {
  gnufft_plan plan = nullptr;
  uint32_t coils = 0;
  uint32_t samples = 0;
  float alpha = 0.0f;
  size_t n = 0;       // image elements (coil-combined)
  float2* buf = nullptr;  // x, r, p, Ap, y (n each) then k ([coils][M])
  float2 *x = nullptr, *r = nullptr, *p = nullptr, *Ap = nullptr, *y = nullptr, *k = nullptr;
  double* s = nullptr;     // rr, pAp, rr_new, unused
  double* part = nullptr;  // kCgBlocks partial sums
  int blocks = kCgBlocks;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t generation = 0;  // the plan's buffer generation the graph was captured at
  bool warm = false;
  // the solver's own stream (default priority, non-blocking: the legacy
  // null stream cannot be captured), ordered after / before the caller's
  // stream with events
  hipStream_t own = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
};

namespace {

// One CG iteration (the body of the captured graph).
int cg_iteration(gnufft_cg c, hipStream_t s) {
  CG_ABI(gnufft_forward(c->plan, c->p, c->k, c->coils, GNUFFT_STOP_DEAPODIZATION, s));
  CG_ABI(gnufft_adjoint(c->plan, c->k, c->Ap, c->coils, GNUFFT_STOP_DEAPODIZATION, s));
  hipLaunchKernelGGL(cg_pap_kernel, dim3(c->blocks), dim3(kCgThreads), 0, s, c->p, c->Ap,
                     c->alpha, c->n, c->part);
  hipLaunchKernelGGL(cg_reduce_kernel, dim3(1), dim3(kCgThreads), 0, s, c->part, c->blocks, c->s, 1);
  hipLaunchKernelGGL(cg_step_kernel, dim3(c->blocks), dim3(kCgThreads), 0, s, c->x, c->r, c->p,
                     c->Ap, c->s, c->n, c->part);
  hipLaunchKernelGGL(cg_reduce_kernel, dim3(1), dim3(kCgThreads), 0, s, c->part, c->blocks, c->s, 2);
  hipLaunchKernelGGL(cg_dir_kernel, dim3(c->blocks), dim3(kCgThreads), 0, s, c->p, c->r, c->s, c->n);
  hipLaunchKernelGGL(cg_shift_kernel, dim3(1), dim3(64), 0, s, c->s);
  CG_HIP(hipGetLastError());
  return GNUFFT_OK;
}

void cg_free(gnufft_cg c) {
  if (c->exec) (void)hipGraphExecDestroy(c->exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->buf) (void)hipFree(c->buf);
  if (c->s) (void)hipFree(c->s);
  if (c->ev_in) (void)hipEventDestroy(c->ev_in);
  if (c->ev_out) (void)hipEventDestroy(c->ev_out);
  if (c->own) (void)hipStreamDestroy(c->own);
  delete c;
}

}  // namespace

extern "C" {

int gnufft_cg_create(gnufft_cg* cg, gnufft_plan plan, uint32_t n_coils, float alpha) {
  if (!cg || !plan || n_coils == 0) return cg_fail(GNUFFT_ERR_INVALID_ARGUMENT, "cg_create: bad argument");
  *cg = nullptr;
  gnufft_plan_info info;
  CG_ABI(gnufft_plan_get_info(plan, &info));
  if (info.n_samples == 0) return cg_fail(GNUFFT_ERR_STATE, "cg_create: plan has no trajectory");
  if (info.n_sens_coils != n_coils)
    return cg_fail(GNUFFT_ERR_INVALID_ARGUMENT, "cg_create: plan needs sensitivities of n_coils coils");
  gnufft_cg c = new (std::nothrow) gnufft_cg_s();
  if (!c) return cg_fail(GNUFFT_ERR_OUT_OF_MEMORY, "cg_create: host allocation");
  c->plan = plan;
  c->coils = n_coils;
  c->samples = info.n_samples;
  c->alpha = alpha;
  c->n = (size_t)info.img_count;
  const size_t total = 5 * c->n + (size_t)n_coils * info.n_samples;
  if (hipMalloc(&c->buf, total * sizeof(float2)) != hipSuccess ||
      hipMalloc(&c->s, (4 + kCgBlocks) * sizeof(double)) != hipSuccess) {
    cg_free(c);
    return cg_fail(GNUFFT_ERR_OUT_OF_MEMORY, "cg_create: device allocation");
  }
  if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_out, hipEventDisableTiming) != hipSuccess) {
    cg_free(c);
    return cg_fail(GNUFFT_ERR_DEVICE, "cg_create: stream / event creation");
  }
  c->x = c->buf;
  c->r = c->x + c->n;
  c->p = c->r + c->n;
  c->Ap = c->p + c->n;
  c->y = c->Ap + c->n;
  c->k = c->y + c->n;
  c->part = c->s + 4;
  *cg = c;
  return GNUFFT_OK;
}

int gnufft_cg_destroy(gnufft_cg cg) {
  if (!cg) return GNUFFT_OK;
  (void)hipDeviceSynchronize();
  cg_free(cg);
  return GNUFFT_OK;
}

int gnufft_cg_solve(gnufft_cg c, const void* kdata, void* x_out, uint32_t iterations, int use_graph,
                    void* stream) {
  if (!c || !kdata || !x_out) return cg_fail(GNUFFT_ERR_INVALID_ARGUMENT, "cg_solve: null argument");
  {
    // the solver's vectors are sized for the plan's sample and coil counts
    gnufft_plan_info info;
    CG_ABI(gnufft_plan_get_info(c->plan, &info));
    if (info.n_samples != c->samples || info.n_sens_coils != c->coils)
      return cg_fail(GNUFFT_ERR_STATE,
                     "cg_solve: the plan's trajectory length or sensitivity coils changed since "
                     "cg_create; create a new solver");
  }
  // the work runs on the solver's stream, after the caller's stream's
  // earlier work and before its later work
  const hipStream_t caller = (hipStream_t)stream;
  const hipStream_t s = c->own;
  CG_HIP(hipEventRecord(c->ev_in, caller));
  CG_HIP(hipStreamWaitEvent(s, c->ev_in, 0));
  // right-hand side A^H y, then x = 0, r = p = A^H y, rr = |r|^2
  CG_ABI(gnufft_adjoint(c->plan, kdata, c->y, c->coils, GNUFFT_STOP_DEAPODIZATION, s));
  hipLaunchKernelGGL(cg_init_kernel, dim3(c->blocks), dim3(kCgThreads), 0, s, c->x, c->r, c->p,
                     c->y, c->n, c->part);
  hipLaunchKernelGGL(cg_reduce_kernel, dim3(1), dim3(kCgThreads), 0, s, c->part, c->blocks, c->s, 0);
  CG_HIP(hipGetLastError());
  // The captured iteration holds the plan's device pointers (tables, work
  // grid, scratch, sensitivities). When any of them was freed or replaced
  // since the capture (set_trajectory, set_sens, a larger coil batch,
  // set_fft), the graph is dropped and captured again.
  if (use_graph && c->exec && c->generation != gnufft::plan_generation(c->plan)) {
    CG_HIP(hipStreamSynchronize(s));  // earlier replays may still run
    (void)hipGraphExecDestroy(c->exec);
    (void)hipGraphDestroy(c->graph);
    c->exec = nullptr;
    c->graph = nullptr;
    c->warm = false;
  }
  if (use_graph && !c->exec) {
    if (!c->warm) {
      // allocate the plan's work buffers and FFT plans outside the capture
      CG_ABI(gnufft_forward(c->plan, c->p, c->k, c->coils, GNUFFT_STOP_DEAPODIZATION, s));
      c->warm = true;
    }
    const uint64_t gen = gnufft::plan_generation(c->plan);
    CG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = cg_iteration(c, s);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    CG_HIP(e);
    if (gen != gnufft::plan_generation(c->plan)) {
      // a buffer was replaced while capturing: the graph may hold a freed pointer
      (void)hipGraphDestroy(g);
      return cg_fail(GNUFFT_ERR_STATE, "cg_solve: plan buffers changed during graph capture");
    }
    c->graph = g;
    CG_HIP(hipGraphInstantiate(&c->exec, c->graph, nullptr, nullptr, 0));
    c->generation = gen;
  }
  for (uint32_t it = 0; it < iterations; it++) {
    if (use_graph) {
      CG_HIP(hipGraphLaunch(c->exec, s));
    } else {
      CG_ABI(cg_iteration(c, s));
    }
  }
  CG_HIP(hipMemcpyAsync(x_out, c->x, c->n * sizeof(float2), hipMemcpyDeviceToDevice, s));
  CG_HIP(hipEventRecord(c->ev_out, s));
  CG_HIP(hipStreamWaitEvent(caller, c->ev_out, 0));
  return GNUFFT_OK;
}

}  // extern "C"
