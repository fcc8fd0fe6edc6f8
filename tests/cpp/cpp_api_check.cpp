// Self-check of the C++ gpuNUFFT:: API on a GPU (run by
// tests/test_cpp_api_gpu.py). Covers what the Python module does not reach:
// the GpuArray overloads, loadPrecomputedGpuNUFFTOperator from the tables of
// a factory-built operator, the operator flavour per factory flags, stop
// stages, the adjoint identity, and exceptions for invalid arguments.
// Prints "CPP_API_CHECK OK" and exits 0 when every check passes.

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpuNUFFT_operator_factory.hpp"
#include "gpunufft_hip.h"

using namespace gpuNUFFT;
typedef std::complex<float> cf;

static int failures = 0;

static void check(bool ok, const std::string& what) {
  std::printf("%s %s\n", ok ? "ok  " : "FAIL", what.c_str());
  std::fflush(stdout);
  if (!ok) failures++;
}

static double rel_l2(const CufftType* a, const CufftType* b, size_t n) {
  double num = 0, den = 0;
  for (size_t i = 0; i < n; i++) {
    const double dx = (double)a[i].x - b[i].x, dy = (double)a[i].y - b[i].y;
    num += dx * dx + dy * dy;
    den += (double)b[i].x * b[i].x + (double)b[i].y * b[i].y;
  }
  return std::sqrt(num / (den > 0 ? den : 1));
}

static std::complex<double> vdot(const CufftType* a, const CufftType* b, size_t n) {
  std::complex<double> s = 0;
  for (size_t i = 0; i < n; i++) s += std::conj(std::complex<double>(a[i].x, a[i].y)) *
                                      std::complex<double>(b[i].x, b[i].y);
  return s;
}

static void* to_device(const void* h, size_t bytes) {
  void* d = nullptr;
  checkStatus(gnufft_malloc(&d, bytes), "malloc");
  checkStatus(gnufft_memcpy(d, h, bytes, GNUFFT_DEVICE, GNUFFT_HOST, nullptr), "copy in");
  return d;
}

static void to_host(void* h, const void* d, size_t bytes) {
  checkStatus(gnufft_memcpy(h, d, bytes, GNUFFT_HOST, GNUFFT_DEVICE, nullptr), "copy out");
}

int main() {
  // 2-D golden-angle radial trajectory, 2 coils, ramp density, smooth maps
  const IndType W = 40, H = 36, C = 2;
  const int spokes = 48, ns = 64;
  const IndType M = (IndType)(spokes * ns);
  std::vector<DType> traj(2 * M), dens(M);
  for (int s = 0; s < spokes; s++)
    for (int i = 0; i < ns; i++) {
      const double th = s * 1.9416110387254665, r = -0.5 + (double)i / ns;
      traj[s * ns + i] = (DType)(r * std::cos(th));
      traj[M + s * ns + i] = (DType)(r * std::sin(th));
      dens[s * ns + i] = (DType)std::fabs(r) + 0.01f;
    }
  Dimensions img(W, H);
  const size_t Ni = (size_t)W * H;
  std::vector<DType2> sens(C * Ni);
  for (IndType c = 0; c < C; c++)
    for (size_t p = 0; p < Ni; p++) {
      const double x = (double)(p % W) / W, y = (double)(p / W) / H;
      const double m = c == 0 ? 1.0 / (1.0 + x * x + y * y) : 1.0 / (1.0 + (1 - x) * (1 - x) + y);
      sens[c * Ni + p].x = (float)(m * std::cos(0.3 * c + x));
      sens[c * Ni + p].y = (float)(m * std::sin(0.3 * c + x));
    }
  std::mt19937 rng(7);
  std::normal_distribution<float> nd;
  std::vector<DType2> kdata((size_t)C * M), image((size_t)C * Ni);
  for (auto& v : kdata) { v.x = nd(rng); v.y = nd(rng); }
  for (auto& v : image) { v.x = nd(rng); v.y = nd(rng); }

  Array<DType> trajA;
  trajA.data = traj.data();
  trajA.dim.length = M;
  Array<DType> densA;
  densA.data = dens.data();
  densA.dim.length = M;
  Array<DType2> sensA;
  sensA.data = sens.data();
  sensA.dim = img;
  sensA.dim.channels = C;
  Array<DType2> kA;
  kA.data = kdata.data();
  kA.dim.length = M;
  kA.dim.channels = C;
  Array<DType2> imA;
  imA.data = image.data();
  imA.dim = img;
  imA.dim.channels = C;

  try {
    // flavour per factory flags (gpuNUFFT_operator_factory.cpp:324-357)
    const OperatorType expect[4] = {BALANCED_TEXTURE, BALANCED, TEXTURE, DEFAULT};
    for (int f = 0; f < 4; f++) {
      GpuNUFFTOperatorFactory fac(f % 2 == 0, true, f < 2);
      GpuNUFFTOperator* o = fac.createGpuNUFFTOperator(trajA, 3, 8, 2.0f, img);
      check(o->getType() == expect[f], "operator type for factory flags " + std::to_string(f));
      check(o->getGridDims().width == 80 && o->getGridDims().height == 72, "grid dims");
      check(o->getGridSectorDims().width == 10 && o->getGridSectorDims().height == 9,
            "grid sector dims");
      delete o;
    }

    GpuNUFFTOperatorFactory fac(false, true, true);  // balanced, nearest lookup
    GpuNUFFTOperator* op = fac.createGpuNUFFTOperator(trajA, densA, 3, 8, 2.0f, img);
    check(op->applyDensComp() && !op->applySensData(), "dens on, sens off");

    // host paths
    Array<CufftType> adj = op->performGpuNUFFTAdj(kA);
    check(adj.dim.channels == C && adj.count() == C * Ni, "adjoint output shape (no sens)");
    Array<CufftType> fwd = op->performForwardGpuNUFFT(imA);
    check(fwd.dim.channels == C && fwd.dim.length == M, "forward output shape");

    // adjoint identity <A x, y> = <x, A^H y> per coil set
    {
      const std::complex<double> lhs = vdot(fwd.data, kdata.data(), (size_t)C * M);
      const std::complex<double> rhs = vdot(image.data(), adj.data, (size_t)C * Ni);
      check(std::abs(lhs - rhs) <= 1e-4 * std::abs(lhs), "adjoint identity");
    }

    // GpuArray overloads give the host results
    {
      GpuArray<DType2> kd;
      kd.data = (DType2*)to_device(kdata.data(), kdata.size() * sizeof(DType2));
      kd.dim = kA.dim;
      GpuArray<CufftType> od;
      checkStatus(gnufft_malloc((void**)&od.data, C * Ni * sizeof(CufftType)), "malloc");
      od.dim = img;
      od.dim.channels = C;
      op->performGpuNUFFTAdj(kd, od);
      std::vector<CufftType> h(C * Ni);
      to_host(h.data(), od.data, h.size() * sizeof(CufftType));
      check(rel_l2(h.data(), adj.data, h.size()) < 1e-6, "GpuArray adjoint == host adjoint");

      GpuArray<DType2> id;
      id.data = (DType2*)to_device(image.data(), image.size() * sizeof(DType2));
      id.dim = imA.dim;
      GpuArray<CufftType> kout;
      checkStatus(gnufft_malloc((void**)&kout.data, (size_t)C * M * sizeof(CufftType)), "malloc");
      kout.dim = kA.dim;
      op->performForwardGpuNUFFT(id, kout);
      std::vector<CufftType> hk((size_t)C * M);
      to_host(hk.data(), kout.data, hk.size() * sizeof(CufftType));
      check(rel_l2(hk.data(), fwd.data, hk.size()) < 1e-6, "GpuArray forward == host forward");
      gnufft_free(kd.data);
      gnufft_free(od.data);
      gnufft_free(id.data);
      gnufft_free(kout.data);
    }

    // stop stages
    {
      Array<CufftType> g = op->performGpuNUFFTAdj(kA, CONVOLUTION);
      check(g.dim.width == 80 && g.dim.height == 72 && g.dim.channels == C, "CONVOLUTION shape");
      double e = 0;
      for (size_t i = 0; i < g.count(); i++) e += (double)g.data[i].x * g.data[i].x;
      check(e > 0, "CONVOLUTION grid non-zero");
      free(g.data);
      Array<CufftType> f = op->performGpuNUFFTAdj(kA, FFT);
      check(f.count() == C * Ni, "FFT stop allocation");
      free(f.data);
    }

    // tables -> loadPrecomputedGpuNUFFTOperator reproduces the operator
    {
      BalancedOperator* b = dynamic_cast<BalancedOperator*>(op);
      check(b != nullptr, "balanced interface");
      Array<IndType2> order = b->getSectorProcessingOrder();
      Array<IndType> sdc = op->getSectorDataCount();
      IndType items = 0;
      for (IndType s = 0; s + 1 < sdc.count(); s++) {
        const IndType n = sdc.data[s + 1] - sdc.data[s];
        items += n ? (n + MAXIMUM_PAYLOAD - 1) / MAXIMUM_PAYLOAD : 0;
      }
      check(order.count() == items, "processing order length");
      bool desc = true;
      for (IndType i = 1; i < order.count(); i++) {
        const IndType a = order.data[i - 1].x, c = order.data[i].x;
        desc = desc && (sdc.data[a + 1] - sdc.data[a]) >= (sdc.data[c + 1] - sdc.data[c]);
      }
      check(desc, "processing order by descending count");
      Array<DType> tr = op->getKSpaceTraj();
      Array<IndType> idx = op->getDataIndices();
      Array<IndType> cen = op->getSectorCenters();
      Array<DType> dn = op->getDens();
      Array<DType> dp = op->getDeapodizationFunction();
      Array<DType2> noSens;
      GpuNUFFTOperator* op2 = fac.loadPrecomputedGpuNUFFTOperator(
          tr, idx, sdc, order, cen, dn, noSens, dp, 3, 8, 2.0f, img);
      Array<CufftType> adj2 = op2->performGpuNUFFTAdj(kA);
      check(rel_l2(adj2.data, adj.data, C * Ni) < 1e-6, "precomputed operator adjoint");
      Array<CufftType> fwd2 = op2->performForwardGpuNUFFT(imA);
      check(rel_l2(fwd2.data, fwd.data, (size_t)C * M) < 1e-6, "precomputed operator forward");
      free(adj2.data);
      free(fwd2.data);
      delete op2;
    }
    free(adj.data);
    free(fwd.data);

    // sensitivities: coil-combined adjoint, forward replicates one image
    {
      GpuNUFFTOperator* os = fac.createGpuNUFFTOperator(trajA, densA, sensA, 3, 8, 2.0f, img);
      check(os->applySensData(), "sens on");
      Array<CufftType> a = os->performGpuNUFFTAdj(kA);
      check(a.dim.channels == 1 && a.count() == Ni, "coil-combined adjoint shape");
      Array<DType2> one;
      one.data = image.data();
      one.dim = img;
      Array<CufftType> k = os->performForwardGpuNUFFT(one);
      check(k.dim.channels == C, "forward with sens: one output per coil");
      const std::complex<double> lhs = vdot(k.data, kdata.data(), (size_t)C * M);
      const std::complex<double> rhs = vdot(image.data(), a.data, Ni);
      check(std::abs(lhs - rhs) <= 1e-4 * std::abs(lhs), "adjoint identity with sens");
      free(a.data);
      free(k.data);
      delete os;
    }

    // invalid arguments throw std::invalid_argument
    {
      bool threw = false;
      Array<DType> bad = trajA;
      bad.dim.channels = 2;
      try {
        delete fac.createGpuNUFFTOperator(bad, 3, 8, 2.0f, img);
      } catch (const std::invalid_argument&) {
        threw = true;
      }
      check(threw, "trajectory with channels > 1 rejected");
      threw = false;
      try {
        GpuNUFFTOperator o(9, 8, 2.0f, img);
      } catch (const std::runtime_error&) {
        threw = true;
      }
      check(threw, "kernel width 9 rejected");
    }
    delete op;
  } catch (const std::exception& e) {
    std::printf("FAIL exception: %s\n", e.what());
    return 1;
  }
  if (failures) {
    std::printf("CPP_API_CHECK FAILED (%d)\n", failures);
    return 1;
  }
  std::printf("CPP_API_CHECK OK\n");
  return 0;
}
