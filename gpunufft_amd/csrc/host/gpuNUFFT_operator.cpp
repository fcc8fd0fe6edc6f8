// gpuNUFFT C++ API: the operator (reference CUDA/inc/gpuNUFFT_operator.hpp,
// CUDA/src/gpuNUFFT_operator.cpp, CUDA/src/texture_gpuNUFFT_operator.cpp,
// CUDA/src/balanced_*_operator.cpp) implemented over the C ABI of
// include/gpunufft_hip.h. Every call runs on the default stream of the
// current device and returns when its results are in place, as the
// reference's do.

#include "gpuNUFFT_operator.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "balanced_gpuNUFFT_operator.hpp"
#include "balanced_texture_gpuNUFFT_operator.hpp"
#include "gpunufft_hip.h"
#include "texture_gpuNUFFT_operator.hpp"

namespace gpuNUFFT {

void checkStatus(int status, const std::string& where) {
  if (status == GNUFFT_OK) return;
  const std::string msg = where + ": " + gnufft_last_error();
  if (status == GNUFFT_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

namespace {

template <typename T>
T* host_alloc(size_t n) {
  T* p = static_cast<T*>(calloc(n ? n : 1, sizeof(T)));
  if (!p) throw std::runtime_error("gpuNUFFT: host allocation failed");
  return p;
}

// Device buffer released on scope exit.
class DeviceBuffer {
 public:
  explicit DeviceBuffer(size_t bytes) {
    if (bytes) checkStatus(gnufft_malloc(&p_, bytes), "gpuNUFFT: device allocation");
  }
  ~DeviceBuffer() {
    if (p_) gnufft_free(p_);
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  void* get() const { return p_; }

 private:
  void* p_ = nullptr;
};

gnufft_plan_info plan_info(gnufft_plan p) {
  gnufft_plan_info i;
  checkStatus(gnufft_plan_get_info(p, &i), "gpuNUFFT: plan info");
  return i;
}

}  // namespace

// ---------------------------------------------------------------------------
// construction / destruction
// ---------------------------------------------------------------------------
GpuNUFFTOperator::GpuNUFFTOperator(IndType kernelWidth, IndType sectorWidth, DType osf,
                                   Dimensions imgDims, bool loadKernel, OperatorType operatorType,
                                   bool matlabSharedMem)
    : operatorType(operatorType),
      osf(osf),
      kernelWidth(kernelWidth),
      sectorWidth(sectorWidth),
      imgDims(imgDims),
      matlabSharedMem(matlabSharedMem),
      owned(0),
      plan_(nullptr),
      tablesOnDevice(false),
      tablesOnHost(false),
      sensDirty(false),
      nSamples(0) {
  sectorDims.width = sectorWidth;
  sectorDims.height = sectorWidth;
  if (imgDims.depth > 0) sectorDims.depth = sectorWidth;
  const uint32_t dims[3] = {imgDims.width, imgDims.height, imgDims.depth};
  const int lookup = (operatorType == TEXTURE || operatorType == BALANCED_TEXTURE)
                         ? GNUFFT_LOOKUP_LINEAR
                         : GNUFFT_LOOKUP_NEAREST;
  checkStatus(gnufft_plan_create(&plan_, dims, kernelWidth, sectorWidth, osf, lookup, nullptr),
              "GpuNUFFTOperator");
  const gnufft_plan_info i = plan_info(plan_);
  gridSectorDims.width = i.sectors[0];
  gridSectorDims.height = i.sectors[1];
  gridSectorDims.depth = i.dim == 3 ? i.sectors[2] : 0;
  if (loadKernel) GpuNUFFTOperator::initKernel();
}

GpuNUFFTOperator::~GpuNUFFTOperator() {
  releaseOwned(~0u);
  if (plan_) gnufft_plan_destroy(plan_);
}

void GpuNUFFTOperator::releaseOwned(unsigned which) {
  struct Slot {
    unsigned bit;
    void** data;
  } slots[] = {
      {OWN_KERNEL, reinterpret_cast<void**>(&kernel.data)},
      {OWN_TRAJ, reinterpret_cast<void**>(&kSpaceTraj.data)},
      {OWN_DENS, reinterpret_cast<void**>(&dens.data)},
      {OWN_DEAPO, reinterpret_cast<void**>(&deapo.data)},
      {OWN_CENTERS, reinterpret_cast<void**>(&sectorCenters.data)},
      {OWN_INDICES, reinterpret_cast<void**>(&dataIndices.data)},
      {OWN_COUNT, reinterpret_cast<void**>(&sectorDataCount.data)},
  };
  for (const Slot& s : slots)
    if ((which & s.bit) && (owned & s.bit)) {
      free(*s.data);
      *s.data = nullptr;
      owned &= ~s.bit;
    }
}

gnufft_plan_s* GpuNUFFTOperator::plan() { return plan_; }

// 1-D lookup table of the plan (load1DKernel, CUDA/src/gpuNUFFT_operator.cpp:43-51).
void GpuNUFFTOperator::initKernel() {
  releaseOwned(OWN_KERNEL);
  const gnufft_plan_info i = plan_info(plan_);
  kernel = Array<DType>();
  kernel.dim.length = i.kernel_count;
  kernel.data = host_alloc<DType>(i.kernel_count);
  owned |= OWN_KERNEL;
  checkStatus(gnufft_plan_get_tables(plan_, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                     kernel.data),
              "initKernel");
}

// Texture flavours expose the product table T[i]*T[j]
// (load2DKernel, CUDA/src/gpuNUFFT_utils.cpp:109-124); the kernels read the
// 1-D table, whose outer product this is.
void TextureGpuNUFFTOperator::initKernel() {
  GpuNUFFTOperator::initKernel();
  if (interpolationType == CONST_LOOKUP || interpolationType == TEXTURE_LOOKUP) return;
  const IndType K = kernel.dim.length;
  const int planes = interpolationType == TEXTURE3D_LOOKUP ? (int)K : 1;
  DType* t = host_alloc<DType>((size_t)K * K * planes);
  for (int k = 0; k < planes; k++)
    for (IndType j = 0; j < K; j++)
      for (IndType i = 0; i < K; i++)
        t[((size_t)k * K + j) * K + i] =
            planes > 1 ? kernel.data[k] * kernel.data[j] * kernel.data[i] : kernel.data[j] * kernel.data[i];
  releaseOwned(OWN_KERNEL);
  kernel = Array<DType>();
  kernel.dim.width = K;
  kernel.dim.height = K;
  kernel.dim.depth = planes > 1 ? K : 1;
  kernel.data = t;
  owned |= OWN_KERNEL;
}

// ---------------------------------------------------------------------------
// setters: borrowed arrays, uploaded at the next perform call
// ---------------------------------------------------------------------------
void GpuNUFFTOperator::setOsf(DType o) { osf = o; }

void GpuNUFFTOperator::setKSpaceTraj(Array<DType> a) {
  fetchTables();  // keep the other tables when replacing one
  releaseOwned(OWN_TRAJ);
  kSpaceTraj = a;
  tablesOnDevice = false;
}

void GpuNUFFTOperator::setSectorCenters(Array<IndType> a) {
  fetchTables();  // keep the other tables when replacing one
  releaseOwned(OWN_CENTERS);
  sectorCenters = a;
  tablesOnDevice = false;
}

void GpuNUFFTOperator::setSectorDataCount(Array<IndType> a) {
  fetchTables();  // keep the other tables when replacing one
  releaseOwned(OWN_COUNT);
  sectorDataCount = a;
  tablesOnDevice = false;
}

void GpuNUFFTOperator::setDataIndices(Array<IndType> a) {
  fetchTables();  // keep the other tables when replacing one
  releaseOwned(OWN_INDICES);
  dataIndices = a;
  tablesOnDevice = false;
}

void GpuNUFFTOperator::setSens(Array<DType2> a) {
  sens = a;
  sensDirty = true;
}

void GpuNUFFTOperator::setDens(Array<DType> a) {
  fetchTables();  // keep the other tables when replacing one
  releaseOwned(OWN_DENS);
  dens = a;
  tablesOnDevice = false;
}

void GpuNUFFTOperator::setDeapodizationFunction(Array<DType> a) {
  fetchTables();  // keep the other tables when replacing one
  releaseOwned(OWN_DEAPO);
  deapo = a;
  tablesOnDevice = false;
}

void GpuNUFFTOperator::setImageDims(Dimensions d) { imgDims = d; }

void GpuNUFFTOperator::setGridSectorDims(Dimensions d) { gridSectorDims = d; }

// ---------------------------------------------------------------------------
// getters
// ---------------------------------------------------------------------------
Dimensions GpuNUFFTOperator::getGridDims() { return imgDims * osf; }

Dimensions GpuNUFFTOperator::getGridSectorDims() { return gridSectorDims; }

Array<DType> GpuNUFFTOperator::getKernel() { return kernel; }

Array<DType> GpuNUFFTOperator::getKSpaceTraj() {
  fetchTables();
  return kSpaceTraj;
}

Array<DType> GpuNUFFTOperator::getDens() {
  fetchTables();
  return dens;
}

Array<IndType> GpuNUFFTOperator::getSectorDataCount() {
  fetchTables();
  return sectorDataCount;
}

Array<IndType> GpuNUFFTOperator::getSectorCenters() {
  fetchTables();
  return sectorCenters;
}

Array<IndType> GpuNUFFTOperator::getDataIndices() {
  fetchTables();
  return dataIndices;
}

Array<DType> GpuNUFFTOperator::getDeapodizationFunction() {
  fetchTables();
  return deapo;
}

IndType GpuNUFFTOperator::sampleCount() {
  if (tablesOnDevice) return nSamples;
  return kSpaceTraj.dim.length ? kSpaceTraj.dim.length : kSpaceTraj.count();
}

// Tables computed on the device by the factory: copy to host arrays owned
// by the operator, with the reference's dimension conventions
// (CUDA/src/gpuNUFFT_operator_factory.cpp:272-322).
void GpuNUFFTOperator::fetchTables() {
  if (tablesOnHost || !tablesOnDevice) return;
  const gnufft_plan_info i = plan_info(plan_);
  const IndType M = i.n_samples, S = i.sector_count, d = (IndType)i.dim;
  releaseOwned(OWN_TRAJ | OWN_DENS | OWN_DEAPO | OWN_CENTERS | OWN_INDICES | OWN_COUNT);
  kSpaceTraj = Array<DType>();
  kSpaceTraj.data = host_alloc<DType>((size_t)d * M);
  kSpaceTraj.dim.length = M;
  dataIndices = Array<IndType>();
  dataIndices.data = host_alloc<IndType>(M);
  dataIndices.dim.length = M;
  sectorDataCount = Array<IndType>();
  sectorDataCount.data = host_alloc<IndType>((size_t)S + 1);
  sectorDataCount.dim.length = S + 1;
  sectorCenters = Array<IndType>();
  sectorCenters.data = host_alloc<IndType>((size_t)S * d);
  sectorCenters.dim.length = S * d;
  deapo = Array<DType>();
  deapo.data = host_alloc<DType>((size_t)i.img_count);
  deapo.dim.length = (IndType)i.img_count;
  owned |= OWN_TRAJ | OWN_INDICES | OWN_COUNT | OWN_CENTERS | OWN_DEAPO;
  DType* densSorted = nullptr;
  if (i.has_dens) {
    dens = Array<DType>();
    dens.data = densSorted = host_alloc<DType>(M);
    dens.dim.length = M;
    owned |= OWN_DENS;
  }
  checkStatus(gnufft_plan_get_tables(plan_, dataIndices.data, sectorDataCount.data,
                                     sectorCenters.data, kSpaceTraj.data, densSorted, deapo.data,
                                     nullptr),
              "gpuNUFFT: tables");
  tablesOnHost = true;
}

// Precompute on the device from a host trajectory (the factory's
// assignSectors / sort / count / centers, CUDA/src/gpuNUFFT_operator_factory.cpp:450-553).
void GpuNUFFTOperator::precompute(Array<DType>& traj, Array<DType>& densCompData) {
  const IndType M = traj.dim.length ? traj.dim.length : traj.count();
  const float* densPtr = densCompData.data;
  checkStatus(gnufft_plan_set_trajectory(plan_, traj.data, M, densPtr, GNUFFT_HOST, nullptr),
              "createGpuNUFFTOperator");
  nSamples = M;
  tablesOnDevice = true;
  tablesOnHost = false;
  // dens is applied when given, as in the reference (applyDensComp): the
  // host view reports its presence until the sorted copy is fetched
  if (densPtr) {
    dens = Array<DType>();
    dens.data = densCompData.data;
    dens.dim = densCompData.dim;
  }
}

// Upload tables set through the setters (loadPrecomputedGpuNUFFTOperator)
// and sensitivities.
void GpuNUFFTOperator::syncDevice() {
  if (!tablesOnDevice) {
    const IndType M = sampleCount();
    if (!kSpaceTraj.data || !sectorDataCount.data || !sectorCenters.data ||
        (M && !dataIndices.data))
      throw std::invalid_argument("gpuNUFFT: operator has no trajectory tables");
    checkStatus(gnufft_plan_load_precomputed(plan_, kSpaceTraj.data, dataIndices.data,
                                             sectorDataCount.data, sectorCenters.data,
                                             applyDensComp() ? dens.data : nullptr,
                                             deapo.data, M, GNUFFT_HOST, nullptr),
                "loadPrecomputedGpuNUFFTOperator");
    nSamples = M;
    tablesOnDevice = true;
    tablesOnHost = true;  // host arrays are the caller's tables
  }
  if (sensDirty) {
    if (applySensData()) {
      IndType coils = DEFAULT_VALUE(sens.dim.channels);
      checkStatus(gnufft_plan_set_sens(plan_, sens.data, coils, GNUFFT_HOST, nullptr), "setSens");
    } else {
      checkStatus(gnufft_plan_set_sens(plan_, nullptr, 0, GNUFFT_HOST, nullptr), "setSens");
    }
    sensDirty = false;
  }
}

// ---------------------------------------------------------------------------
// adjoint
// ---------------------------------------------------------------------------
// Output elements the reference writes for a stop stage
// (CUDA/src/gpuNUFFT_operator.cpp:735-751, :800-812, :863-887).
static size_t adjoint_out_count(const gnufft_plan_info& i, IndType coils, bool sens,
                                GpuNUFFTOutput stop) {
  switch (stop) {
    case CONVOLUTION:
    case DENSITY_ESTIMATION: return (size_t)coils * i.grid_count;
    case FFT: return (size_t)i.img_count;
    default: return (size_t)(sens ? 1 : coils) * i.img_count;
  }
}

void GpuNUFFTOperator::adjoint(const void* kdata, void* out, IndType nCoils, GpuNUFFTOutput stop,
                               bool host, size_t outCount) {
  syncDevice();
  const gnufft_plan_info i = plan_info(plan_);
  const size_t M = i.n_samples;
  const bool sensOn = applySensData() && stop == DEAPODIZATION;
  // the C ABI returns every coil at the FFT stop; the reference returns the first
  const size_t abiCount = stop == FFT ? (size_t)nCoils * i.img_count
                                      : adjoint_out_count(i, nCoils, sensOn, stop);
  if (host) {
    DeviceBuffer k(M * nCoils * sizeof(CufftType));
    DeviceBuffer o(abiCount * sizeof(CufftType));
    checkStatus(gnufft_memcpy(k.get(), kdata, M * nCoils * sizeof(CufftType), GNUFFT_DEVICE,
                              GNUFFT_HOST, nullptr),
                "performGpuNUFFTAdj");
    checkStatus(gnufft_adjoint(plan_, k.get(), o.get(), nCoils, (int)stop, nullptr),
                "performGpuNUFFTAdj");
    checkStatus(gnufft_memcpy(out, o.get(), std::min(outCount, abiCount) * sizeof(CufftType),
                              GNUFFT_HOST, GNUFFT_DEVICE, nullptr),
                "performGpuNUFFTAdj");
    return;
  }
  if (stop == FFT && nCoils > 1) {
    DeviceBuffer o(abiCount * sizeof(CufftType));
    checkStatus(gnufft_adjoint(plan_, kdata, o.get(), nCoils, (int)stop, nullptr),
                "performGpuNUFFTAdj");
    checkStatus(gnufft_memcpy(out, o.get(), (size_t)i.img_count * sizeof(CufftType), GNUFFT_DEVICE,
                              GNUFFT_DEVICE, nullptr),
                "performGpuNUFFTAdj");
  } else {
    checkStatus(gnufft_adjoint(plan_, kdata, out, nCoils, (int)stop, nullptr), "performGpuNUFFTAdj");
  }
  checkStatus(gnufft_stream_synchronize(nullptr), "performGpuNUFFTAdj");
}

void GpuNUFFTOperator::performGpuNUFFTAdj(Array<DType2> kspaceData, Array<CufftType>& imgData,
                                          GpuNUFFTOutput stop) {
  const IndType coils = DEFAULT_VALUE(kspaceData.dim.channels);
  syncDevice();
  const gnufft_plan_info i = plan_info(plan_);
  const size_t need = adjoint_out_count(i, coils, applySensData() && stop == DEAPODIZATION, stop);
  if (!imgData.data) throw std::invalid_argument("performGpuNUFFTAdj: no output array");
  adjoint(kspaceData.data, imgData.data, coils, stop, true, need);
}

void GpuNUFFTOperator::performGpuNUFFTAdj(GpuArray<DType2> kspaceData_gpu,
                                          GpuArray<CufftType>& imgData_gpu, GpuNUFFTOutput stop) {
  const IndType coils = DEFAULT_VALUE(kspaceData_gpu.dim.channels);
  if (!imgData_gpu.data) throw std::invalid_argument("performGpuNUFFTAdj: no output array");
  adjoint(kspaceData_gpu.data, imgData_gpu.data, coils, stop, false, 0);
}

Array<CufftType> GpuNUFFTOperator::performGpuNUFFTAdj(Array<DType2> kspaceData,
                                                      GpuNUFFTOutput stop) {
  Array<CufftType> img;
  if (stop == CONVOLUTION || stop == DENSITY_ESTIMATION) {
    img.dim = getGridDims();
    img.dim.channels = kspaceData.dim.channels;
  } else {
    img.dim = getImageDims();
    // with sensitivities the coils are summed into one image
    img.dim.channels = applySensData() ? 1 : kspaceData.dim.channels;
  }
  img.data = host_alloc<CufftType>(img.count());
  try {
    performGpuNUFFTAdj(kspaceData, img, stop);
  } catch (...) {
    free(img.data);
    throw;
  }
  return img;
}

Array<CufftType> GpuNUFFTOperator::performGpuNUFFTAdj(Array<DType2> kspaceData) {
  return performGpuNUFFTAdj(kspaceData, DEAPODIZATION);
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
void GpuNUFFTOperator::forward(const void* img, void* out, IndType nCoils, GpuNUFFTOutput stop,
                               bool host, IndType imgChannels) {
  syncDevice();
  const gnufft_plan_info i = plan_info(plan_);
  const size_t M = i.n_samples;
  // DENSITY_ESTIMATION reads the image as the grid; every other stop is the
  // full forward operator (CUDA/src/gpuNUFFT_operator.cpp:1143-1261)
  const int abiStop = stop == DENSITY_ESTIMATION ? GNUFFT_STOP_DENSITY_ESTIMATION
                                                 : GNUFFT_STOP_DEAPODIZATION;
  if (host) {
    const size_t imgBytes = (size_t)imgChannels * i.img_count * sizeof(CufftType);
    DeviceBuffer im(imgBytes);
    DeviceBuffer k(M * nCoils * sizeof(CufftType));
    checkStatus(gnufft_memcpy(im.get(), img, imgBytes, GNUFFT_DEVICE, GNUFFT_HOST, nullptr),
                "performForwardGpuNUFFT");
    checkStatus(gnufft_forward(plan_, im.get(), k.get(), nCoils, abiStop, nullptr),
                "performForwardGpuNUFFT");
    checkStatus(gnufft_memcpy(out, k.get(), M * nCoils * sizeof(CufftType), GNUFFT_HOST,
                              GNUFFT_DEVICE, nullptr),
                "performForwardGpuNUFFT");
    return;
  }
  checkStatus(gnufft_forward(plan_, img, out, nCoils, abiStop, nullptr), "performForwardGpuNUFFT");
  checkStatus(gnufft_stream_synchronize(nullptr), "performForwardGpuNUFFT");
}

void GpuNUFFTOperator::performForwardGpuNUFFT(Array<DType2> imgData, Array<CufftType>& kspaceData,
                                              GpuNUFFTOutput stop) {
  if (!kspaceData.data) throw std::invalid_argument("performForwardGpuNUFFT: no output array");
  const IndType imgCh = DEFAULT_VALUE(imgData.dim.channels);
  const IndType coils = applySensData() ? DEFAULT_VALUE(sens.dim.channels) : imgCh;
  forward(imgData.data, kspaceData.data, coils, stop, true, applySensData() ? 1 : imgCh);
}

void GpuNUFFTOperator::performForwardGpuNUFFT(GpuArray<DType2> imgData_gpu,
                                              GpuArray<CufftType>& kspaceData_gpu,
                                              GpuNUFFTOutput stop) {
  if (!kspaceData_gpu.data) throw std::invalid_argument("performForwardGpuNUFFT: no output array");
  const IndType imgCh = DEFAULT_VALUE(imgData_gpu.dim.channels);
  const IndType coils = applySensData() ? DEFAULT_VALUE(sens.dim.channels) : imgCh;
  forward(imgData_gpu.data, kspaceData_gpu.data, coils, stop, false, imgCh);
}

Array<CufftType> GpuNUFFTOperator::performForwardGpuNUFFT(Array<DType2> imgData,
                                                          GpuNUFFTOutput stop) {
  syncDevice();
  Array<CufftType> k;
  k.dim.length = sampleCount();
  k.dim.channels = applySensData() ? sens.dim.channels : imgData.dim.channels;
  k.data = host_alloc<CufftType>((size_t)sampleCount() * DEFAULT_VALUE(k.dim.channels));
  try {
    performForwardGpuNUFFT(imgData, k, stop);
  } catch (...) {
    free(k.data);
    throw;
  }
  return k;
}

Array<CufftType> GpuNUFFTOperator::performForwardGpuNUFFT(Array<DType2> imgData) {
  return performForwardGpuNUFFT(imgData, CONVOLUTION);
}

// ---------------------------------------------------------------------------
// balanced flavours
// ---------------------------------------------------------------------------
// Sectors by sample count, descending (std::sort with the reference's
// comparator, so ties fall as in the reference), empty sectors dropped,
// split every MAXIMUM_PAYLOAD samples
// (CUDA/src/gpuNUFFT_operator_factory.cpp:74-126).
Array<IndType2> computeSectorProcessingOrder(Array<IndType> sdc) {
  std::vector<IndPair> perSector;
  const IndType S = sdc.count() ? sdc.count() - 1 : 0;
  perSector.reserve(S);
  for (IndType s = 0; s < S; s++) perSector.push_back(IndPair(s, sdc.data[s + 1] - sdc.data[s]));
  std::sort(perSector.begin(), perSector.end(), std::greater<IndPair>());
  std::vector<IndType2> order;
  for (const IndPair& p : perSector) {
    if (p.second == 0) break;
    for (IndType off = 0; off == 0 || off < p.second; off += MAXIMUM_PAYLOAD) {
      IndType2 item;
      item.x = p.first;
      item.y = off;
      order.push_back(item);
    }
  }
  Array<IndType2> out;
  out.data = host_alloc<IndType2>(order.size());
  out.dim.length = (IndType)order.size();
  if (!order.empty()) memcpy(out.data, order.data(), order.size() * sizeof(IndType2));
  return out;
}

BalancedGpuNUFFTOperator::~BalancedGpuNUFFTOperator() {
  if (ownOrder) free(sectorProcessingOrder.data);
}

Array<IndType2> BalancedGpuNUFFTOperator::getSectorProcessingOrder() {
  if (!sectorProcessingOrder.data) {
    sectorProcessingOrder = computeSectorProcessingOrder(getSectorDataCount());
    ownOrder = true;
  }
  return sectorProcessingOrder;
}

void BalancedGpuNUFFTOperator::setSectorProcessingOrder(Array<IndType2> order) {
  if (ownOrder) free(sectorProcessingOrder.data);
  ownOrder = false;
  sectorProcessingOrder = order;
}

BalancedTextureGpuNUFFTOperator::~BalancedTextureGpuNUFFTOperator() {
  if (ownOrder) free(sectorProcessingOrder.data);
}

Array<IndType2> BalancedTextureGpuNUFFTOperator::getSectorProcessingOrder() {
  if (!sectorProcessingOrder.data) {
    sectorProcessingOrder = computeSectorProcessingOrder(getSectorDataCount());
    ownOrder = true;
  }
  return sectorProcessingOrder;
}

void BalancedTextureGpuNUFFTOperator::setSectorProcessingOrder(Array<IndType2> order) {
  if (ownOrder) free(sectorProcessingOrder.data);
  ownOrder = false;
  sectorProcessingOrder = order;
}

}  // namespace gpuNUFFT
