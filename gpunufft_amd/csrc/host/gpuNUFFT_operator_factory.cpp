// gpuNUFFT C++ API: the operator factory (reference
// CUDA/src/gpuNUFFT_operator_factory.cpp) over the C ABI of
// include/gpunufft_hip.h. The sector precompute and the deapodization
// function are computed on the device by the plan.

#include "gpuNUFFT_operator_factory.hpp"

#include <cmath>
#include <sstream>
#include <stdexcept>

#include "gpunufft_hip.h"

namespace gpuNUFFT {

Dimensions GpuNUFFTOperatorFactory::computeSectorCountPerDimension(Dimensions dim,
                                                                   IndType sectorWidth) {
  Dimensions s;
  s.width = (IndType)std::ceil((DType)dim.width / (DType)sectorWidth);
  s.height = (IndType)std::ceil((DType)dim.height / (DType)sectorWidth);
  s.depth = (IndType)std::ceil((DType)dim.depth / (DType)sectorWidth);
  return s;
}

GpuNUFFTOperator* GpuNUFFTOperatorFactory::createNewGpuNUFFTOperator(IndType kernelWidth,
                                                                     IndType sectorWidth,
                                                                     DType osf,
                                                                     Dimensions imgDims) {
  if (balanceWorkload) {
    if (useTextures)
      return new BalancedTextureGpuNUFFTOperator(kernelWidth, sectorWidth, osf, imgDims,
                                                 TEXTURE2D_LOOKUP, matlabSharedMem);
    return new BalancedGpuNUFFTOperator(kernelWidth, sectorWidth, osf, imgDims, matlabSharedMem);
  }
  if (useTextures)
    return new TextureGpuNUFFTOperator(kernelWidth, sectorWidth, osf, imgDims, TEXTURE2D_LOOKUP,
                                       matlabSharedMem);
  return new GpuNUFFTOperator(kernelWidth, sectorWidth, osf, imgDims, true, DEFAULT,
                              matlabSharedMem);
}

// Estimate of the device memory an operator needs, checked against the
// device's total (same estimate as CUDA/src/gpuNUFFT_operator_factory.cpp:628-663).
void GpuNUFFTOperatorFactory::checkMemoryConsumption(Dimensions& kSpaceDims,
                                                     const IndType& /*sectorWidth*/,
                                                     const DType& osf, Dimensions& imgDims,
                                                     Dimensions& densDims, Dimensions& sensDims) {
  const size_t real = sizeof(DType), cplx = 2 * sizeof(DType);
  size_t est = 2 * (size_t)(imgDims.count() * std::pow(osf, 3)) * cplx;  // grid + FFT
  est += (size_t)kSpaceDims.count() * (3 * real + cplx);                 // coords + samples
  est += (size_t)densDims.count() * real;
  if (sensDims.count() > 0) est += 3 * (size_t)imgDims.count() * cplx;   // sens, deapo, sum
  size_t freeB = 0, total = 0;
  checkStatus(gnufft_mem_info(&freeB, &total), "checkMemoryConsumption");
  if (est > total) {
    std::stringstream ss;
    ss << "Required device memory too large for selected device!\n"
       << "Total available memory: " << total << "\nRequired memory: " << est << "\n";
    throw std::runtime_error(ss.str());
  }
}

GpuNUFFTOperator* GpuNUFFTOperatorFactory::createGpuNUFFTOperator(
    Array<DType>& kSpaceTraj, Array<DType>& densCompData, Array<DType2>& sensData,
    const IndType& kernelWidth, const IndType& sectorWidth, const DType& osf, Dimensions& imgDims) {
  checkMemoryConsumption(kSpaceTraj.dim, sectorWidth, osf, imgDims, densCompData.dim,
                         sensData.dim);
  if (kSpaceTraj.dim.channels > 1)
    throw std::invalid_argument(
        "Trajectory dimension must not contain a channel size greater than 1!");
  if (imgDims.channels > 1)
    throw std::invalid_argument("Image dimensions must not contain a channel size greater than 1!");
  GpuNUFFTOperator* op = createNewGpuNUFFTOperator(kernelWidth, sectorWidth, osf, imgDims);
  try {
    op->precompute(kSpaceTraj, densCompData);
    if (sensData.data != NULL) op->setSens(sensData);
  } catch (...) {
    delete op;
    throw;
  }
  return op;
}

GpuNUFFTOperator* GpuNUFFTOperatorFactory::createGpuNUFFTOperator(
    Array<DType>& kSpaceTraj, Array<DType>& densCompData, const IndType& kernelWidth,
    const IndType& sectorWidth, const DType& osf, Dimensions& imgDims) {
  Array<DType2> noSens;
  return createGpuNUFFTOperator(kSpaceTraj, densCompData, noSens, kernelWidth, sectorWidth, osf,
                                imgDims);
}

GpuNUFFTOperator* GpuNUFFTOperatorFactory::createGpuNUFFTOperator(
    Array<DType>& kSpaceTraj, const IndType& kernelWidth, const IndType& sectorWidth,
    const DType& osf, Dimensions& imgDims) {
  Array<DType> noDens;
  return createGpuNUFFTOperator(kSpaceTraj, noDens, kernelWidth, sectorWidth, osf, imgDims);
}

GpuNUFFTOperator* GpuNUFFTOperatorFactory::loadPrecomputedGpuNUFFTOperator(
    Array<DType>& kSpaceTraj, Array<IndType>& dataIndices, Array<IndType>& sectorDataCount,
    Array<IndType2>& sectorProcessingOrder, Array<IndType>& sectorCenters,
    Array<DType2>& sensData, Array<DType>& deapoData, const IndType& kernelWidth,
    const IndType& sectorWidth, const DType& osf, Dimensions& imgDims) {
  GpuNUFFTOperator* op = createNewGpuNUFFTOperator(kernelWidth, sectorWidth, osf, imgDims);
  op->setGridSectorDims(computeSectorCountPerDimension(op->getGridDims(), op->getSectorWidth()));
  op->setKSpaceTraj(kSpaceTraj);
  op->setDataIndices(dataIndices);
  op->setSectorDataCount(sectorDataCount);
  if (BalancedOperator* b = dynamic_cast<BalancedOperator*>(op))
    b->setSectorProcessingOrder(sectorProcessingOrder);
  op->setSectorCenters(sectorCenters);
  op->setSens(sensData);
  op->setDeapodizationFunction(deapoData);
  return op;
}

GpuNUFFTOperator* GpuNUFFTOperatorFactory::loadPrecomputedGpuNUFFTOperator(
    Array<DType>& kSpaceTraj, Array<IndType>& dataIndices, Array<IndType>& sectorDataCount,
    Array<IndType2>& sectorProcessingOrder, Array<IndType>& sectorCenters,
    Array<DType>& densCompData, Array<DType2>& sensData, Array<DType>& deapoData,
    const IndType& kernelWidth, const IndType& sectorWidth, const DType& osf,
    Dimensions& imgDims) {
  GpuNUFFTOperator* op = loadPrecomputedGpuNUFFTOperator(
      kSpaceTraj, dataIndices, sectorDataCount, sectorProcessingOrder, sectorCenters, sensData,
      deapoData, kernelWidth, sectorWidth, osf, imgDims);
  op->setDens(densCompData);
  return op;
}

}  // namespace gpuNUFFT
