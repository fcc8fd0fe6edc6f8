/*
 * gpuNUFFT_operator_factory.hpp — operator factory of the C++ API
 * (reference CUDA/inc/gpuNUFFT_operator_factory.hpp:40-240).
 *
 * createGpuNUFFTOperator runs the sector precompute (assign, sort, count,
 * centers) and the deapodization precompute on the device; the reference
 * does most of it on the host (CUDA/src/gpuNUFFT_operator_factory.cpp:450-553).
 * useGpu is accepted for compatibility (the device path is always used).
 */
#ifndef GPUNUFFT_AMD_OPERATOR_FACTORY_HPP
#define GPUNUFFT_AMD_OPERATOR_FACTORY_HPP

#include <string>

#include "balanced_gpuNUFFT_operator.hpp"
#include "balanced_texture_gpuNUFFT_operator.hpp"
#include "gpuNUFFT_operator.hpp"
#include "texture_gpuNUFFT_operator.hpp"

namespace gpuNUFFT {

class GpuNUFFTOperatorFactory {
 public:
  GpuNUFFTOperatorFactory(const bool useTextures = true, const bool useGpu = true,
                          bool balanceWorkload = true, bool matlabSharedMem = false)
      : useTextures(useTextures), useGpu(useGpu), balanceWorkload(balanceWorkload),
        matlabSharedMem(matlabSharedMem) {}
  ~GpuNUFFTOperatorFactory() {}

  /* kSpaceTraj: float SoA [x..., y..., (z...)], dim.length = M, in
   * [-0.5, 0.5]; densCompData: float[M] (sqrt is applied per sample);
   * sensData: complex [coils][img]. The caller deletes the operator. */
  GpuNUFFTOperator* createGpuNUFFTOperator(Array<DType>& kSpaceTraj, const IndType& kernelWidth,
                                           const IndType& sectorWidth, const DType& osf,
                                           Dimensions& imgDims);
  GpuNUFFTOperator* createGpuNUFFTOperator(Array<DType>& kSpaceTraj, Array<DType>& densCompData,
                                           const IndType& kernelWidth, const IndType& sectorWidth,
                                           const DType& osf, Dimensions& imgDims);
  GpuNUFFTOperator* createGpuNUFFTOperator(Array<DType>& kSpaceTraj, Array<DType>& densCompData,
                                           Array<DType2>& sensData, const IndType& kernelWidth,
                                           const IndType& sectorWidth, const DType& osf,
                                           Dimensions& imgDims);

  /* Operator from precomputed tables (sorted trajectory, data indices,
   * sector data count, processing order, centers, sorted density, sens,
   * deapodization) — CUDA/src/gpuNUFFT_operator_factory.cpp:555-626. */
  GpuNUFFTOperator* loadPrecomputedGpuNUFFTOperator(
      Array<DType>& kSpaceTraj, Array<IndType>& dataIndices, Array<IndType>& sectorDataCount,
      Array<IndType2>& sectorProcessingOrder, Array<IndType>& sectorCenters,
      Array<DType>& densCompData, Array<DType2>& sensData, Array<DType>& deapoData,
      const IndType& kernelWidth, const IndType& sectorWidth, const DType& osf,
      Dimensions& imgDims);
  GpuNUFFTOperator* loadPrecomputedGpuNUFFTOperator(
      Array<DType>& kSpaceTraj, Array<IndType>& dataIndices, Array<IndType>& sectorDataCount,
      Array<IndType2>& sectorProcessingOrder, Array<IndType>& sectorCenters,
      Array<DType2>& sensData, Array<DType>& deapoData, const IndType& kernelWidth,
      const IndType& sectorWidth, const DType& osf, Dimensions& imgDims);

  void setUseTextures(bool useTextures) { this->useTextures = useTextures; }
  void setBalanceWorkload(bool balanceWorkload) { this->balanceWorkload = balanceWorkload; }

 protected:
  /* operator flavour from the factory flags
   * (CUDA/src/gpuNUFFT_operator_factory.cpp:324-357) */
  GpuNUFFTOperator* createNewGpuNUFFTOperator(IndType kernelWidth, IndType sectorWidth, DType osf,
                                              Dimensions imgDims);

  /* throws std::runtime_error when the estimated device memory exceeds the
   * device's (CUDA/src/gpuNUFFT_operator_factory.cpp:628-663) */
  void checkMemoryConsumption(Dimensions& kSpaceDims, const IndType& sectorWidth, const DType& osf,
                              Dimensions& imgDims, Dimensions& densDims, Dimensions& sensDims);

  /* sectors per dimension of a grid (CUDA/src/gpuNUFFT_operator_factory.cpp:23-34) */
  static Dimensions computeSectorCountPerDimension(Dimensions dim, IndType sectorWidth);

 private:
  bool useTextures;
  bool useGpu;
  bool balanceWorkload;
  bool matlabSharedMem;
};

}  // namespace gpuNUFFT

#endif /* GPUNUFFT_AMD_OPERATOR_FACTORY_HPP */
