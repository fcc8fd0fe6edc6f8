/*
 * gpuNUFFT_operator.hpp — the gpuNUFFT operator of the C++ API
 * (reference CUDA/inc/gpuNUFFT_operator.hpp:40-590).
 *
 * Same public interface as the reference: constructor, setters/getters of
 * the precomputed tables, and the five shapes each of performGpuNUFFTAdj /
 * performForwardGpuNUFFT (host Arrays, device GpuArrays, stop stage). Work
 * runs on the current HIP device through the C ABI of include/gpunufft_hip.h
 * (one gnufft_plan per operator: every table lives in device memory; no
 * process-global state, so operators may be used from different threads).
 *
 * Errors throw std::invalid_argument (bad arguments) or std::runtime_error
 * (device, FFT, memory), where the reference exits the process on device
 * errors (CUDA/inc/cuda_utils.hpp:22-30).
 */
#ifndef GPUNUFFT_AMD_OPERATOR_HPP
#define GPUNUFFT_AMD_OPERATOR_HPP

#include <string>

#include "gpuNUFFT_types.hpp"

struct gnufft_plan_s;

namespace gpuNUFFT {

class GpuNUFFTOperatorFactory;

class GpuNUFFTOperator {
 public:
  /* kernelWidth, sectorWidth in grid units; osf oversampling factor;
   * imgDims image size (depth 0 = 2-D). loadKernel is kept for interface
   * compatibility: the lookup table is always built with the plan.
   * operatorType TEXTURE / BALANCED_TEXTURE selects the linear lookup. */
  GpuNUFFTOperator(IndType kernelWidth, IndType sectorWidth, DType osf, Dimensions imgDims,
                   bool loadKernel = true, OperatorType operatorType = DEFAULT,
                   bool matlabSharedMem = false);
  virtual ~GpuNUFFTOperator();

  GpuNUFFTOperator(const GpuNUFFTOperator&) = delete;
  GpuNUFFTOperator& operator=(const GpuNUFFTOperator&) = delete;

  friend class GpuNUFFTOperatorFactory;

  /* ---- setters (CUDA/inc/gpuNUFFT_operator.hpp:92-134). Arrays are
   * borrowed: the operator reads them at the next perform call. */
  void setOsf(DType osf);
  void setKSpaceTraj(Array<DType> kSpaceTraj);
  void setSectorCenters(Array<IndType> sectorCenters);
  void setSectorDataCount(Array<IndType> sectorDataCount);
  void setDataIndices(Array<IndType> dataIndices);
  void setSens(Array<DType2> sens);
  void setDens(Array<DType> dens);
  void setDeapodizationFunction(Array<DType> deapo);
  void setImageDims(Dimensions dims);
  void setGridSectorDims(Dimensions dims);

  /* ---- getters (CUDA/inc/gpuNUFFT_operator.hpp:136-216). Tables built on
   * the device are copied to host memory owned by the operator on first
   * use. */
  DType getOsf() { return osf; }
  Array<DType> getKSpaceTraj();
  Array<DType2> getSens() { return sens; }
  Array<DType> getDens();
  Array<DType> getKernel();
  Array<IndType> getSectorDataCount();
  IndType getKernelWidth() { return kernelWidth; }
  IndType getSectorWidth() { return sectorWidth; }
  Dimensions getImageDims() { return imgDims; }
  Dimensions getGridDims();
  Dimensions getGridSectorDims();
  Dimensions getSectorDims() { return sectorDims; }
  Array<IndType> getSectorCenters();
  IndType* getSectorCentersData() { return getSectorCenters().data; }
  Array<IndType> getDataIndices();
  Array<DType> getDeapodizationFunction();
  bool is2DProcessing() { return imgDims.depth == 0; }
  bool is3DProcessing() { return !is2DProcessing(); }
  int getImageDimensionCount() { return is2DProcessing() ? 2 : 3; }

  /* ---- adjoint (CUDA/inc/gpuNUFFT_operator.hpp:218-287) ------------------
   * kspaceData [channels][M] complex in the trajectory's original order.
   * Output: CONVOLUTION / DENSITY_ESTIMATION the oversampled grid
   * [channels][grid]; FFT the cropped, scaled image of the first coil;
   * DEAPODIZATION the image, coil-combined with conj(sens) when
   * sensitivities are set, else [channels][img]. The returning shapes
   * allocate with calloc (caller frees). */
  Array<CufftType> performGpuNUFFTAdj(Array<DType2> kspaceData);
  virtual void performGpuNUFFTAdj(Array<DType2> kspaceData, Array<CufftType>& imgData,
                                  GpuNUFFTOutput gpuNUFFTOut = DEAPODIZATION);
  virtual void performGpuNUFFTAdj(GpuArray<DType2> kspaceData_gpu,
                                  GpuArray<CufftType>& imgData_gpu,
                                  GpuNUFFTOutput gpuNUFFTOut = DEAPODIZATION);
  Array<CufftType> performGpuNUFFTAdj(Array<DType2> kspaceData, GpuNUFFTOutput gpuNUFFTOut);

  /* ---- forward (CUDA/inc/gpuNUFFT_operator.hpp:289-346) ------------------
   * imgData [channels][img] (one image replicated over the sensitivity
   * coils when they are set); output [coils][M] in the original order.
   * DENSITY_ESTIMATION interpolates from the deapodized image used as the
   * grid (defined for grid == image size). */
  Array<CufftType> performForwardGpuNUFFT(Array<DType2> imgData);
  virtual void performForwardGpuNUFFT(Array<DType2> imgData, Array<CufftType>& kspaceData,
                                      GpuNUFFTOutput gpuNUFFTOut = DEAPODIZATION);
  virtual void performForwardGpuNUFFT(GpuArray<DType2> imgData_gpu,
                                      GpuArray<CufftType>& kspaceData_gpu,
                                      GpuNUFFTOutput gpuNUFFTOut = DEAPODIZATION);
  Array<CufftType> performForwardGpuNUFFT(Array<DType2> imgData, GpuNUFFTOutput gpuNUFFTOut);

  bool applyDensComp() { return dens.data != NULL && dens.count() > 1; }
  bool applySensData() { return sens.data != NULL && sens.count() > 1; }

  virtual OperatorType getType() { return operatorType; }

  /* C ABI handle of this operator (include/gpunufft_hip.h); valid for the
   * operator's lifetime. Not part of the reference interface. */
  gnufft_plan_s* plan();

 protected:
  OperatorType operatorType;
  DType osf;
  IndType kernelWidth;
  IndType sectorWidth;
  Dimensions imgDims;
  Dimensions sectorDims;
  Dimensions gridSectorDims;
  bool matlabSharedMem;

  Array<DType> kernel;
  Array<DType> kSpaceTraj;
  Array<DType> dens;
  Array<DType> deapo;
  Array<DType2> sens;
  Array<IndType> sectorCenters;
  Array<IndType> dataIndices;
  Array<IndType> sectorDataCount;

  /* builds the host copy of the lookup table (2-D product table for the
   * texture flavours, CUDA/src/texture_gpuNUFFT_operator.cpp:3-27) */
  virtual void initKernel();

  /* number of samples of the trajectory */
  IndType sampleCount();

  /* ensure the device plan holds the tables / sensitivities last set */
  void syncDevice();

  /* copy the device-built tables into operator-owned host arrays */
  void fetchTables();

  /* factory hook: run the device precompute on a host trajectory */
  void precompute(Array<DType>& kSpaceTraj, Array<DType>& densCompData);

  /* host arrays allocated by the operator itself */
  enum Owned {
    OWN_KERNEL = 1, OWN_TRAJ = 2, OWN_DENS = 4, OWN_DEAPO = 8, OWN_CENTERS = 16,
    OWN_INDICES = 32, OWN_COUNT = 64
  };
  unsigned owned;
  void releaseOwned(unsigned which);

 private:
  gnufft_plan_s* plan_;
  bool tablesOnDevice;  /* plan holds the current tables */
  bool tablesOnHost;    /* host arrays mirror the device tables */
  bool sensDirty;
  IndType nSamples;

  void adjoint(const void* kdata, void* out, IndType nCoils, GpuNUFFTOutput stop, bool host,
               size_t outCount);
  void forward(const void* img, void* out, IndType nCoils, GpuNUFFTOutput stop, bool host,
               IndType imgChannels);
};

/* status -> exception (std::invalid_argument / std::runtime_error) */
void checkStatus(int status, const std::string& where);

}  // namespace gpuNUFFT

#endif /* GPUNUFFT_AMD_OPERATOR_HPP */
