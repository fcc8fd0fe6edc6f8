/*
 * texture_gpuNUFFT_operator.hpp — operator with the linear (texture
 * emulating) kernel lookup (reference CUDA/inc/texture_gpuNUFFT_operator.hpp:19-70).
 *
 * The reference reads the kernel through a CUDA texture with hardware
 * linear filtering; here the fetch is emulated in the gridding kernels
 * (clamped addressing, texel centre at i + 0.5, 8-bit fractional weight).
 */
#ifndef GPUNUFFT_AMD_TEXTURE_OPERATOR_HPP
#define GPUNUFFT_AMD_TEXTURE_OPERATOR_HPP

#include "gpuNUFFT_operator.hpp"

namespace gpuNUFFT {

class TextureGpuNUFFTOperator : public GpuNUFFTOperator {
 public:
  TextureGpuNUFFTOperator(IndType kernelWidth, IndType sectorWidth, DType osf, Dimensions imgDims,
                          InterpolationType interpolationType = TEXTURE2D_LOOKUP,
                          bool matlabSharedMem = false)
      : GpuNUFFTOperator(kernelWidth, sectorWidth, osf, imgDims, false, TEXTURE, matlabSharedMem),
        interpolationType(interpolationType) {
    initKernel();
  }

  virtual OperatorType getType() { return TEXTURE; }

 protected:
  InterpolationType interpolationType;
  void initKernel();
};

}  // namespace gpuNUFFT

#endif /* GPUNUFFT_AMD_TEXTURE_OPERATOR_HPP */
