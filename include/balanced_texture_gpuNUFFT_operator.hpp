/*
 * balanced_texture_gpuNUFFT_operator.hpp — load-balanced operator, linear
 * (texture emulating) lookup (reference
 * CUDA/inc/balanced_texture_gpuNUFFT_operator.hpp:24-85).
 */
#ifndef GPUNUFFT_AMD_BALANCED_TEXTURE_OPERATOR_HPP
#define GPUNUFFT_AMD_BALANCED_TEXTURE_OPERATOR_HPP

#include "balanced_operator.hpp"
#include "texture_gpuNUFFT_operator.hpp"

namespace gpuNUFFT {

class BalancedTextureGpuNUFFTOperator : public TextureGpuNUFFTOperator, public BalancedOperator {
 public:
  BalancedTextureGpuNUFFTOperator(IndType kernelWidth, IndType sectorWidth, DType osf,
                                  Dimensions imgDims,
                                  InterpolationType interpolationType = TEXTURE2D_LOOKUP,
                                  bool matlabSharedMem = false)
      : TextureGpuNUFFTOperator(kernelWidth, sectorWidth, osf, imgDims, interpolationType,
                                matlabSharedMem),
        ownOrder(false) {}
  ~BalancedTextureGpuNUFFTOperator();

  Array<IndType2> getSectorProcessingOrder();
  void setSectorProcessingOrder(Array<IndType2> sectorProcessingOrder);

  OperatorType getType() { return BALANCED_TEXTURE; }

 protected:
  Array<IndType2> sectorProcessingOrder;
  bool ownOrder;
};

}  // namespace gpuNUFFT

#endif /* GPUNUFFT_AMD_BALANCED_TEXTURE_OPERATOR_HPP */
