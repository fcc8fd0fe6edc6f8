/*
 * balanced_gpuNUFFT_operator.hpp — load-balanced operator, nearest lookup
 * (reference CUDA/inc/balanced_gpuNUFFT_operator.hpp:20-80).
 */
#ifndef GPUNUFFT_AMD_BALANCED_GPUNUFFT_OPERATOR_HPP
#define GPUNUFFT_AMD_BALANCED_GPUNUFFT_OPERATOR_HPP

#include "balanced_operator.hpp"
#include "gpuNUFFT_operator.hpp"

namespace gpuNUFFT {

class BalancedGpuNUFFTOperator : public GpuNUFFTOperator, public BalancedOperator {
 public:
  BalancedGpuNUFFTOperator(IndType kernelWidth, IndType sectorWidth, DType osf, Dimensions imgDims,
                           bool matlabSharedMem = false)
      : GpuNUFFTOperator(kernelWidth, sectorWidth, osf, imgDims, true, BALANCED, matlabSharedMem),
        ownOrder(false) {}
  ~BalancedGpuNUFFTOperator();

  /* computed from the sector data count on first use unless set */
  Array<IndType2> getSectorProcessingOrder();
  void setSectorProcessingOrder(Array<IndType2> sectorProcessingOrder);

  OperatorType getType() { return BALANCED; }

 protected:
  Array<IndType2> sectorProcessingOrder;
  bool ownOrder;
};

}  // namespace gpuNUFFT

#endif /* GPUNUFFT_AMD_BALANCED_GPUNUFFT_OPERATOR_HPP */
