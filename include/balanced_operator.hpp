/*
 * balanced_operator.hpp — interface of the load-balanced operators
 * (reference CUDA/inc/balanced_operator.hpp:13-20).
 *
 * The sector processing order lists (sector, first sample offset) work
 * items: sectors by sample count, descending, each split every
 * MAXIMUM_PAYLOAD samples (CUDA/src/gpuNUFFT_operator_factory.cpp:74-126).
 * The HIP kernels balance work with their own chunk list (sector runs of at
 * most 1024 samples, spread over the XCDs by cost); the order is kept for
 * the interface and for the precomputed-operator format.
 */
#ifndef GPUNUFFT_AMD_BALANCED_OPERATOR_HPP
#define GPUNUFFT_AMD_BALANCED_OPERATOR_HPP

#include "gpuNUFFT_types.hpp"

namespace gpuNUFFT {

class BalancedOperator {
 public:
  virtual ~BalancedOperator() {}
  virtual Array<IndType2> getSectorProcessingOrder() = 0;
  virtual void setSectorProcessingOrder(Array<IndType2> sectorProcessingOrder) = 0;
};

/* The balanced order of a sector data count table (S + 1 entries):
 * malloc-ed array the caller frees. */
Array<IndType2> computeSectorProcessingOrder(Array<IndType> sectorDataCount);

}  // namespace gpuNUFFT

#endif /* GPUNUFFT_AMD_BALANCED_OPERATOR_HPP */
