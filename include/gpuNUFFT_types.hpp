/*
 * gpuNUFFT_types.hpp — data model of the gpuNUFFT C++ API
 * (reference CUDA/inc/gpuNUFFT_types.hpp:14-214).
 *
 * Dimensions / Array / GpuArray keep the reference's field names and
 * semantics: an Array is a non-owning (pointer, Dimensions) view; count()
 * multiplies every extent, treating 0 as 1; GpuArray marks a pointer to
 * device memory of the current device.
 */
#ifndef GPUNUFFT_AMD_TYPES_HPP
#define GPUNUFFT_AMD_TYPES_HPP

#include <cstddef>
#include <utility>

#include "config.hpp"

/* extent used by count() for an unset (zero) dimension */
#define DEFAULT_VALUE(a) ((a == 0) ? 1 : a)

/* samples per work item of the balanced sector order
 * (CUDA/inc/gpuNUFFT_types.hpp:22) */
#define MAXIMUM_PAYLOAD 256

namespace gpuNUFFT {

/* (index, key) pair ordered by key (CUDA/inc/gpuNUFFT_types.hpp:33-49). */
struct IndPair : std::pair<IndType, IndType> {
  IndPair(IndType index, IndType key) : std::pair<IndType, IndType>(index, key) {}
  bool operator<(const IndPair& o) const { return second < o.second; }
  bool operator>(const IndPair& o) const { return second > o.second; }
};

/* Extents of an array (CUDA/inc/gpuNUFFT_types.hpp:51-135). */
struct Dimensions {
  Dimensions() : width(0), height(0), depth(0), channels(1), frames(1), length(0) {}
  Dimensions(IndType w, IndType h, IndType d)
      : width(w), height(h), depth(d), channels(1), frames(1), length(0) {}
  Dimensions(IndType w, IndType h)
      : width(w), height(h), depth(0), channels(1), frames(1), length(0) {}

  IndType width;
  IndType height;
  IndType depth;
  IndType channels;
  IndType frames;
  IndType length; /* 1-D extent */

  IndType count() {
    return DEFAULT_VALUE(length) * DEFAULT_VALUE(width) * DEFAULT_VALUE(height) *
           DEFAULT_VALUE(depth) * DEFAULT_VALUE(channels) * DEFAULT_VALUE(frames);
  }

  /* spatial extents scaled (truncating), channels/frames reset */
  Dimensions operator*(const DType alpha) {
    Dimensions r;
    r.width = (IndType)(width * alpha);
    r.height = (IndType)(height * alpha);
    r.depth = (IndType)(depth * alpha);
    r.length = (IndType)(length * alpha);
    return r;
  }

  /* non-zero spatial extents grown by alpha, channels/frames reset */
  Dimensions operator+(const IndType alpha) {
    Dimensions r;
    if (width > 0) r.width = width + alpha;
    if (height > 0) r.height = height + alpha;
    if (depth > 0) r.depth = depth + alpha;
    if (length > 0) r.length = length + alpha;
    return r;
  }
};

/* Non-owning host view (CUDA/inc/gpuNUFFT_types.hpp:137-153). */
template <typename T>
struct Array {
  Array() : data(NULL) {}
  T* data;
  Dimensions dim;
  IndType count() { return dim.count(); }
};

/* Non-owning view of device memory (CUDA/inc/gpuNUFFT_types.hpp:155-162). */
template <typename T>
struct GpuArray : Array<T> {};

/* Kernel lookup (CUDA/inc/gpuNUFFT_types.hpp:164-177). CONST_LOOKUP uses the
 * nearest entry of the 1-D table; TEXTURE2D_LOOKUP the emulated bilinear
 * texture fetch. TEXTURE_LOOKUP / TEXTURE3D_LOOKUP are accepted and run as
 * TEXTURE2D_LOOKUP (identical weights up to the fetch rounding). */
enum InterpolationType { CONST_LOOKUP, TEXTURE_LOOKUP, TEXTURE2D_LOOKUP, TEXTURE3D_LOOKUP };

/* Stage after which an operation stops (CUDA/inc/gpuNUFFT_types.hpp:179-193). */
enum GpuNUFFTOutput { CONVOLUTION, FFT, DEAPODIZATION, DENSITY_ESTIMATION };

/* fftshift direction (CUDA/inc/gpuNUFFT_types.hpp:195-200). */
enum FFTShiftDir { FORWARD, INVERSE };

/* Operator flavour (CUDA/inc/gpuNUFFT_types.hpp:202-214). */
enum OperatorType { DEFAULT, TEXTURE, BALANCED, BALANCED_TEXTURE };

}  // namespace gpuNUFFT

#endif /* GPUNUFFT_AMD_TYPES_HPP */
