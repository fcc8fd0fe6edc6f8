/*
 * config.hpp — scalar and vector types of the gpuNUFFT C++ API.
 *
 * Mirrors the single-precision configuration of the reference
 * (CUDA/inc/config.hpp.cmake:17-36): DType float, complex samples as
 * {x = re, y = im} pairs, unsigned int indices. The vector types are plain
 * host structs here (no CUDA/HIP vector types), so the C++ API compiles with
 * any host compiler; their layout is that of float2/uint2/..., so device
 * buffers can be reinterpreted either way.
 */
#ifndef GPUNUFFT_AMD_CONFIG_HPP
#define GPUNUFFT_AMD_CONFIG_HPP

struct gnufft_float2 {
  float x, y;
};
struct gnufft_float3 {
  float x, y, z;
};
struct gnufft_uint2 {
  unsigned int x, y;
};
struct gnufft_uint3 {
  unsigned int x, y, z;
};

typedef float DType;
typedef gnufft_float2 DType2;
typedef gnufft_float3 DType3;
typedef gnufft_float2 CufftType; /* interleaved complex64, layout of cufftComplex */

typedef unsigned int SizeType;
typedef unsigned int IndType;
typedef gnufft_uint2 IndType2;
typedef gnufft_uint3 IndType3;

#endif /* GPUNUFFT_AMD_CONFIG_HPP */
