// scx_device.h -- device helpers (wave64 reductions/scans, hashing, packed cell keys).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace scx {

__device__ __forceinline__ uint32_t wave_inclusive_scan_u32(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_reduce_sum_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum over blockDim.x == 256 threads; result valid in every thread.
__device__ __forceinline__ uint32_t block_reduce_sum_u32(uint32_t v) {
  __shared__ uint32_t s[4];
  v = wave_reduce_sum_u32(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  return s[0] + s[1] + s[2] + s[3];
}

// Block-wide exclusive scan over 256 threads; *total = block sum.
__device__ __forceinline__ uint32_t block_exclusive_scan_u32(uint32_t v, uint32_t* total) {
  __shared__ uint32_t s[4];
  const uint32_t incl = wave_inclusive_scan_u32(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 63) s[threadIdx.x >> 6] = incl;
  __syncthreads();
  const int w = threadIdx.x >> 6;
  uint32_t before = 0;
  for (int k = 0; k < w; ++k) before += s[k];
  *total = s[0] + s[1] + s[2] + s[3];
  return before + incl - v;
}

__device__ __forceinline__ double wave_reduce_max_f64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ double wave_reduce_min_f64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ double wave_reduce_sum_f64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Cell coordinates packed 20 bits per axis (60 bits); bits 60..62 carry, for a grid
// entry, whether the cell is the first cell of its box along x / y / z.
constexpr int kKeyBits = 20;
constexpr uint32_t kKeyMax = (1u << kKeyBits) - 1;
constexpr unsigned long long kKeyMask = (1ull << 60) - 1;

__device__ __forceinline__ unsigned long long pack_cell(uint32_t x, uint32_t y, uint32_t z) {
  return (unsigned long long)x | ((unsigned long long)y << 20) | ((unsigned long long)z << 40);
}

__device__ __forceinline__ uint32_t hash_cell(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return (uint32_t)k;
}

// Contact priority for the colouring rounds: (fmix32(c * golden + salt) << 32) | c.
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
__device__ __forceinline__ unsigned long long prio_key(uint32_t c) {
  return ((unsigned long long)fmix32(c * 0x9E3779B9u + 0x632BE5ABu) << 32) | c;
}

// Cell index of a coordinate: floor((x - min) / h) as the reference's unsigned cast
// (ball2d/SpatialGridDetector.cpp:29-34), clamped to the packed key range.
__device__ __forceinline__ uint32_t cell_of(double x, double mn, double h) {
  const double t = (x - mn) / h;
  if (!(t > 0.0)) return 0u;
  if (!(t < (double)kKeyMax)) return kKeyMax;
  return (uint32_t)t;
}

}  // namespace scx
