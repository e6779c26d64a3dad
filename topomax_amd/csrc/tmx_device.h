// Device-side helpers shared by the topomax_amd HIP kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tmx_kernels.h"

namespace tmx {

constexpr int kWave = 64;
constexpr int kStripCols = kWave - 2;  // output columns per wavefront strip
constexpr int kBlock = 256;            // threads per block for sweep / vector kernels

template <typename T>
struct alignas(2 * sizeof(T)) V2 {
  T x, y;
};

template <typename T>
__device__ __forceinline__ V2<T> operator+(V2<T> a, V2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ V2<T> operator-(V2<T> a, V2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ V2<T>& operator+=(V2<T>& a, V2<T> b) {
  a.x += b.x;
  a.y += b.y;
  return a;
}

// Cross-lane moves inside one wavefront.  down(v): lane i receives lane i+1's value;
// up(v): lane i receives lane i-1's value.  Edge lanes receive their own value, which
// the sweep kernels never use for an output node.
template <typename T>
__device__ __forceinline__ T lane_down(T v) { return __shfl_down(v, 1, kWave); }
template <typename T>
__device__ __forceinline__ T lane_up(T v) { return __shfl_up(v, 1, kWave); }
template <typename T>
__device__ __forceinline__ V2<T> lane_down(V2<T> v) { return {lane_down(v.x), lane_down(v.y)}; }
template <typename T>
__device__ __forceinline__ V2<T> lane_up(V2<T> v) { return {lane_up(v.x), lane_up(v.y)}; }

__device__ __forceinline__ bool node_fixed(unsigned fixed, int ix, int iy, int nx, int ny) {
  return ((fixed & kLeft) && ix == 0) || ((fixed & kRight) && ix == nx) ||
         ((fixed & kBottom) && iy == 0) || ((fixed & kTop) && iy == ny);
}

// Complete homogeneous symmetric polynomial h_p(a, b, c) by the recurrences
// h_k(a,b) = a^k + b h_{k-1}(a,b) and h_k(a,b,c) = h_k(a,b) + c h_{k-1}(a,b,c).
template <typename T>
__device__ __forceinline__ T hsym3(T a, T b, T c, int p) {
  T pa = T(1), hab = T(1), h = T(1);
  for (int k = 1; k <= p; ++k) {
    pa *= a;
    hab = pa + b * hab;
    h = hab + c * h;
  }
  return h;
}

// ---------------------------------------------------------------- reductions
// Sum a double over the block, write the block's partial, and let the last block to
// finish add all partials in block order (deterministic) into *result.
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  return v;
}

struct Reduction {
  double* partials;   // >= gridDim.x entries
  unsigned* counter;  // zero before the launch; the last block resets it
  double* result;     // where the final sum goes
};

__device__ __forceinline__ double load_coherent(const double* p) {
  unsigned long long bits =
      __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED,
                        __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double((long long)bits);
}

template <int NRED>
__device__ void block_reduce_finalize(const double (&v)[NRED], const Reduction (&red)[NRED]) {
  __shared__ double sh[NRED][kBlock / kWave];
  __shared__ bool last;
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
#pragma unroll
  for (int k = 0; k < NRED; ++k) {
    double s = wave_sum(v[k]);
    if (lane == 0) sh[k][w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NRED; ++k) {
      double s = 0.0;
      for (int i = 0; i < nw; ++i) s += sh[k][i];
      __hip_atomic_store(reinterpret_cast<unsigned long long*>(&red[k].partials[blockIdx.x]),
                         (unsigned long long)__double_as_longlong(s), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    __threadfence();
    unsigned prev = atomicAdd(red[0].counter, 1u);
    last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
#pragma unroll
  for (int k = 0; k < NRED; ++k) {
    double s = 0.0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += blockDim.x) s += load_coherent(&red[k].partials[i]);
    // fixed-order tree over the block
    s = wave_sum(s);
    __syncthreads();
    if (lane == 0) sh[k][w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int i = 0; i < nw; ++i) t += sh[k][i];
      *red[k].result = t;
    }
  }
  if (threadIdx.x == 0) *red[0].counter = 0u;
}

}  // namespace tmx
