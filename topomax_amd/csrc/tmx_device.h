// Device-side helpers shared by the topomax_amd HIP kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tmx_kernels.h"
#include "tmx_multi.h"

namespace tmx {

constexpr int kWave = 64;
constexpr int kStripCols = kWave - 2;  // output columns per wavefront strip
constexpr int kBlock = 256;            // threads per block for sweep / vector kernels

// Index of the calling wavefront among the launch's (kBlock-thread workgroups), read from the
// first lane so the compiler keeps it, and what is derived from it (strip, chunk, row range),
// in scalar registers: row loops bounded by it then branch on scalars instead of carrying a
// divergent exit mask.
__device__ __forceinline__ int wave_index(unsigned block) {
  return __builtin_amdgcn_readfirstlane((int)(block * (kBlock / kWave) + threadIdx.x / kWave));
}

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
// the sweep kernels never use for an output node.  Doubles move as two DPP moves (GFX9
// wavefront shifts wave_shl:1 = 0x130 and wave_shr:1 = 0x138, with the lane's own value kept
// where the source lane is out of the wave); floats use the LDS permute, which measured
// faster for the FP32 sweeps (the DPP form costs hazard wait states there).
__device__ __forceinline__ int dpp_down_i(int v) {
  return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xF, 0xF, false);
}
__device__ __forceinline__ int dpp_up_i(int v) {
  return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ float lane_down(float v) { return __shfl_down(v, 1, kWave); }
__device__ __forceinline__ float lane_up(float v) { return __shfl_up(v, 1, kWave); }
__device__ __forceinline__ double lane_down(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_down_i((int)(b & 0xffffffffll)), hi = dpp_down_i((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane_up(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_up_i((int)(b & 0xffffffffll)), hi = dpp_up_i((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
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

// Triangle factor used by an operator: as is when wmin <= 0 (the operator being solved);
// |w| raised to at least wmin when wmin > 0 (the preconditioner's SPD stand-in for a factor
// that an odd penalty of a negative filtered density made negative).
template <typename T>
__device__ __forceinline__ T pc_weight(T w, T wmin) {
  return wmin > T(0) ? max(w < T(0) ? -w : w, wmin) : w;
}

// 1/x for the smoothers' 2x2 block inverses: the hardware reciprocal (1 ulp) in FP32, where
// the V-cycle is a preconditioner; correctly rounded division in FP64
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }

// ---------------------------------------------------------------- reductions
// Each block writes its partial sum with a plain store; a separate one-block kernel adds the
// partials in block order (deterministic).  Kernel boundaries order the two, so no fences
// or atomics are needed.
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  return v;
}

// block bid of the launch writes its partials to out[k][bid] (workgroups of at most NT threads)
template <int NRED, int NT = kBlock>
__device__ __forceinline__ void block_partials(const double (&v)[NRED], double* const (&out)[NRED],
                                               unsigned bid) {
  __shared__ double sh[NRED][NT / kWave];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
#pragma unroll
  for (int k = 0; k < NRED; ++k) {
    const double s = wave_sum(v[k]);
    if (lane == 0) sh[k][w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NRED; ++k) {
      double s = 0.0;
      for (int i = 0; i < nw; ++i) s += sh[k][i];
      out[k][bid] = s;
    }
  }
}
template <int NRED>
__device__ __forceinline__ void block_partials(const double (&v)[NRED], double* const (&out)[NRED]) {
  block_partials<NRED>(v, out, blockIdx.x);
}

// fixed-order sum of n partials by one block (all threads receive the result)
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int n) {
  __shared__ double sh[kBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  double s = 0.0;
  // eight loads in flight per thread (the same left-to-right order of additions as one at a
  // time), so a few thousand partials cost one memory latency rather than one per stride
  const int bd = (int)blockDim.x;
  int i = (int)threadIdx.x;
  for (; i + 7 * bd < n; i += 8 * bd) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = part[i + k * bd];
#pragma unroll
    for (int k = 0; k < 8; ++k) s += v[k];
  }
  for (; i < n; i += bd) s += part[i];
  s = wave_sum(s);
  __syncthreads();
  if (lane == 0) sh[w] = s;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < (int)(blockDim.x / kWave); ++i) t += sh[i];
  return t;
}

// ---------------------------------------------------------------- load rules
// Body force / traction rules of the reference (FEM_src/elasisity_problem.py:20-73) at a point,
// with IEEE semantics (no contraction): the body force is value inside the open disc
// |x - c| < r; a traction adds its value on the named side, |coordinate - centre| within
// length/2 + DOLFIN_EPS (df.between).  Vertex and midpoint coordinates as dolfin computes them.
static __device__ __noinline__ void body_value(const LoadDesc& d, double x, double y, double& fx,
                                        double& fy) {
#pragma clang fp contract(off)
  fx = 0.0;
  fy = 0.0;
  if (!d.has_body) return;
  const double dx = x - d.bc[0], dy = y - d.bc[1];
  const double dx2 = __dmul_rn(dx, dx), dy2 = __dmul_rn(dy, dy);
  const double dist = __dsqrt_rn(__dadd_rn(dx2, dy2));
  if (dist < d.br) {
    fx = d.bv[0];
    fy = d.bv[1];
  }
}

static __device__ __noinline__ void traction_value(const LoadDesc& d, double x, double y, double& fx,
                                            double& fy) {
#pragma clang fp contract(off)
  const double eps = 3.0e-16;
  fx = 0.0;
  fy = 0.0;
  for (int i = 0; i < d.ntr; ++i) {
    bool on = false;
    switch (d.tside[i]) {
      case 0: on = (x == 0.0) && (y >= d.tlo[i] - eps) && (y <= d.thi[i] + eps); break;
      case 1: on = (x == d.width) && (y >= d.tlo[i] - eps) && (y <= d.thi[i] + eps); break;
      case 2: on = (y == d.height) && (x >= d.tlo[i] - eps) && (x <= d.thi[i] + eps); break;
      default: on = (y == 0.0) && (x >= d.tlo[i] - eps) && (x <= d.thi[i] + eps); break;
    }
    if (on) {
      fx = __dadd_rn(fx, d.tv[i][0]);
      fy = __dadd_rn(fy, d.tv[i][1]);
    }
  }
}

static __device__ __forceinline__ double vx(const LoadDesc& d, int ix) { return __dadd_rn(0.0, __dmul_rn(d.ab, (double)ix)); }
static __device__ __forceinline__ double vy(const LoadDesc& d, int iy) { return __dadd_rn(0.0, __dmul_rn(d.cd, (double)iy)); }
static __device__ __forceinline__ double mid(double a, double b) { return __dadd_rn(__dmul_rn(0.5, a), __dmul_rn(0.5, b)); }


// ---------------------------------------------------------------- bilinear transfers
// 1-D interpolation from coarse nodes to fine node f (nf fine cells, nc coarse cells).
struct Interp1 {
  int n;
  int c[2];
  double w[2];
};

__device__ __forceinline__ Interp1 interp1(int f, int nf, int nc) {
  Interp1 r;
  if (nf == nc) {  // not coarsened in this direction
    r.n = 1; r.c[0] = f; r.w[0] = 1.0;
  } else if ((f & 1) == 0) {
    r.n = 1; r.c[0] = f >> 1; r.w[0] = 1.0;
  } else if (f == nf) {  // odd fine count: last fine node is the last coarse node
    r.n = 1; r.c[0] = nc; r.w[0] = 1.0;
  } else {
    r.n = 2; r.c[0] = (f - 1) >> 1; r.c[1] = (f + 1) >> 1; r.w[0] = r.w[1] = 0.5;
  }
  return r;
}

__device__ __forceinline__ int cpos(int I, int nf, int nc) { return nf == nc ? I : min(2 * I, nf); }

__device__ __forceinline__ double interp1_weight(int f, int nf, int nc, int I) {
  const Interp1 r = interp1(f, nf, nc);
  return (r.c[0] == I ? r.w[0] : 0.0) + (r.n > 1 && r.c[1] == I ? r.w[1] : 0.0);
}

}  // namespace tmx
