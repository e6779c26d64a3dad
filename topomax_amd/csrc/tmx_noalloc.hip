// tmx_integrate_noalloc: the integral of a CG1 field without a problem handle and without any
// device allocation (include/tmx.h).
//
// tmx_integrate reduces m . v with the handle's lumped-mass vector and the handle's workspace
// for the block partials; both come from tmx_problem_create, which allocates.  Here the lumped
// mass m_i = (hx hy / 6) (triangles at vertex i) of the 'right' mesh is formed from (ix, iy)
// inside the kernel, and the block partials live in `scratch`, a file-scope device array that
// every GPU holds once for the life of the process.  Kernels receive its address as
// hipGetSymbolAddress returns it for the current device; the array's name is never passed to
// a kernel or a copy from host code.
#include "../../include/tmx.h"

#include <mutex>
#include <string>

#include "tmx_device.h"
#include "tmx_solver.h"

namespace tmx {

// vertices per thread and blocks of one launch: scratch holds one partial sum per block, so
// one call reduces at most kNoAllocBlocks * kBlock * kNoAllocPerThread = 8,388,608 vertices
// (the N = 1000 cantilever has 3,004,001) with 32 KiB of device memory per GPU
constexpr int kNoAllocPerThread = 8;
constexpr int kNoAllocBlocks = 4096;
constexpr long long kNoAllocMaxVertices = (long long)kNoAllocBlocks * kBlock * kNoAllocPerThread;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocBlocks];

// partial sums of (hx hy / 6) * (triangles at the vertex) * v over kBlock * kNoAllocPerThread
// consecutive vertices per block.  A vertex (ix, iy) is v0 of cell (ix, iy) (two triangles),
// v1 of cell (ix-1, iy) (triangle A), v2 of cell (ix, iy-1) (triangle B) and v3 of cell
// (ix-1, iy-1) (two triangles)
__global__ __launch_bounds__(kBlock) void k_integrate_partials(int nx, int ny, double w6,
                                                               const double* __restrict__ v,
                                                               double* __restrict__ part) {
  const unsigned nvx = (unsigned)nx + 1u;
  const unsigned nv = nvx * ((unsigned)ny + 1u);
  const unsigned base = blockIdx.x * (unsigned)(kBlock * kNoAllocPerThread) + threadIdx.x;
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < kNoAllocPerThread; ++k) {
    const unsigned i = base + (unsigned)(k * kBlock);
    if (i < nv) {
      const int iy = (int)(i / nvx), ix = (int)(i - (unsigned)iy * nvx);
      const bool cx0 = ix < nx, cx1 = ix > 0, cy0 = iy < ny, cy1 = iy > 0;
      const int tris = 2 * (cx0 && cy0) + (cx1 && cy0) + (cx0 && cy1) + 2 * (cx1 && cy1);
      acc += (w6 * (double)tris) * v[i];
    }
  }
  const double r[1] = {acc};
  double* const out[1] = {part};
  block_partials<1>(r, out);
}

__global__ __launch_bounds__(kBlock) void k_integrate_total(const double* __restrict__ part,
                                                            int n, double* __restrict__ out) {
  const double s = sum_partials(part, n);
  if (threadIdx.x == 0) *out = s;
}

// address of the current device's copy of `scratch`, looked up once per device
static int scratch_address(double** out) {
  static std::mutex mu;
  static double* addr[kNoAllocMaxDevices] = {};
  int dev = -1;
  TMX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) {
    set_error("tmx_integrate_noalloc: device ordinal " + std::to_string(dev) + " out of range");
    return TMX_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("tmx_integrate_noalloc: hipGetSymbolAddress of the static partial "
                            "buffer failed: ") + hipGetErrorString(e));
      return TMX_ERR_HIP;
    }
    addr[dev] = static_cast<double*>(p);
  }
  *out = addr[dev];
  return TMX_OK;
}

}  // namespace tmx

using namespace tmx;

extern "C" int tmx_integrate_noalloc(int nx, int ny, double width, double height,
                                     const double* values, double* out, void* stream) {
  if (!values || !out) {
    set_error("tmx_integrate_noalloc: null argument");
    return TMX_ERR_INVALID;
  }
  if (nx < 1 || ny < 1 || !(width > 0.0) || !(height > 0.0)) {
    set_error("tmx_integrate_noalloc: the mesh needs nx, ny >= 1 and a positive width and height");
    return TMX_ERR_INVALID;
  }
  const long long nv = ((long long)nx + 1) * ((long long)ny + 1);
  if (nv > kNoAllocMaxVertices) {
    set_error("tmx_integrate_noalloc: " + std::to_string(nv) + " vertices exceed the maximum of " +
              std::to_string(kNoAllocMaxVertices) + " (the static buffer holds " +
              std::to_string(kNoAllocBlocks) + " block partial sums); use tmx_integrate");
    return TMX_ERR_INVALID;
  }
  double* part = nullptr;
  TMX_TRY(scratch_address(&part));
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = (int)((nv + kBlock * kNoAllocPerThread - 1) / (kBlock * kNoAllocPerThread));
  const double hx = width / nx, hy = height / ny;
  k_integrate_partials<<<nb, kBlock, 0, s>>>(nx, ny, hx * hy / 6.0, values, part);
  TMX_HIP(hipGetLastError());
  k_integrate_total<<<1, kBlock, 0, s>>>(part, nb, out);
  TMX_HIP(hipGetLastError());
  return TMX_OK;
}
