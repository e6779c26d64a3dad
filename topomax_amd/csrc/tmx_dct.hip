// DCT-I / cyclic-reduction solver of the filter's tensor-lumped operator (see tmx_dct.h).
//
// Data layout: vectors are node-major rows, v = iy (nx + 1) + ix, so a y-column is strided and a
// mode row (fixed y-mode, all x) is contiguous.  Seen column-major, a block of rows is an
// (nx + 1) x rows matrix with leading dimension nx + 1, and the y transforms are plain GEMMs
// with the folded DCT-I matrices on the right.
#include "tmx_dct.h"

#include <cmath>
#include <vector>

#include "tmx_device.h"

namespace tmx {

// s[j] = r[j] + r[n - j] (j < h), s[h] = r[h]; d[j] = r[j] - r[n - j] (j < h); rows of mx values
template <typename T>
__global__ void k_dct_fold(int mx, int n, const T* __restrict__ r, T* __restrict__ s,
                           T* __restrict__ d) {
  const int h = n / 2;
  const size_t total = (size_t)(h + 1) * mx;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(t / mx), ix = (int)(t - (size_t)j * mx);
    const T a = r[(size_t)j * mx + ix];
    if (j < h) {
      const T b = r[(size_t)(n - j) * mx + ix];
      s[t] = a + b;
      d[t] = a - b;
    } else {
      s[t] = a;
    }
  }
}

// z[j] = e[j] + o[j], z[n - j] = e[j] - o[j] (j < h), z[h] = e[h]; partials of r.z
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dct_unfold(int mx, int n, const T* __restrict__ e,
                                                       const T* __restrict__ o,
                                                       const T* __restrict__ r,
                                                       T* __restrict__ z,
                                                       double* __restrict__ part) {
  const int h = n / 2;
  const size_t total = (size_t)(h + 1) * mx;
  double acc = 0.0;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(t / mx), ix = (int)(t - (size_t)j * mx);
    const size_t lo = (size_t)j * mx + ix;
    if (j < h) {
      const size_t hi = (size_t)(n - j) * mx + ix;
      const T zl = e[t] + o[t], zh = e[t] - o[t];
      z[lo] = zl;
      z[hi] = zh;
      acc += (double)r[lo] * (double)zl + (double)r[hi] * (double)zh;
    } else {
      const T zl = e[t];
      z[lo] = zl;
      acc += (double)r[lo] * (double)zl;
    }
  }
  if (part) {
    const double v[1] = {acc};
    double* const out[1] = {part};
    block_partials<1>(v, out);
  }
}

// One workgroup per mode row: rows 0..h of `even` are the modes 0, 2, .., n, rows 0..h-1 of
// `odd` the modes 1, 3, .., n-1 (blocks h+1 ..).  Each solves (a T_x + c_m W_x) u = f / D_m in
// place by parallel cyclic reduction in LDS, keeping 1/b beside b so that each equation costs
// one division per step (five arrays of floats: 60 KB of LDS for the longest row).
constexpr int kPcrThreads = 1024;

template <typename T>
constexpr int kDctMaxRow = DctPrecond<T>::kMaxRow;

template <typename T>
__global__ __launch_bounds__(kPcrThreads) void k_dct_pcr(int mx, int n, T a_off, double hxhy,
                                                         double cy, T* __restrict__ even,
                                                         T* __restrict__ odd) {
  constexpr int kRow = kDctMaxRow<T>;
  constexpr int kPcrPer = kRow / kPcrThreads;  // equations per thread
  __shared__ T la[kRow], lb[kRow], lr[kRow], lc[kRow], lf[kRow];  // sub, diag, 1/diag, super, rhs
  const int h = n / 2;
  const int blk = (int)blockIdx.x;
  const int mode = blk <= h ? 2 * blk : 2 * (blk - h - 1) + 1;
  T* row = blk <= h ? even + (size_t)blk * mx : odd + (size_t)(blk - h - 1) * mx;
  const double lam = 2.0 - 2.0 * std::cos(M_PI * (double)mode / (double)n);
  const T c = (T)(hxhy + cy * lam);
  const T inv_d = (mode == 0 || mode == n) ? T(1) / (T)n : T(2) / (T)n;
  const int tid = threadIdx.x;
  for (int i = tid; i < mx; i += kPcrThreads) {
    const bool edge = i == 0 || i == mx - 1;
    la[i] = i > 0 ? -a_off : T(0);
    lc[i] = i < mx - 1 ? -a_off : T(0);
    const T bi = a_off * (edge ? T(1) : T(2)) + c * (edge ? T(0.5) : T(1));
    lb[i] = bi;
    lr[i] = T(1) / bi;
    lf[i] = row[i] * inv_d;
  }
  __syncthreads();
  for (int s = 1; s < mx; s *= 2) {
    T na[kPcrPer], nb[kPcrPer], nc[kPcrPer], nf[kPcrPer];
#pragma unroll
    for (int q = 0; q < kPcrPer; ++q) {
      const int i = tid + q * kPcrThreads;
      if (i >= mx) continue;
      const int im = i - s, ip = i + s;
      T ai = T(0), ci = T(0);
      T bi = lb[i], fi = lf[i];
      if (im >= 0) {
        const T k1 = la[i] * lr[im];
        ai = -la[im] * k1;
        bi -= lc[im] * k1;
        fi -= lf[im] * k1;
      }
      if (ip < mx) {
        const T k2 = lc[i] * lr[ip];
        ci = -lc[ip] * k2;
        bi -= la[ip] * k2;
        fi -= lf[ip] * k2;
      }
      na[q] = ai;
      nb[q] = bi;
      nc[q] = ci;
      nf[q] = fi;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPcrPer; ++q) {
      const int i = tid + q * kPcrThreads;
      if (i >= mx) continue;
      la[i] = na[q];
      lb[i] = nb[q];
      lr[i] = T(1) / nb[q];
      lc[i] = nc[q];
      lf[i] = nf[q];
    }
    __syncthreads();
  }
  for (int i = tid; i < mx; i += kPcrThreads) row[i] = lf[i] * lr[i];
}

template <typename T>
bool DctPrecond<T>::supported(int nx, int ny) {
  return ny >= 2 && !(ny & 1) && nx >= 1 && nx + 1 <= kMaxRow;
}

template <typename T>
DctPrecond<T>::~DctPrecond() {
  if (blas_) (void)rocblas_destroy_handle(blas_);
}

// plain GEMM in the element type: C (m x n) = op(A) op(B), column-major
static rocblas_status gemm(rocblas_handle b, rocblas_operation tb, int m, int n, int k,
                           const float* A, int lda, const float* B, int ldb, float* C, int ldc) {
  const float one = 1.0f, zero = 0.0f;
  return rocblas_sgemm(b, rocblas_operation_none, tb, m, n, k, &one, A, lda, B, ldb, &zero, C, ldc);
}
static rocblas_status gemm(rocblas_handle b, rocblas_operation tb, int m, int n, int k,
                           const double* A, int lda, const double* B, int ldb, double* C,
                           int ldc) {
  const double one = 1.0, zero = 0.0;
  return rocblas_dgemm(b, rocblas_operation_none, tb, m, n, k, &one, A, lda, B, ldb, &zero, C, ldc);
}

static int blas_check(rocblas_status st, const char* what) {
  if (st == rocblas_status_success) return 0;
  set_error(std::string("rocBLAS ") + what + " failed: " + rocblas_status_to_string(st));
  return -1;
}

template <typename T>
int DctPrecond<T>::init(int nx, int ny, double hx, double hy, double eps, hipStream_t s) {
  if (!supported(nx, ny)) {
    set_error("internal: DCT preconditioner needs an even ny and nx + 1 <= " +
              std::to_string(kMaxRow));
    return -1;
  }
  nx_ = nx;
  ny_ = ny;
  h_ = ny / 2;
  mx_ = nx + 1;
  hx_ = hx;
  hy_ = hy;
  eps_ = eps;
  const int h = h_, n = ny;
  std::vector<T> ce((size_t)(h + 1) * (h + 1)), co((size_t)h * h);
  // column-major: ce[j + m (h + 1)] = cos(pi j 2m / n); co[j + m h] = cos(pi j (2m + 1) / n)
  for (int m = 0; m <= h; ++m)
    for (int j = 0; j <= h; ++j)
      ce[(size_t)j + (size_t)m * (h + 1)] =
          (T)std::cos(M_PI * (double)((long)j * 2 * m % (2 * n)) / (double)n);
  for (int m = 0; m < h; ++m)
    for (int j = 0; j < h; ++j)
      co[(size_t)j + (size_t)m * h] =
          (T)std::cos(M_PI * (double)((long)j * (2 * m + 1) % (2 * n)) / (double)n);
  TMX_TRY(ce_.alloc(ce.size()));
  TMX_TRY(co_.alloc(co.size()));
  TMX_TRY(fs_.alloc((size_t)(h + 1) * mx_));
  TMX_TRY(fd_.alloc((size_t)(h + 1) * mx_));
  TMX_TRY(ts_.alloc((size_t)(h + 1) * mx_));
  TMX_TRY(td_.alloc((size_t)(h + 1) * mx_));
  TMX_HIP(hipMemcpyAsync(ce_.get(), ce.data(), ce.size() * sizeof(T), hipMemcpyHostToDevice, s));
  TMX_HIP(hipMemcpyAsync(co_.get(), co.data(), co.size() * sizeof(T), hipMemcpyHostToDevice, s));
  TMX_HIP(hipStreamSynchronize(s));
  if (!blas_) TMX_TRY(blas_check(rocblas_create_handle(&blas_), "create_handle"));
  return 0;
}

template <typename T>
int DctPrecond<T>::apply(const T* r, T* z, RedSlot* rz, hipStream_t s) {
  const int h = h_, n = ny_, mx = mx_;
  const size_t total = (size_t)(h + 1) * mx;
  const unsigned g = (unsigned)std::min<size_t>(4096, (total + kBlock - 1) / kBlock);
  k_dct_fold<T><<<g, kBlock, 0, s>>>(mx, n, r, fs_.get(), fd_.get());
  TMX_HIP(hipGetLastError());
  TMX_TRY(blas_check(rocblas_set_stream(blas_, s), "set_stream"));
  // forward: T_e (mx x h+1) = S (mx x h+1) Ce; T_o (mx x h) = D (mx x h) Co
  TMX_TRY(blas_check(gemm(blas_, rocblas_operation_none, mx, h + 1, h + 1, fs_.get(), mx,
                          ce_.get(), h + 1, ts_.get(), mx), "gemm"));
  TMX_TRY(blas_check(gemm(blas_, rocblas_operation_none, mx, h, h, fd_.get(), mx, co_.get(), h,
                          td_.get(), mx), "gemm"));
  // tridiagonal solves along x, one workgroup per mode
  const T a_off = (T)(eps_ * eps_ * hy_ / hx_);
  const double cy = eps_ * eps_ * hx_ / hy_;
  k_dct_pcr<T><<<2 * h + 1, kPcrThreads, 0, s>>>(mx, n, a_off, hx_ * hy_, cy, ts_.get(),
                                                 td_.get());
  TMX_HIP(hipGetLastError());
  // inverse: E (mx x h+1) = T_e Ce^T; O (mx x h) = T_o Co^T
  TMX_TRY(blas_check(gemm(blas_, rocblas_operation_transpose, mx, h + 1, h + 1, ts_.get(), mx,
                          ce_.get(), h + 1, fs_.get(), mx), "gemm"));
  TMX_TRY(blas_check(gemm(blas_, rocblas_operation_transpose, mx, h, h, td_.get(), mx, co_.get(),
                          h, fd_.get(), mx), "gemm"));
  const unsigned gu = (unsigned)std::min<size_t>(2048, (total + kBlock - 1) / kBlock);
  double* part = nullptr;
  if (rz) {
    rz->nblocks = (int)gu;
    part = rz->partials;
  }
  k_dct_unfold<T><<<gu, kBlock, 0, s>>>(mx, n, fs_.get(), fd_.get(), r, z, part);
  TMX_HIP(hipGetLastError());
  return 0;
}

template class DctPrecond<float>;
template class DctPrecond<double>;

}  // namespace tmx
