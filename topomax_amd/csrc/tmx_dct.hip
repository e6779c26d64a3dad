// DCT-I / tridiagonal solver of the filter's tensor-lumped operator (see tmx_dct.h).
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

// s[j] = r[j] + r[n - j] (j < h), s[h] = r[h]; d[j] = r[j] - r[n - j] (j < h), d[h] = 0; rows
// of mx values.  Row h of d is the odd half's zero padding: it is rewritten on every call, so a
// non-finite value that an earlier call left in the buffer cannot reach this one's transforms
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
      d[t] = T(0);
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

// One tridiagonal solve along x per mode row: rows 0..h of `even` are the modes 0, 2, .., n,
// rows 0..h-1 of `odd` the modes 1, 3, .., n-1 (workgroups h+1 ..).  Each solves
// (a T_x + c_m W_x) u = f / D_m in place, one workgroup of kTriThreads per mode, by the
// partition method: thread t owns the k = ceil(mx / kTriThreads) consecutive rows s..e of one
// block.  A downward elimination writes each of its rows as
//   x_i + al_i x_{s-1} + ga_i x_{i+1} = ph_i,
// an upward one as x_i = P_i - A_i x_{s-1} - G_i x_e (P_e = A_e = 0, G_e = -1).  Substituting
// the next block's first row into this block's last one leaves a tridiagonal system in the
// blocks' last unknowns y_t = x_{e_t}:
//   al_e y_{t-1} + (1 - ga_e A'_s) y_t - ga_e G'_s y_{t+1} = ph_e - ga_e P'_s   (' : block t+1),
// solved by parallel cyclic reduction in LDS (log2(kTriThreads) steps, one equation per
// thread); then x_i = P_i - A_i y_{t-1} - G_i y_t.  The operator is diagonally dominant, so the
// eliminations need no pivoting.  ph / P stay in the staged row; al / A and ga / G sit in LDS
// transposed (index j * nblk + t), so a wave's accesses fall in distinct banks.  LDS is sized
// to the row (40 KB in FP32 at 3001 nodes: four workgroups per CU).
constexpr int kTriThreads = 256;

// LDS of k_dct_tri for rows of mx nodes: the row, two coefficient arrays of k * nblk values
// (k rows per block, nblk blocks) and four arrays of the interface system
template <typename T>
static size_t dct_tri_lds(int mx) {
  const int k = (mx + kTriThreads - 1) / kTriThreads;
  const int nblk = (mx + k - 1) / k;
  return sizeof(T) * ((size_t)mx + 2 * (size_t)k * nblk + 4 * (size_t)kTriThreads);
}

template <typename T>
__global__ __launch_bounds__(kTriThreads) void k_dct_tri(int mx, int n, T a_off, double hxhy,
                                                         double cy, T* __restrict__ even,
                                                         T* __restrict__ odd) {
  extern __shared__ __align__(16) unsigned char tri_lds[];
  const int k = (mx + kTriThreads - 1) / kTriThreads;
  const int nblk = (mx + k - 1) / k;
  T* lrow = reinterpret_cast<T*>(tri_lds);  // right-hand side, then ph, P and x
  T* lal = lrow + mx;                       // al then A, at j * nblk + t
  T* lga = lal + (size_t)k * nblk;          // ga then G, at j * nblk + t
  T* ra = lga + (size_t)k * nblk;           // interface system; ra..rc first hold the first
  T* rb = ra + kTriThreads;                 // rows' P, A, G
  T* rc = rb + kTriThreads;
  T* rf = rc + kTriThreads;
  const int h = n / 2;
  const int blk = (int)blockIdx.x;
  const int mode = blk <= h ? 2 * blk : 2 * (blk - h - 1) + 1;
  T* row = blk <= h ? even + (size_t)blk * mx : odd + (size_t)(blk - h - 1) * mx;
  const double lam = 2.0 - 2.0 * std::cos(M_PI * (double)mode / (double)n);
  const T c = (T)(hxhy + cy * lam);
  const T inv_d = (mode == 0 || mode == n) ? T(1) / (T)n : T(2) / (T)n;
  const T b_in = a_off * T(2) + c, b_edge = a_off + c * T(0.5);
  const int tid = threadIdx.x;
  for (int i = tid; i < mx; i += kTriThreads) lrow[i] = row[i] * inv_d;
  const int s0 = tid * k;
  const int jl = tid < nblk ? min(s0 + k, mx) - 1 - s0 : -1;  // last row of the block, local
  __syncthreads();
  // downward over rows s0 .. s0 + jl
  T al = T(0), ga = T(0), ph = T(0);
  for (int j = 0; j <= jl; ++j) {
    const int i = s0 + j;
    const T ai = i > 0 ? -a_off : T(0);
    const T ci = i < mx - 1 ? -a_off : T(0);
    const T bi = (i == 0 || i == mx - 1) ? b_edge : b_in;
    const T inv = T(1) / (bi - ai * ga);  // ga = 0 at j = 0
    al = j == 0 ? ai * inv : -ai * al * inv;
    ph = (lrow[i] - ai * ph) * inv;
    ga = ci * inv;
    lrow[i] = ph;
    lal[j * nblk + tid] = al;
    lga[j * nblk + tid] = ga;
  }
  const T ale = al, gae = ga, phe = ph;  // the block's last row
  // upward: (ph, al, ga) -> (P, A, G) in place
  T P = T(0), A = T(0), G = T(-1);
  if (jl >= 0) {
    lrow[s0 + jl] = P;
    lal[jl * nblk + tid] = A;
    lga[jl * nblk + tid] = G;
  }
  for (int j = jl - 1; j >= 0; --j) {
    const int q = j * nblk + tid;
    const T g = lga[q];
    P = lrow[s0 + j] - g * P;
    A = lal[q] - g * A;
    G = -g * G;
    lrow[s0 + j] = P;
    lal[q] = A;
    lga[q] = G;
  }
  ra[tid] = P;
  rb[tid] = A;
  rc[tid] = G;
  __syncthreads();
  // interface system in the blocks' last unknowns (identity rows past the last block)
  {
    T ai = T(0), bi = T(1), ci = T(0), fi = T(0);
    if (tid < nblk) {
      ai = ale;
      fi = phe;
      if (tid + 1 < nblk) {
        bi = T(1) - gae * rb[tid + 1];
        ci = -gae * rc[tid + 1];
        fi = phe - gae * ra[tid + 1];
      }
    }
    __syncthreads();  // ra..rc are overwritten below
    ra[tid] = ai;
    rb[tid] = bi;
    rc[tid] = ci;
    rf[tid] = fi;
  }
  __syncthreads();
  for (int st = 1; st < kTriThreads; st *= 2) {
    const int im = tid - st, ip = tid + st;
    T ai = T(0), bi = rb[tid], ci = T(0), fi = rf[tid];
    if (im >= 0) {
      const T k1 = ra[tid] / rb[im];
      ai = -ra[im] * k1;
      bi -= rc[im] * k1;
      fi -= rf[im] * k1;
    }
    if (ip < kTriThreads) {
      const T k2 = rc[tid] / rb[ip];
      ci = -rc[ip] * k2;
      bi -= ra[ip] * k2;
      fi -= rf[ip] * k2;
    }
    __syncthreads();
    ra[tid] = ai;
    rb[tid] = bi;
    rc[tid] = ci;
    rf[tid] = fi;
    __syncthreads();
  }
  const T y = rf[tid] / rb[tid];  // y_t
  __syncthreads();
  ra[tid] = y;
  __syncthreads();
  const T ym = tid > 0 ? ra[tid - 1] : T(0);
  for (int j = 0; j <= jl; ++j) {
    const int q = j * nblk + tid;
    lrow[s0 + j] = lrow[s0 + j] - lal[q] * ym - lga[q] * y;
  }
  __syncthreads();
  for (int i = tid; i < mx; i += kTriThreads) row[i] = lrow[i];
}

template <typename T>
bool DctPrecond<T>::supported(int nx, int ny) {
  return ny >= 2 && !(ny & 1) && nx >= 1 && nx + 1 <= kMaxRow;
}

template <typename T>
DctPrecond<T>::~DctPrecond() {
  if (blas_) (void)rocblas_destroy_handle(blas_);
}

// two plain GEMMs in the element type in one call: C_i (m x n) = A_i op(B_i), column-major,
// the second operands at the given strides
static rocblas_status gemm2(rocblas_handle b, rocblas_operation tb, int m, int n, int k,
                            const float* A, int lda, long sa, const float* B, int ldb, long sb,
                            float* C, int ldc, long sc) {
  const float one = 1.0f, zero = 0.0f;
  return rocblas_sgemm_strided_batched(b, rocblas_operation_none, tb, m, n, k, &one, A, lda, sa,
                                       B, ldb, sb, &zero, C, ldc, sc, 2);
}
static rocblas_status gemm2(rocblas_handle b, rocblas_operation tb, int m, int n, int k,
                            const double* A, int lda, long sa, const double* B, int ldb, long sb,
                            double* C, int ldc, long sc) {
  const double one = 1.0, zero = 0.0;
  return rocblas_dgemm_strided_batched(b, rocblas_operation_none, tb, m, n, k, &one, A, lda, sa,
                                       B, ldb, sb, &zero, C, ldc, sc, 2);
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
  const size_t hh = (size_t)(h + 1) * (h + 1);
  // column-major (h+1) x (h+1) blocks: even block Ce[j + m (h+1)] = cos(pi j 2m / n)
  // (symmetric); odd block Co[j + m (h+1)] = cos(pi j (2m + 1) / n) for j, m < h, zero in row
  // and column h, so that the odd half runs as a GEMM of the even half's shape.  Stored as
  // [Ce, Co, Ce, Co^T]: the first pair serves the forward transform with B untransposed and
  // the inverse with B transposed, the second pair the other way round
  std::vector<T> cmat(4 * hh, T(0));
  for (int m = 0; m <= h; ++m)
    for (int j = 0; j <= h; ++j) {
      const T v = (T)std::cos(M_PI * (double)((long)j * 2 * m % (2 * n)) / (double)n);
      cmat[(size_t)j + (size_t)m * (h + 1)] = v;
      cmat[2 * hh + (size_t)j + (size_t)m * (h + 1)] = v;
    }
  for (int m = 0; m < h; ++m)
    for (int j = 0; j < h; ++j) {
      const T v = (T)std::cos(M_PI * (double)((long)j * (2 * m + 1) % (2 * n)) / (double)n);
      cmat[hh + (size_t)j + (size_t)m * (h + 1)] = v;
      cmat[3 * hh + (size_t)m + (size_t)j * (h + 1)] = v;
    }
  TMX_TRY(c_.alloc(cmat.size()));
  TMX_TRY(f_.alloc(2 * (size_t)(h + 1) * mx_));
  TMX_TRY(t_.alloc(2 * (size_t)(h + 1) * mx_));
  TMX_HIP(hipMemcpyAsync(c_.get(), cmat.data(), cmat.size() * sizeof(T), hipMemcpyHostToDevice, s));
  // (k_dct_fold rewrites the odd half's zero padding on every application)
  TMX_HIP(hipMemsetAsync(f_.get(), 0, f_.size() * sizeof(T), s));
  TMX_HIP(hipMemsetAsync(t_.get(), 0, t_.size() * sizeof(T), s));
  TMX_HIP(hipStreamSynchronize(s));
  // rows up to 3072 nodes need more than 64 KB of dynamic LDS in FP64
  TMX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_dct_tri<T>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)dct_tri_lds<T>(kMaxRow)));
  if (!blas_) TMX_TRY(blas_check(rocblas_create_handle(&blas_), "create_handle"));
  return 0;
}

template <typename T>
int DctPrecond<T>::apply(const T* r, T* z, RedSlot* rz, hipStream_t s) {
  const int h = h_, n = ny_, mx = mx_;
  const size_t total = (size_t)(h + 1) * mx;
  const unsigned g = (unsigned)std::min<size_t>(4096, (total + kBlock - 1) / kBlock);
  const long half = (long)(h + 1) * mx, hh = (long)(h + 1) * (h + 1);
  T* fe = f_.get();
  T* fo = fe + half;
  T* te = t_.get();
  T* to = te + half;
  k_dct_fold<T><<<g, kBlock, 0, s>>>(mx, n, r, fe, fo);
  TMX_HIP(hipGetLastError());
  TMX_TRY(blas_check(rocblas_set_stream(blas_, s), "set_stream"));
  // rocBLAS's kernels for these shapes on gfx950 (3001 x 501 x 501 at N=1000): with B
  // transposed 34 us against 44 us in FP32, untransposed 62 us against 72 us in FP64
  const rocblas_operation op = sizeof(T) == 4 ? rocblas_operation_transpose
                                              : rocblas_operation_none;
  const T* c_fwd = op == rocblas_operation_none ? c_.get() : c_.get() + 2 * hh;
  const T* c_inv = op == rocblas_operation_none ? c_.get() + 2 * hh : c_.get();
  // forward: T_e = S Ce, T_o = D Co (mx x h+1 each, one batched call)
  TMX_TRY(blas_check(gemm2(blas_, op, mx, h + 1, h + 1, fe, mx, half, c_fwd, h + 1, hh, te, mx,
                           half), "gemm"));
  // tridiagonal solves along x, one workgroup per mode
  const T a_off = (T)(eps_ * eps_ * hy_ / hx_);
  const double cy = eps_ * eps_ * hx_ / hy_;
  k_dct_tri<T><<<2 * h + 1, kTriThreads, dct_tri_lds<T>(mx), s>>>(mx, n, a_off, hx_ * hy_, cy,
                                                                   te, to);
  TMX_HIP(hipGetLastError());
  // inverse: E = T_e Ce^T, O = T_o Co^T
  TMX_TRY(blas_check(gemm2(blas_, op, mx, h + 1, h + 1, te, mx, half, c_inv, h + 1, hh, fe, mx,
                           half), "gemm"));
  const unsigned gu = (unsigned)std::min<size_t>(2048, (total + kBlock - 1) / kBlock);
  double* part = nullptr;
  if (rz) {
    rz->nblocks = (int)gu;
    part = rz->partials;
  }
  k_dct_unfold<T><<<gu, kBlock, 0, s>>>(mx, n, fe, fo, r, z, part);
  TMX_HIP(hipGetLastError());
  return 0;
}

template class DctPrecond<float>;
template class DctPrecond<double>;

}  // namespace tmx
