// Exact fast solver of the filter's spectrally equivalent operator, used as its PCG
// preconditioner (FP32 in the mixed scheme, FP64 in the FP64 scheme).
//
// With natural boundary conditions on the structured 'right' mesh the filter stiffness is a
// sum of Kronecker products of 1D Neumann Laplacians T (tridiag(-1, 2, -1), boundary diagonal
// 1) with the boundary half-weights W = diag(1/2, 1, ..., 1, 1/2):
//   eps^2 K = eps^2 (hy/hx) W_y (x) T_x + eps^2 (hx/hy) T_y (x) W_x   (y outer, x inner).
// Replacing the consistent mass by hx hy W_y (x) W_x gives
//   A' = W_y (x) B_x + T_y (x) C_x,  B_x = eps^2 (hy/hx) T_x + hx hy W_x,  C_x = eps^2 (hx/hy) W_x,
// which the DCT-I in y diagonalises: T_y v_m = lambda_m W_y v_m with v_m[j] = cos(pi j m / n),
// lambda_m = 2 - 2 cos(pi m / n), v_m' W_y v_k = D_m delta_mk (D_m = n / 2, or n for m = 0, n).
// So A'^-1 r = sum_m v_m (x) (B_x + lambda_m C_x)^-1 (v_m' r) / D_m: a DCT-I along y, one
// tridiagonal solve along x per mode (partition method: block eliminations, cyclic reduction
// of the blocks' interface system in LDS), and the same DCT-I again (the matrix
// C[j, m] = cos(pi j m / n) is symmetric).  A'^-1 A has its spectrum in [0.95, 1.003] for
// eps/h = 2 and closer to 1 as eps/h grows (the mass is a small part of the operator there),
// so PCG needs 3-5 iterations to 1e-9 whatever the grid.
//
// The DCT-I of the n + 1 values of a y-column is computed from a complex FFT of length 2n in
// LDS (k_dct1_cols): two columns a, b are packed as a + i b and extended evenly
// (z[2n - j] = z[j]); the DFT of a real even sequence is real, so the real and imaginary parts
// of the FFT are the two columns' transforms, up to the end-point terms
//   T[m] = (Y[m] + r_0 + (-1)^m r_n) / 2,   Y = DFT_2n of the even extension.
// The FFT is a mixed-radix Stockham transform (radices 2, 3, 4, 5, 7), so 2n must have no
// prime factor above 7; other grids keep the multigrid V-cycle.
#pragma once

#include "tmx_kernels.h"
#include "tmx_solver.h"

namespace tmx {

// radix plan of one FFT length (stages applied in order)
struct FftPlan {
  static constexpr int kMaxStages = 16;
  int L = 0;
  int nstages = 0;
  int radix[kMaxStages] = {};
};

// the radix plan of length L from radices {4, 2, 3, 5, 7}; false when L has a larger prime
bool fft_plan(int L, FftPlan& plan);

template <typename T>
class DctPrecond {
 public:
  // nx, ny: cells; 2 ny factors into 2, 3, 5, 7, nx + 1 <= kMaxRow, and the FFT of 2 ny
  // points fits the LDS budget; eps: the filter's length scale
  static bool supported(int nx, int ny);
  int init(int nx, int ny, double hx, double hy, double eps, hipStream_t s);
  // z = A'^-1 r (vectors of (nx + 1)(ny + 1) values), partials of r.z into rz if given
  int apply(const T* r, T* z, RedSlot* rz, hipStream_t s);

  // longest x row: one workgroup of 256 threads solves it, 12 rows per thread, with the row
  // and two coefficient arrays in LDS (40 KB in FP32, 80 KB in FP64 at this length)
  static constexpr int kMaxRow = 3072;
  // fewest y-columns per FFT workgroup (two per complex sequence of 2 ny points; ny <= 2048);
  // four when their sequences and the twiddle table fit 64 KB of LDS
  static constexpr int kCols = 2;

 private:
  int nx_ = 0, ny_ = 0, mx_ = 0;
  double hx_ = 0, hy_ = 0, eps_ = 0;
  FftPlan plan_;
  DevBuf<T> tw_;  // twiddles exp(-2 pi i k / 2n), k < 2n, as (re, im) pairs
  DevBuf<T> t_;   // transformed modes: row m = mode m (n + 1 rows of nx + 1 values)
};

}  // namespace tmx
