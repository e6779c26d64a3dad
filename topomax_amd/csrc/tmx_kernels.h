// Host-callable launchers for the topomax_amd gfx950 kernels.
//
// Grid conventions (all levels): nx x ny cells, (nx+1) x (ny+1) nodes, node
// v = iy*(nx+1) + ix; block size BS = 2 (elasticity, node-interleaved (ux,uy)) or
// BS = 1 (Helmholtz filter).  Every launcher enqueues on `stream` and returns
// hipGetLastError().
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace tmx {

// Dirichlet side bits (same values as TMX_FIX_* in include/tmx.h)
enum SideBits : unsigned { kLeft = 1u, kRight = 2u, kTop = 4u, kBottom = 8u };

struct RedSlot {          // one deterministic reduction target
  double* partials;       // scratch, >= kMaxPartials doubles
  unsigned* counter;      // zero-initialised, reset by the kernel
  double* result;         // device scalar
};

constexpr int kMaxPartials = 65536;

// ------------------------------------------------------------------ elasticity
template <typename T>
struct ElastConst {       // level constants, see ElastConst<T>::make
  T l2m_a, lam_h, l2m_b, mu_h, mu_a, mu_b;  // element force coefficients
  T Pxx, Pyy, Qxx, Qyy, Rxx, Rxy, Ryy;      // diagonal 2x2 blocks (sym)
  static ElastConst make(double lam, double mu, double hx, double hy);
};

struct SweepShape {
  int nx, ny;
  int rows;      // node rows per wavefront chunk
  int nstrips, nchunks;
  static SweepShape make(int nx, int ny, int rows);
  int blocks() const;
};

enum SweepOp : int { kApply = 0, kResidual = 1, kSmooth = 2, kSmoothZero = 3 };

template <typename T>
struct ElastOpArgs {
  SweepShape sh;
  unsigned fixed;
  const T* xi;          // nodal xi (level 0) or nullptr
  const T* cellw;       // per-cell (wA, wB) means of r (coarser levels) or nullptr
  int p;
  T m, cp;              // rbar = m + cp * h_p(xi_a, xi_b, xi_c)
  ElastConst<T> k;
};

template <typename T>
hipError_t elast_sweep(const ElastOpArgs<T>& a, int op, const T* x, const T* b, T* y, T omega,
                       const RedSlot* dot_xy, hipStream_t s);

// per-cell mean r from nodal xi (for restriction of level-0 weights)
template <typename T>
hipError_t elast_cell_weights(int nx, int ny, const T* xi, int p, T m, T cp, T* cellw,
                              hipStream_t s);
// coarse cell means = mean of the 4 nested fine triangles of each coarse triangle
template <typename T>
hipError_t elast_restrict_weights(int nxc, int nyc, const T* fine, T* coarse, hipStream_t s);
// explicit 9-point block stencil (BS=2, SoA: st[k*4+a*2+b][node]) of an element level
template <typename T>
hipError_t elast_to_stencil(const ElastOpArgs<T>& a, T* st, hipStream_t s);

// sensitivity rhs b_v = -p(1-m) sum_T e_T int_T xi^(p-1) lambda_v (P1)
hipError_t elast_sensitivity(const SweepShape& sh, const double* xi, const double* u, int p,
                             double m, double lam, double mu, double hx, double hy, double* out,
                             hipStream_t s);

// load vector of body force + tractions (P2-interpolated rules, exact P2xP1 integrals)
struct LoadDesc {
  int nx, ny;
  double width, height, ab, cd;  // ab = width/nx, cd = height/ny
  int has_body;
  double bc[2], br, bv[2];
  int ntr;
  int tside[16];
  double tlo[16], thi[16], tv[16][2];
};
hipError_t elast_assemble_load(const LoadDesc& d, double* F, hipStream_t s);

// ------------------------------------------------------------------ filter (scalar P1)
template <typename T>
struct FilterConst {
  T ka, kb, mA;                 // eps^2 a/2, eps^2 b/2, |T|/12
  static FilterConst make(double eps, double hx, double hy);
};

template <typename T>
hipError_t filter_sweep(const SweepShape& sh, const FilterConst<T>& k, int op, const T* x,
                        const T* b, T* y, T omega, const RedSlot* dot_xy, hipStream_t s);
template <typename T>
hipError_t filter_to_stencil(int nx, int ny, const FilterConst<T>& k, T* st, hipStream_t s);

// ------------------------------------------------------------------ grid transfers
// nested P1 ('right' mesh refinement): fine (2nxc, 2nyc) <-> coarse (nxc, nyc)
template <typename T, int BS>
hipError_t restrict_nested(int nxc, int nyc, unsigned fixed, const T* rf, T* bc, hipStream_t s);
template <typename T, int BS>
hipError_t prolong_nested(int nxc, int nyc, unsigned fixed, const T* ec, T* xf, hipStream_t s);
// tensor-product 1-D linear interpolation, coarse cells = ceil(fine/2) (1 stays 1)
template <typename T, int BS>
hipError_t restrict_bilinear(int nxf, int nyf, int nxc, int nyc, unsigned fixed, const T* rf,
                             T* bc, hipStream_t s);
template <typename T, int BS>
hipError_t prolong_bilinear(int nxf, int nyf, int nxc, int nyc, unsigned fixed, const T* ec,
                            T* xf, hipStream_t s);

// ------------------------------------------------------------------ 9-point stencil levels
template <typename T, int BS>
hipError_t stencil_op(int nx, int ny, const T* st, int op, const T* x, const T* b, T* y,
                      T omega, hipStream_t s);
// Galerkin coarse stencil Ac = P^T Af P for the given transfer (nested or bilinear)
template <int BS>
hipError_t stencil_rap(int nxf, int nyf, int nxc, int nyc, int nested, unsigned fixed,
                       const double* stf, double* stc, hipStream_t s);
template <int BS>
hipError_t stencil_to_dense(int nx, int ny, const double* st, double* dense, hipStream_t s);
hipError_t dense_invert(int n, double* a, double* work, hipStream_t s);
template <typename T>
hipError_t dense_apply(int n, const double* ainv, const T* b, T* x, hipStream_t s);
template <typename TI, typename TO>
hipError_t convert(size_t n, const TI* in, TO* out, hipStream_t s);

// ------------------------------------------------------------------ vectors
hipError_t dot(size_t n, const double* a, const double* b, const RedSlot& r, hipStream_t s);
template <typename T>
hipError_t dot_t(size_t n, const T* a, const T* b, const RedSlot& r, hipStream_t s);
// PCG: alpha = sc[rz]/sc[pq]; x += alpha p; r -= alpha q; rr = r.r
template <typename T>
hipError_t pcg_update_xr(size_t n, const double* sc, int i_rz, int i_pq, T* x, T* r, const T* p,
                         const T* q, const RedSlot& rr, hipStream_t s);
// PCG: beta = sc[rznew]/sc[rz]; p = z + beta p (the caller alternates the two rz slots)
template <typename T>
hipError_t pcg_update_p(size_t n, double* sc, int i_rznew, int i_rz, const T* z, T* p,
                        hipStream_t s);
template <typename T>
hipError_t copy_vec(size_t n, const T* a, T* b, hipStream_t s);
// x64 += sqrt(sc[i]) * d (d in T): undoes the scaling of scale_to
template <typename T>
hipError_t axpy_scaled(size_t n, const double* sc, int i, const T* d, double* x, hipStream_t s);
// out = in / sqrt(sc[i]) converted to T (scale for mixed-precision refinement)
template <typename T>
hipError_t scale_to(size_t n, const double* sc, int i, const double* in, T* out, hipStream_t s);
hipError_t fill_zero_bytes(void* p, size_t bytes, hipStream_t s);
hipError_t fill(size_t n, double val, double* p, hipStream_t s);
// zero both components of the Dirichlet nodes of an interleaved vector
hipError_t mask_fixed(int nx, int ny, unsigned fixed, double* v, hipStream_t s);

}  // namespace tmx
