// P2 displacement element on the structured 'right' mesh: the reference's own element
// (FEM_src/elasisity_problem.py:194-196, VectorElement("CG", triangle, 2)).
//
// Unknowns live on the (2nx+1) x (2ny+1) lattice of vertices and edge midpoints, node
// J*(2nx+1) + I at doubled grid indices (I, J); displacements are node-interleaved.  A cell
// (cx, cy) spans lattice [2cx, 2cx+2] x [2cy, 2cy+2] with triangles A = (v0, v1, v3) and
// B = (v0, v2, v3); local P2 node order is the three vertices, then the midpoints opposite
// vertex 0, 1, 2.
#pragma once

#include <hip/hip_runtime.h>

#include "tmx_kernels.h"

namespace tmx {

struct P2Geom {
  int nx, ny;          // cells
  double hx, hy;       // cell size
  double lam, mu;      // Lame parameters
  unsigned fixed;      // fixed sides (SideBits)
  size_t nodes() const { return (size_t)(2 * nx + 1) * (2 * ny + 1); }
};

// Per-triangle weights W_lk = int_T r(xi_h) l_l l_k (6 per triangle: 00 01 02 11 12 22, 12 per
// cell, A then B), exact for integer p: r = m + (1-m) xi^p with xi the P1 interpolant of the
// nodal xi.  positive_part: use max(xi, 0) (the preconditioner's SPD stand-in).
hipError_t p2_weights(const P2Geom& g, const double* xi, int p, double m, bool positive_part,
                      double* W, hipStream_t s);
// y = K u for the Dirichlet-masked P2 stiffness (identity rows and columns on fixed nodes)
hipError_t p2_apply(const P2Geom& g, const double* W, const double* u, double* y, hipStream_t s);
// MINRES's operator step: y = K v - (*beta / *oldb) r1 (0 when *oldb is 0) with the partials of
// v.y in part (one per tile; their count in *nparts)
hipError_t p2_apply_lanczos(const P2Geom& g, const double* W, const double* v, const double* r1,
                            double* y, const double* beta, const double* oldb, double* part,
                            int* nparts, hipStream_t s);
// whether the grid's tiles fit the partials buffer of p2_apply_lanczos
bool p2_apply_lanczos_ok(const P2Geom& g);
// 2x2 diagonal blocks of K (3 per node: xx, xy, yy)
hipError_t p2_diag(const P2Geom& g, const double* W, double* D, hipStream_t s);
// one damped block-Jacobi sweep y = x + omega D^-1 (b - K x) (x = nullptr: zero); fixed nodes
// 0.  T = double, or float for the mixed scheme's preconditioner (W, D and vectors in float).
template <typename T>
hipError_t p2_smooth(const P2Geom& g, const T* W, const T* D, const T* x, const T* b, T* y,
                     double omega, hipStream_t s);
// the float cycle's first sweep (from zero) reading a double b, which it also writes rounded
// to b32, and its last sweep writing a double y
hipError_t p2_smooth_first32(const P2Geom& g, const float* W, const float* D, const double* b,
                             float* b32, float* y, double omega, hipStream_t s);
hipError_t p2_smooth_last32(const P2Geom& g, const float* W, const float* D, const float* x,
                            const float* b, double* y, double omega, hipStream_t s);
// r = b - K x
template <typename T>
hipError_t p2_residual(const P2Geom& g, const T* W, const T* x, const T* b, T* r, hipStream_t s);
// P1 <-> P2 transfers on the same mesh (P1 is a subspace of P2): r1 = P^T r2, x2 += P e1
template <typename T>
hipError_t p2_restrict(const P2Geom& g, const T* r2, T* r1, hipStream_t s);
template <typename T>
hipError_t p2_prolong_add(const P2Geom& g, const T* e1, T* x2, hipStream_t s);
// load vector of the P2 test functions (body force and traction P2 interpolants, exact)
hipError_t p2_assemble_load(const LoadDesc& d, double* F, hipStream_t s);
// zero both components of every fixed lattice node
hipError_t p2_mask_fixed(const P2Geom& g, double* v, hipStream_t s);
// sensitivity rhs on CG1: b_v = int -r'(xi) sigma0(u):eps(u) phi_v with u in P2 (exact)
hipError_t p2_sensitivity(const P2Geom& g, const double* xi, const double* u, int p, double m,
                          double* b, hipStream_t s);

}  // namespace tmx
