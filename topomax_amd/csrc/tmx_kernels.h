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

// Deterministic two-stage reductions: a kernel writes one partial per block into
// `partials` (kMaxPartials doubles) and records the block count; a one-block kernel adds them.
struct RedSlot {
  double* partials;
  int nblocks;
};

constexpr int kMaxPartials = 65536;

// Device scalar slots of one PCG run (indices into a double array).
struct PcgSlots {
  int rz, pq, rr, alpha, beta, flag, tol2, iters;
  int seq;  // iterations finished, stopped or not (reported to the host with the ring)
  // 1 once an iteration has run entirely after the stop flag was set (its deferred x update
  // done): the iterations after it change nothing, and a merged batch step skips them
  int skip;
};

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
  int stripw;    // output columns per wavefront (64 minus the halo lanes)
  int nstrips, nchunks;
  static SweepShape make(int nx, int ny, int rows, int stripw = 62);
  int blocks() const;
};

// kResidualRestrict: y = P^T (b - A x) on the nested coarse grid (nx, ny even);
// kProlongSmooth: one Jacobi sweep applied to x + P e (e = nested coarse correction).
// kApplyUpdate (PCG direction update fused with the operator): reads p_old = x and z = b
// (or, when ex.zlow is set, the float z = ex.zlow), forms p = z + beta p_old with
// beta = *ex.beta, writes p to ex.pnew and y = A p.
// A dot slot receives partials of x.y (kApply / kApplyUpdate: p.Ap) or b.y (sweeps: r.z).
enum SweepOp : int {
  kApply = 0, kResidual = 1, kSmooth = 2, kSmoothZero = 3, kResidualRestrict = 4,
  kProlongSmooth = 5, kApplyUpdate = 6
};

template <typename T>
struct SweepExtra {
  const T* ec = nullptr;         // kProlongSmooth: coarse correction
  T* pnew = nullptr;             // kApplyUpdate: updated direction
  const double* beta = nullptr;  // kApplyUpdate: device scalar
  const float* zlow = nullptr;   // kApplyUpdate: z in float (FP32 preconditioner output)
  // kApplyUpdate: the previous iteration's iterate update, deferred to here because this
  // kernel reads p_old anyway: xk += (*alpha) p_old
  T* xk = nullptr;
  const double* alpha = nullptr;
};

template <typename T>
struct ElastOpArgs {
  SweepShape sh;
  unsigned fixed;
  const T* xi;          // nodal xi (level 0) or nullptr
  const T* cellw;       // per-cell (wA, wB) means of r (coarser levels) or nullptr
  int p;
  T m, cp;              // rbar = m + cp * h_p(xi_a, xi_b, xi_c)
  T wmin;               // rbar is raised to at least wmin (preconditioner operators use m)
  ElastConst<T> k;
};

template <typename T>
hipError_t elast_sweep(const ElastOpArgs<T>& a, int op, const T* x, const T* b, T* y, T omega,
                       RedSlot* dot, hipStream_t s, const SweepExtra<T>& ex = SweepExtra<T>());

// Two block-Jacobi sweeps in one launch: op1 = kSmoothZero (y = S2 S1 0) or kProlongSmooth
// (y = S2 S1 (x + P ec)), with damping om1, om2 and node-row chunks of `rows`; a dot slot
// receives partials of b.y.  Same results as the two launches up to round-off.
template <typename T>
hipError_t elast_sweep_pair(const ElastOpArgs<T>& a, int op1, int rows, const T* x, const T* b,
                            T* y, T om1, T om2, RedSlot* dot, hipStream_t s, const T* ec);

// The zero-start pair (y = S2 S1 0) with the residual b - A y restricted to the nested coarse
// grid in the same launch (yc); nx, ny even, node-row chunks of `rows` (rounded down to even).
template <typename T>
hipError_t elast_sweep_pair_rr(const ElastOpArgs<T>& a, int rows, const T* b, T* y, T* yc,
                               T om1, T om2, hipStream_t s);

// per-cell mean r from nodal xi (for restriction of level-0 weights); with cellw32, also
// their FP32 copy
template <typename T>
hipError_t elast_cell_weights(int nx, int ny, const T* xi, int p, T m, T cp, T wmin, T* cellw,
                              hipStream_t s, float* cellw32 = nullptr);
// coarse cell means = mean of the 4 nested fine triangles of each coarse triangle (and with
// coarse32 their FP32 copy)
template <typename T>
hipError_t elast_restrict_weights(int nxc, int nyc, const T* fine, T* coarse, hipStream_t s,
                                  float* coarse32 = nullptr);
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
                        const T* b, T* y, T omega, RedSlot* dot, hipStream_t s,
                        const SweepExtra<T>& ex = SweepExtra<T>());

// The filter's finest level as two column pipelines (nx, ny even): down = nu block-Jacobi
// sweeps from zero (step sizes omega[0..nu)), y the smoothed iterate, yc the restricted
// residual; up = x + P ec, then nu sweeps into y (partials of b.y to dot).  nu <= kMaxPipeNu.
constexpr int kMaxPipeNu = 4;
template <typename T>
struct PipeOmega {
  T w[kMaxPipeNu];
};
template <typename T>
hipError_t filter_pipe_down(int nx, int ny, const FilterConst<T>& k, int nu, const T* omega,
                            int rows, const T* b, T* y, T* yc, hipStream_t s);
template <typename T>
hipError_t filter_pipe_up(int nx, int ny, const FilterConst<T>& k, int nu, const T* omega,
                          int rows, const T* x, const T* ec, const T* b, T* y, RedSlot* dot,
                          hipStream_t s);
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
// fused bilinear transfers of a 9-point level: bc = P^T (b - A x); y = sweep(x + P e)
template <typename T, int BS>
hipError_t stencil_resrestrict(int nxf, int nyf, int nxc, int nyc, unsigned fixed, const T* st,
                               const T* x, const T* b, T* bc, hipStream_t s);
template <typename T, int BS>
hipError_t stencil_prolong_smooth(int nxf, int nyf, int nxc, int nyc, unsigned fixed, const T* st,
                                  const T* x, const T* ec, const T* b, T* y, T omega,
                                  hipStream_t s);
// Galerkin coarse stencil Ac = P^T Af P for the given transfer (nested or bilinear); with
// stc32, also its FP32 copy
template <int BS>
hipError_t stencil_rap(int nxf, int nyf, int nxc, int nyc, int nested, unsigned fixed,
                       const double* stf, double* stc, hipStream_t s, float* stc32 = nullptr);
template <int BS>
hipError_t stencil_to_dense(int nx, int ny, const double* st, double* dense, hipStream_t s);

// Bottom of a V-cycle in one workgroup of kTailThreads: 9-point levels lev[0..nlev-2]
// (bilinear transfers, at most kTailThreads nodes each) and the dense coarsest lev[nlev-1]
// (x = Ainv b).  Every level's vectors live in LDS (lev[l].off, in elements of T: iterate,
// ping-pong and right-hand side, BS*nn each; the dense level holds b and x); b_in (global)
// is the top level's right-hand side and x_out (global) receives its result.
constexpr int kMaxTail = 8;
constexpr int kTailThreads = 1024;
constexpr int kTailNodes = kTailThreads;
constexpr size_t kTailLdsBytes = 64 * 1024;
template <typename T>
struct TailLevel {
  int nx, ny;
  const T* st;  // 9-point stencil (global, SoA as for stencil_op)
  int off;      // LDS offset of this level's vectors, in elements of T
};
constexpr int kMaxNu = 8;  // smoothing sweeps per pass in the tail kernel
template <typename T>
struct TailDesc {
  int nlev, cn, nu_pre, nu_post;
  unsigned fixed;
  T omega[kMaxNu];  // step size of sweep k of a smoothing pass
  const double* dense;
  const T* b_in;
  T* x_out;
  TailLevel<T> lev[kMaxTail];
};
// LDS bytes a tail over these levels needs (9-point levels nn[0..nlev-2], dense level nn[nlev-1])
size_t tail_lds_bytes(const size_t* nn, int nlev, int bs, size_t elem);
template <typename T, int BS>
hipError_t tail_vcycle(const TailDesc<T>& d, size_t lds_bytes, hipStream_t s);
hipError_t dense_invert(int n, double* a, double* work, hipStream_t s);
template <typename T>
hipError_t dense_apply(int n, const double* ainv, const T* b, T* x, hipStream_t s);
template <typename TI, typename TO>
hipError_t convert(size_t n, const TI* in, TO* out, hipStream_t s);
// out64 = in and out32 = (float) in in one pass
hipError_t copy_split(size_t n, const double* in, double* out64, float* out32, hipStream_t s);

// ------------------------------------------------------------------ vectors
// independent vector copies in one launch (tasks: device array of count entries, <= 65535)
struct CopyTask {
  const double* src;
  double* dst;
  size_t n;
};
hipError_t copy_tasks(const CopyTask* tasks, int count, hipStream_t s);
// *host_dst = *src by a kernel (host_dst: device address of pinned host memory)
hipError_t publish_scalar(const double* src, double* host_dst, hipStream_t s);
template <typename T>
hipError_t dot_t(size_t n, const T* a, const T* b, RedSlot& r, hipStream_t s);
// y += (*a) x with a device scalar; with gate, only where *gate != 0 when the kernel runs
template <typename T>
hipError_t axpy_dev(size_t n, const double* a, const T* x, T* y, hipStream_t s,
                    const double* gate = nullptr);
// out = sum of r's partials (one block)
// (with host_out, the sum is also written to that mapped host memory, as publish_scalar does)
hipError_t reduce_to(const RedSlot& r, double* out, hipStream_t s, double* host_out = nullptr);
// reliable-update PCG: x += xs + (*alpha) p, xs = 0; and r = (float) r64 with partials of |r64|^2
hipError_t ru_flush(size_t n, const double* alpha, const float* p, float* xs, double* x,
                    hipStream_t s);
hipError_t ru_replace(size_t n, const double* r64, float* r, RedSlot& rr, hipStream_t s);
// y += a
hipError_t add_vec(size_t n, const double* a, double* y, hipStream_t s);
// dot_t + reduce_to
hipError_t dot(size_t n, const double* a, const double* b, RedSlot& r, double* out,
               hipStream_t s, double* host_out = nullptr);
// PCG scalars on the device.  flag: 0 running, 1 converged, 2 p.Ap <= 0, 3 non-finite.
// alpha: pq = sum; alpha = rz/pq (0 once stopped).  rr: rr = sum; converged if rr <= tol2.
// beta: rznew = sum; beta = rznew/rz; rz = rznew.
hipError_t pcg_alpha(const RedSlot& r, double* sc, const PcgSlots& q, hipStream_t s);
hipError_t pcg_rr(const RedSlot& r, double* sc, const PcgSlots& q, hipStream_t s);
hipError_t pcg_beta(const RedSlot& r, double* sc, const PcgSlots& q, hipStream_t s);
// x += alpha p; r -= alpha q (alpha = sc[i_alpha]); partials of r.r
template <typename T>
hipError_t pcg_update_xr(size_t n, const double* sc, int i_alpha, T* x, T* r, const T* p,
                         const T* q, RedSlot& rr, hipStream_t s);
// as pcg_update_xr, with alpha = rz / sum(pq partials) formed by every block (0 once stopped
// or if p.Ap <= 0); rlow (optional) receives a float copy of the new r.  x = nullptr: the x
// update is deferred (alpha is stored in sc[q.alpha] for the next direction update)
template <typename T>
hipError_t pcg_update_xr_fused(size_t n, double* sc, const PcgSlots& q, const RedSlot& pq,
                               T* x, T* r, const T* p, const T* qv, RedSlot& rr, hipStream_t s,
                               float* rlow = nullptr);
// end of a fused iteration: pq, rr, rz from partials; flags; beta = rz_new / rz; rz = rz_new
hipError_t pcg_finalize(const RedSlot& pq, const RedSlot& rr, const RedSlot& rz, double* sc,
                        const PcgSlots& q, hipStream_t s, double* host_out = nullptr);
// start of a fused run: rz from partials, beta = 0, flag = 0, iters = 0
hipError_t pcg_start(const RedSlot& rz, double* sc, const PcgSlots& q, hipStream_t s);
// PCG set-up from x0 = 0 without a host round trip: r = b, x = 0, rlow = float(b) (optional)
// and partials of b.b in one pass; then one block forms bb, r.r = bb, tol2 = rtol2 * bb and the
// start flag (1: b = 0, nothing to solve; 3: non-finite b; else 0), stores bb in sc[bb_slot]
// and in pinned host memory (host_bb[0])
template <typename T>
hipError_t pcg_begin(size_t n, const T* b, T* x, T* r, float* rlow, RedSlot& bb, hipStream_t s);
hipError_t pcg_begin_final(const RedSlot& bb, double* sc, const PcgSlots& q, double rtol2,
                           int bb_slot, double* host_bb, hipStream_t s);
// p = z + beta p, beta = sc[i_beta]
template <typename T>
hipError_t pcg_update_p(size_t n, const double* sc, int i_beta, const T* z, T* p, hipStream_t s);
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
// sc[i] = v (kernel argument, safe to enqueue before earlier work completes)
hipError_t set_scalar(double* sc, int i, double v, hipStream_t s);
// zero both components of the Dirichlet nodes of an interleaved vector
hipError_t mask_fixed(int nx, int ny, unsigned fixed, double* v, hipStream_t s);

// entropic mirror descent (nodal designs, lumped-mass integrals)
hipError_t emd_half_step(size_t n, const double* psi, const double* w, double step, double* half,
                         hipStream_t s);
// out0 = int expit(half + c), out1 = int expit'(half + c) (device scalars)
hipError_t emd_volume(size_t n, const double* half, double c, const double* m, RedSlot& r0,
                      RedSlot& r1, double* out0, double* out1, hipStream_t s);
hipError_t emd_apply(size_t n, const double* half, double c, double* psi, double* rho,
                     hipStream_t s);
// out = int (rho - expit(psi_prev))^2 (device scalar)
hipError_t emd_change(size_t n, const double* rho, const double* psi_prev, const double* m,
                      RedSlot& r, double* out, hipStream_t s);
hipError_t logit(size_t n, const double* rho, double* psi, hipStream_t s);

// MINRES (indefinite K): scalar slots and steps, see tmx_kernels.hip
struct MrSlots {
  int beta1, beta, oldb, alfa, dbar, epsln, oldeps, delta, phibar, phi, cs, sn, denom, iters,
      tol, flag;
  int skip;  // 1 while an iteration runs after the stop: it changes no iterate
  int seq;   // iterations finished, stopped or not (reported to the host with the ring)
};
hipError_t mr_start(const RedSlot& rz, double* sc, const MrSlots& q, double tol, hipStream_t s);
hipError_t mr_scale(size_t n, const double* sc, const MrSlots& q, const double* z, double* v,
                    hipStream_t s);
// y -= (beta/oldb) r1; alfa = v.y; y -= (alfa/beta) r2
hipError_t mr_lanczos(size_t n, double* sc, const MrSlots& q, const double* v, const double* r1,
                      const double* r2, double* y, RedSlot& part, hipStream_t s);
// its second half alone: alfa from the partials of v.y in part; y -= (alfa/beta) r2 (for an
// operator whose launch already formed y - (beta/oldb) r1 and those partials)
hipError_t mr_lanczos2(size_t n, double* sc, const MrSlots& q, const RedSlot& part,
                       const double* r2, double* y, hipStream_t s);
hipError_t mr_step(const RedSlot& rz, double* sc, const MrSlots& q, double* host_out,
                   hipStream_t s);
hipError_t mr_update(size_t n, const double* sc, const MrSlots& q, double* v, const double* z,
                     double* w_out, const double* w_older, const double* w_newer, double* x,
                     hipStream_t s);

// out = a - b
hipError_t sub_vec(size_t n, const double* a, const double* b, double* out, hipStream_t s);

// fine CG1 values from a coarser CG1 function on the same rectangle ('right' meshes)
hipError_t interp_p1(int nxc, int nyc, double hxc, double hyc, int nxf, int nyf, double hxf,
                     double hyf, const double* rc, double* rf, hipStream_t s);

// v[i] = pseudo-random value in [-1, 1] from a hash of i (power-iteration start vector)
hipError_t fill_hash(size_t n, double* v, hipStream_t s);

// out = c * in
hipError_t scale_copy(size_t n, double c, const double* in, double* out, hipStream_t s);

}  // namespace tmx
