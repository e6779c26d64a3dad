// Host-side multigrid hierarchy and Krylov drivers for topomax_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <functional>
#include <vector>

#include "tmx_kernels.h"

namespace tmx {

// ------------------------------------------------------------------ errors
struct Error {
  int code;
  std::string msg;
};
void set_error(const std::string& m);

#define TMX_HIP(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      ::tmx::set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr);  \
      return -2;                                                                             \
    }                                                                                        \
  } while (0)

#define TMX_TRY(expr)        \
  do {                       \
    int rc_ = (expr);        \
    if (rc_ != 0) return rc_; \
  } while (0)

// ------------------------------------------------------------------ device buffers
template <typename T>
class DevBuf {
 public:
  DevBuf() = default;
  ~DevBuf() { reset(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { reset(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  int alloc(size_t n) {
    reset();
    if (n == 0) return 0;
    if (hipMalloc(&p_, n * sizeof(T)) != hipSuccess) {
      p_ = nullptr;
      set_error("hipMalloc failed for " + std::to_string(n * sizeof(T)) + " bytes");
      return -2;
    }
    n_ = n;
    return 0;
  }
  void reset() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

// Reduction scratch and device scalars of one stream of work.
struct Workspace {
  static constexpr int kRegions = 4;
  static constexpr int kScalars = 64;
  DevBuf<double> partials;  // kRegions x kMaxPartials
  DevBuf<double> scalars;
  double* host = nullptr;   // pinned: [0, kScalars) mirror + ring for lagged reads
  double* host_dev = nullptr;  // device address of `host`, for kernels writing the ring
  int init(hipStream_t s);
  ~Workspace();
  RedSlot region(int k) const { return RedSlot{partials.get() + (size_t)k * kMaxPartials, 0}; }
  double* sc() const { return scalars.get(); }
  int read(hipStream_t s, int first, int count);  // scalars[first..) -> host[first..), sync
};

// scalar slots used outside PCG
enum : int { kScBB = 0, kScRR64 = 1, kScJ = 2, kScInt = 3, kScTmp = 4, kScVerify = 6 };
// pinned ring (after the kScalars mirror): PCG report slots at 0 and 4, MINRES at 8, |b|^2 of
// the current PCG solve at kRingBB
constexpr int kRingBB = 24;
// PCG slots (flag, rr, iters contiguous for one async read)
constexpr PcgSlots kPcg = {10, 11, 16, 12, 13, 15, 14, 17, 18, 19};

enum LevelKind : int { kElem = 0, kStencil = 1, kDense = 2 };
enum Phys : int { kElasticity = 0, kFilter = 1 };

struct Level {
  int nx = 0, ny = 0;
  double hx = 0, hy = 0;
  int kind = kElem;
  int nested_from_prev = 0;  // transfer from level-1 to this level is nested P1
  size_t nn = 0;
  SweepShape sh64, sh32;
  // elasticity element levels > 0: per-cell (wA, wB)
  DevBuf<double> cw64;
  DevBuf<float> cw32;
  // 9-point stencils
  DevBuf<double> st64;
  DevBuf<float> st32;
  // V-cycle vectors (levels > 0 use b/x; all levels use t and r)
  DevBuf<double> b64, x64, t64, r64;
  DevBuf<float> b32, x32, t32, r32;
};

// solve precision: FP64 PCG with an FP64 V-cycle, FP64 PCG with an FP32 V-cycle, or FP32 PCG
// inside FP64 iterative refinement
enum SolvePrecision { kSolveFP64 = 0, kSolveMixed = 1, kSolveMixedIR = 2, kSolveMixedRU = 3 };

// Waits for the stream's work by polling it: the host is back within microseconds of the last
// kernel (a blocking synchronisation wakes the thread up tens of microseconds later).  Returns
// the stream's error, if any.
hipError_t stream_wait(hipStream_t s);

struct SolveParams {
  double rtol = 1e-10;
  int max_iter = 500;
  int precision = kSolveFP64;
  double inner_rtol = 1e-4;  // measured fastest at N=1000 (fewer inner iterations in total)
  int max_outer = 30;
  // FP64 / mixed: report the true residual |b - A x| / |b| (one more operator application and
  // a host round trip) rather than PCG's recursively updated one
  bool verify = true;
  // FP64 / mixed: false returns with the solution's last kernels (and the true residual)
  // still queued, so the caller's next work follows without a host round trip; the caller
  // synchronises the stream and then calls SolveStats::complete()
  bool sync = true;
};

struct SolveStats {
  int iterations = 0, outer = 0, levels = 0;
  double rel_residual = 0, setup_ms = 0, solve_ms = 0;
  // left by a solve with sync = false: the pinned |b - A x|^2 (with |b|^2) and the events
  // around the solve's device work, read by complete() once the stream has drained
  const double* pending_rr = nullptr;
  double pending_bb = 0;
  hipEvent_t pending_ev[2] = {nullptr, nullptr};
  void complete();
};

// Preconditioned MINRES (FP64) on A x = b from x0 = 0, restarted on the true residual until
// |b - A x| <= rtol |b|.  apply(x, y): y = A x.  precond(r, z, rz): z = M^-1 r with M SPD and
// the partials of r.z in rz.  bufs: eight vectors of n doubles (two residuals, two work
// vectors, the Lanczos vector and three directions); the result is left in x.
struct MinresBufs {
  double *r1, *r2, *y, *z, *v, *w0, *w1, *w2;
};
// apply_l1 (optional): y = A v - (*beta / *oldb) r1 (0 when *oldb is 0) with the partials of
// v.y in part, in one launch; used in place of apply and the first Lanczos update.
using MinresApplyL1 = std::function<int(const double* v, const double* r1, double* y,
                                        RedSlot* part, const double* beta, const double* oldb)>;
int minres_core(Workspace& ws, size_t n, const double* b, double* x, MinresBufs bufs,
                const std::function<int(const double*, double*)>& apply,
                const std::function<int(const double*, double*, RedSlot*)>& precond,
                double rtol, int max_iter, double bnorm2, int* iters, double* rel,
                hipStream_t s, bool trace, const MinresApplyL1& apply_l1 = nullptr);

// Geometric multigrid + PCG for one operator family on one grid.
template <typename T>
class DctPrecond;

class Hierarchy {
 public:
  int init(Phys phys, int nx, int ny, double hx, double hy, unsigned fixed, double lam,
           double mu, double eps, int coarse_max_nodes, hipStream_t s);
  // elasticity: copy nodal xi (device, fp64) and rebuild the coarse operators, penalty p.
  // With overlap, the operators below the cell-weight levels (element-to-stencil, Galerkin
  // products, coarsest inverse) are built on the hierarchy's side stream, concurrently with
  // what s runs next; vcycle() makes s wait for them before the first level that reads them,
  // and solve() before it returns (join_coarse).
  int update_elasticity(const double* xi, int p, double m, hipStream_t s, bool overlap = false);
  // make s wait for a coarse-operator build still running on the side stream
  int join_coarse(hipStream_t s);
  // the preconditioner's per-cell weights of the finest cells from the current xi, p, m
  int fine_cell_weights(hipStream_t s);
  // elasticity fine operator only (copies xi; no coarse rebuild)
  int set_fine_operator(const double* xi, int p, double m, hipStream_t s);
  // filter: build the constant hierarchy once
  int build_filter(hipStream_t s);

  template <typename T>
  int apply(const T* x, T* y, RedSlot* dot, hipStream_t s);
  template <typename T>
  int residual(const T* x, const T* b, T* r, hipStream_t s);
  template <typename T>
  int smooth0(const T* x, const T* b, T* y, hipStream_t s);
  // two sweeps of elasticity element level l (default the finest) in one launch: op1 =
  // kSmoothZero (x unused) or kProlongSmooth (x + P e with e = level l + 1's x), chunk height
  // rows
  template <typename T>
  int sweep_pair0(int op1, int rows, const T* x, const T* b, T* y, double om1, double om2,
                  RedSlot* dot, hipStream_t s, int l = 0);
  // one launch of the filter's finest-level pipeline (down or up) on the hierarchy's own
  // level buffers, for timing (tmx_autotune)
  template <typename T>
  int filter_pipe0(bool down, int rows, hipStream_t s);
  // the zero-start pair with residual + restriction to level l + 1's right-hand side fused
  template <typename T>
  int sweep_pair_rr0(int rows, const T* b, T* y, double om1, double om2, hipStream_t s,
                     int l = 0);
  // z = M^-1 b; at level 0 the last smoothing sweep also writes partials of b.z to dot_bz
  template <typename T>
  int vcycle(int l, const T* b, T* out, hipStream_t s, RedSlot* dot_bz = nullptr);
  // q = A p with p = z + beta p_old written to p_new; partials of p.q to dot.  z is read from
  // zlow (float) when that is given.
  // With xk, also xk += (*alpha) p_old (the previous iteration's deferred iterate update).
  template <typename T>
  int apply_update(const T* p_old, const T* z, T* q, T* p_new, const double* beta, RedSlot* dot,
                   hipStream_t s, const float* zlow = nullptr, T* xk = nullptr,
                   const double* alpha = nullptr);

  // PCG from x0 = 0 on A x = b with Krylov vectors in T and the V-cycle preconditioner in PC;
  // result left in the hierarchy's Krylov x buffer (x_buf<T>())
  // (|b|^2 is formed on the device and returned in *bnorm2)
  template <typename T, typename PC = T>
  int pcg(Workspace& ws, const T* b, double rtol, int max_iter, double* bnorm2, int* iters,
          double* rel, hipStream_t s);
  template <typename T>
  T* x_buf();
  // The launches of pcg() without its host loop, for callers that drive the iterations
  // themselves (merged batches): the start (r = b, x = 0, |b|^2 and the tolerance into the
  // device scalars and the pinned ring, then, with init, z = M r and the first direction's
  // scalars), one iteration of the given parity (the ring slot 4 * parity receives its report),
  // and the end (the last iteration's deferred x += alpha p, given the number of iterations
  // launched).  pcg() itself is built from them.
  template <typename T, typename PC = T>
  int pcg_enqueue_begin(Workspace& ws, const T* b, double rtol, bool init, hipStream_t s);
  template <typename T, typename PC = T>
  int pcg_enqueue_iteration(Workspace& ws, int parity, hipStream_t s);
  // the deferred x += alpha p of iteration `launched`; gated: only if that iteration has set
  // the stop flag by the time the update runs (a no-op otherwise: the next iteration's
  // direction update applies it)
  template <typename T>
  int pcg_enqueue_end(Workspace& ws, int launched, hipStream_t s, bool gated = false);
  // Reliable-update PCG on A x = b from x0 = 0 (van der Vorst and Ye 2000, residual
  // replacement): Krylov vectors, operator and V-cycle in FP32, the iterate accumulated in
  // FP64 (x64_), and the recursive FP32 residual replaced by the true FP64 one b - A x64
  // every ru_every_ iterations until it nears the tolerance; the search direction is kept,
  // unlike a refinement restart.
  int pcg_ru(Workspace& ws, const double* b, double rtol, int max_iter, double* bnorm2,
             int* iters, double* rel, hipStream_t s);
  int ru_every_ = 4;  // iterations between replacements
  // preconditioned MINRES (FP64, FP64 V-cycle) on A x = b from x0 = 0, restarted on the true
  // residual until |b - A x| <= rtol |b|; result in x_buf<double>().  For operators that are
  // not positive definite (PCG breakdown).
  int minres(Workspace& ws, const double* b, double rtol, int max_iter, double bnorm2,
             int* iters, double* rel, hipStream_t s);
  int solve(Workspace& ws, const double* b, double* x, const SolveParams& p, SolveStats* st,
            hipStream_t s);

  size_t n() const { return (size_t)bs_ * levels_[0].nn; }
  int levels() const { return (int)levels_.size() + 1; }
  void set_rows(int rows64, int rows32);
  int rows64() const { return levels_.empty() ? 0 : levels_[0].sh64.rows; }
  int rows32() const { return levels_.empty() ? 0 : levels_[0].sh32.rows; }

  Phys phys_ = kElasticity;
  int bs_ = 2;
  unsigned fixed_ = 0;
  double lam_ = 0, mu_ = 0, eps_ = 0;
  int p_ = 0;
  double m_ = 0, cp_ = 0;
  double omega_ = 0.7;
  // Chebyshev step sizes: with cheb_lmax_ > 0 the nu Jacobi sweeps of one smoothing pass use
  // omega_k = 1 / theta_k, theta_k the roots of the degree-nu Chebyshev polynomial on
  // [cheb_lmax_ / cheb_ratio_, cheb_lmax_] (an upper bound of lambda_max(D^-1 A)), so one pass
  // damps that whole band; the pass polynomial has commuting factors, so pre- and
  // post-smoothing stay adjoint and the V-cycle symmetric.  0: damped Jacobi with omega_.
  double cheb_lmax_ = 0.0, cheb_ratio_ = 8.0;
  double sweep_omega(int k, int nu) const;
  // power-iteration estimate of max over levels of lambda_max(D^-1 A) (preconditioner
  // operators, FP64); sets cheb_lmax_ to it times the margin
  int estimate_lmax(Workspace& ws, hipStream_t s, double margin = 1.1);
  bool cheb_auto_ = true;     // estimate once, at the first operator
  bool cheb_done_ = false;
  int nu_pre_ = 2, nu_post_ = 2;  // Jacobi sweeps on the finest level
  int nu_coarse_ = 2;              // pre- and post-sweeps on every coarser level
  bool fused_ = true, use_tail_ = true;
  // tiled multi-sweep kernels (one launch per half-visit of a level): on levels >= band_first_
  // (FP32 V-cycles) or >= band_first_fp64_ (FP64) above the tail, and on the filter's finest
  // level when band_fine_filter_ is set; a first level of 0 switches them off
  int band_first_ = 1, band_first_fp64_ = 1;
  // PCG iterations replayed as captured graphs, or launched one by one (the default: on
  // ROCm 7.2 / MI355X direct launches measured 2-3% faster per solve and 5% faster in batches,
  // the host keeping an iteration ahead either way)
  bool use_graphs_ = false;
  bool band_fine_filter_ = true;
  // elasticity cell-weight band levels use the 1024-thread one-node-per-thread kernels
  // the elasticity's cell-weight band levels: 0: 256-thread workgroups; 1: 1024 threads, one
  // node each (square tiles); 2 / 3: 512 / 1024 threads on the standard tiles.  2 (the
  // default) measured 0.15 ms faster per mixed step at N=1000 than 0, 3 slower than both.
  int band_wide_ = 2;
  // > 0: the finest elasticity level's first two pre-sweeps (pair_rows_), and its prolongation
  // with the first two post-sweeps (pair_rows_post_), run as one launch each with node-row
  // chunks of that height; 0: one launch per sweep.  Index 0: FP32 V-cycles, 1: FP64.
  int pair_rows_[2] = {8, 8}, pair_rows_post_[2] = {12, 16};
  // > 0: with two pre-sweeps, the zero-start pair also forms the residual and restricts it
  // (chunks of this many rows); 0: a separate residual + restriction launch
  int pair_rr_rows_[2] = {12, 12};
  // elasticity element levels 0 .. pair_levels_ - 1 use the paired column-strip kernels (the
  // coarser ones with chunk heights pair_rows_coarse_ [pre, post, fused restriction] and
  // pair_nu_coarse_ sweeps each way, 0: nu_coarse_); the levels below them the band kernels
  int pair_levels_ = 1;
  int pair_rows_coarse_[3] = {8, 8, 8};
  int pair_nu_coarse_ = 0;
  // > 0: the filter's finest level runs as two column pipelines (tmx_kernels.h
  // filter_pipe_down / _up) with node-row chunks of these heights ([FP32, FP64][down, up]);
  // 0: band kernels
  int fpipe_rows_[2][2] = {{12, 12}, {12, 12}};
  // filter preconditioner: the exact DCT-I / tridiagonal solver of the tensor-lumped
  // operator (tmx_dct.h) instead of the V-cycle, where the grid allows it
  bool use_dct_ = true;
  std::unique_ptr<DctPrecond<float>> dct32_;
  std::unique_ptr<DctPrecond<double>> dct64_;
  template <typename T>
  std::unique_ptr<DctPrecond<T>>& dct_slot() {
    if constexpr (sizeof(T) == 8) return dct64_;
    else return dct32_;
  }
  hipEvent_t solve_ev_[2] = {nullptr, nullptr};  // around an unsynchronised solve
  // side stream of the overlapped coarse-operator build (default priority), the events that
  // fork it from and join it to the solve's stream, and the first level whose operator it
  // builds
  hipStream_t side_ = nullptr;
  hipEvent_t fork_ev_ = nullptr, join_ev_ = nullptr;
  bool coarse_pending_ = false;
  int coarse_first_ = 0;
  // build the coarse operators on the side stream (TMX_OPT_SETUP_OVERLAP)
  bool overlap_setup_ = true;
  // pcg(): queue the awaited iteration's x update gated on its stop flag when nothing is
  // queued behind it
  bool gate_end_ = true;
  // In-step timing of the PCG direction update + operator (apply_update): with au_timing_ on,
  // each launch inside a PCG iteration is bracketed by a pair of events on the solve's stream
  // (up to kAuEvents launches between two reads; direct launches only, not under graphs)
  static constexpr int kAuEvents = 256;
  bool au_timing_ = false;
  std::vector<hipEvent_t> au_ev_;
  int au_used_ = 0;
  int au_record(int which, hipStream_t s);  // which: 0 before, 1 after the launch
  // sum of the recorded launches' durations (us) and their count since the last read
  int au_read(double* total_us, int* launches);
  bool pair_post_ok(int l = 0) const;  // level l has even sides and a nested level below it
  int tail_nodes_ = kTailNodes;  // largest 9-point level handled by the tail kernel
  void find_tail();
  int tail_ = -1;               // first level handled by the one-workgroup tail kernel
  std::vector<Level> levels_;   // levels_[0] = finest
  // coarsest dense level
  int cnx_ = 0, cny_ = 0, cn_ = 0;
  DevBuf<double> dense_;
  DevBuf<double> cst64_;        // coarsest stencil (fp64)
  DevBuf<double> tmpst_;        // explicit stencil of the last element level
  DevBuf<double> cw0_;          // level-0 cell weights (restriction source)
  DevBuf<float> cw0_32_;        // the same in FP32
  // Finest-level V-cycle sweeps read the preconditioner's per-cell weights (cw0_ / cw0_32_,
  // one pair per cell, formed once per operator) instead of re-forming them from the nodal xi
  // in every sweep; bit 0: FP32 V-cycles, bit 1: FP64.  has_cw0_: the weights match the
  // current operator.
  unsigned fine_cw_ = 1u;
  bool has_cw0_ = false, graphs_cw0_ = false;  // graphs_cw0_: has_cw0_ when graphs were captured
  const double* xi64_ = nullptr;  // = xi64buf_ once an operator is set
  DevBuf<double> xi64buf_;
  DevBuf<float> xi32_;
  bool has_operator_ = false;
  // Krylov vectors
  DevBuf<double> x64_, r64_, z64_, p64_, q64_, pb64_;
  double* xit_ = nullptr;  // the FP64 PCG iterate's buffer while solve() iterates in its output
  DevBuf<float> r32_, z32_, p32_, q32_, d32_, pb32_;
  DevBuf<double> rr64_;         // refinement residual
  DevBuf<double> mrw_[3];       // MINRES direction vectors (elasticity only)
  // captured PCG graphs: [fp64 / fp32 / fp64 with fp32 V-cycle][0 = init, 1 = even
  // iteration, 2 = odd iteration]
  hipGraphExec_t graphs_[3][3] = {};
  void drop_graphs();
  ~Hierarchy();

 private:
  int alloc_vectors();
  int build_coarse_stencils(hipStream_t s);
};

}  // namespace tmx
