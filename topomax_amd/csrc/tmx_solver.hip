// Geometric multigrid hierarchy, PCG and mixed-precision iterative refinement.
//
// Levels: the fine grid, then nested P1 coarsenings (nx, ny even) that keep the element form
// (the Galerkin coarse operator of nested P1 spaces is the coarse-mesh operator whose
// triangle factor is the mean of the 4 fine triangles it contains), then tensor-product
// linear coarsenings whose Galerkin operators are explicit 9-point block stencils, down to a
// coarsest grid solved by a dense inverse.  Smoother: damped block Jacobi.
#include "tmx_solver.h"
#include "tmx_dct.h"
#include "tmx_band.h"
#include "tmx_multi.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <type_traits>

namespace tmx {

static thread_local std::string g_last_error;
static const bool g_trace = std::getenv("TMX_TRACE") != nullptr;
void set_error(const std::string& m) { g_last_error = m; }
const char* last_error_cstr() { return g_last_error.c_str(); }

// ------------------------------------------------------------------ workspace
int Workspace::init(hipStream_t s) {
  TMX_TRY(partials.alloc((size_t)kRegions * kMaxPartials));
  TMX_TRY(scalars.alloc(kScalars));
  TMX_HIP(hipHostMalloc(&host, sizeof(double) * 2 * kScalars, hipHostMallocMapped));
  std::memset(host, 0, sizeof(double) * 2 * kScalars);
  TMX_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&host_dev), host, 0));
  TMX_HIP(fill_zero_bytes(scalars.get(), sizeof(double) * kScalars, s));
  return 0;
}

Workspace::~Workspace() {
  if (host) (void)hipHostFree(host);
}

hipError_t stream_wait(hipStream_t s) {
  for (;;) {
    const hipError_t q = hipStreamQuery(s);
    if (q != hipErrorNotReady) return q;
  }
}

int Workspace::read(hipStream_t s, int first, int count) {
  // a one-thread kernel per scalar writes the mapped pinned copy: cheaper to queue than a
  // device-to-host copy, whose host-side cost the device would otherwise sit idle through
  for (int k = 0; k < count; ++k)
    TMX_HIP(publish_scalar(scalars.get() + first + k, host_dev + first + k, s));
  TMX_HIP(stream_wait(s));
  return 0;
}

// ------------------------------------------------------------------ hierarchy setup
// node rows per wavefront chunk: enough wavefronts (~4096) to fill 256 CUs, 2..16 rows
static int rows_default(int nx, int ny) {
  const long strips = (nx + 1 + 61) / 62;  // 62 output columns per wavefront
  const long r = (strips * (ny + 1)) / 4096;
  return (int)std::max(2L, std::min(16L, r));
}

int Hierarchy::init(Phys phys, int nx, int ny, double hx, double hy, unsigned fixed, double lam,
                    double mu, double eps, int coarse_max_nodes, hipStream_t s) {
  phys_ = phys;
  bs_ = phys == kElasticity ? 2 : 1;
  fixed_ = fixed;
  lam_ = lam;
  mu_ = mu;
  eps_ = eps;
  omega_ = phys == kElasticity ? 0.7 : 0.8;
  levels_.clear();
  Level L0;
  L0.nx = nx; L0.ny = ny; L0.hx = hx; L0.hy = hy; L0.kind = kElem;
  L0.nn = (size_t)(nx + 1) * (ny + 1);
  levels_.push_back(std::move(L0));
  while (levels_.back().nn > (size_t)coarse_max_nodes) {
    const Level& c = levels_.back();
    Level n;
    if (c.kind == kElem && c.nx % 2 == 0 && c.ny % 2 == 0) {
      n.nx = c.nx / 2; n.ny = c.ny / 2; n.kind = kElem; n.nested_from_prev = 1;
    } else {
      n.nx = c.nx == 1 ? 1 : (c.nx + 1) / 2;
      n.ny = c.ny == 1 ? 1 : (c.ny + 1) / 2;
      n.kind = kStencil;
      n.nested_from_prev = 0;
      if (n.nx == c.nx && n.ny == c.ny) break;
    }
    n.hx = c.hx * c.nx / n.nx;
    n.hy = c.hy * c.ny / n.ny;
    n.nn = (size_t)(n.nx + 1) * (n.ny + 1);
    levels_.push_back(std::move(n));
  }
  for (auto& L : levels_) {
    L.sh64 = SweepShape::make(L.nx, L.ny, rows_default(L.nx, L.ny));
    L.sh32 = SweepShape::make(L.nx, L.ny, rows_default(L.nx, L.ny));
  }
  find_tail();
  if (phys == kElasticity) {  // the overlapped coarse build's stream and events, made up front
                              // so that no solve's setup pays for their creation
    if (!side_) TMX_HIP(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    if (!fork_ev_) TMX_HIP(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
    if (!join_ev_) TMX_HIP(hipEventCreateWithFlags(&join_ev_, hipEventDisableTiming));
  }
  Level& last = levels_.back();
  cnx_ = last.nx;
  cny_ = last.ny;
  cn_ = bs_ * (int)last.nn;
  TMX_TRY(dense_.alloc((size_t)cn_ * cn_));
  TMX_TRY(alloc_vectors());
  (void)s;
  return 0;
}

// tail: the longest suffix of small 9-point levels ending at the dense coarsest whose
// vectors fit the tail kernel's LDS budget (sized for FP64, so both precisions fit)
void Hierarchy::find_tail() {
  const int lastl = (int)levels_.size() - 1;
  tail_ = -1;
  if (lastl >= 1 && bs_ * levels_[lastl].nn <= 128) {
    const size_t cap = (size_t)std::min(tail_nodes_, kTailThreads);
    int t = lastl;
    while (t - 1 >= 1 && lastl - (t - 1) + 1 <= kMaxTail && levels_[t - 1].kind == kStencil &&
           levels_[t - 1].nn <= cap && !levels_[t].nested_from_prev) {
      size_t nn[kMaxTail];
      for (int l = t - 1; l <= lastl; ++l) nn[l - (t - 1)] = levels_[l].nn;
      if (tail_lds_bytes(nn, lastl - (t - 1) + 1, bs_, sizeof(double)) > kTailLdsBytes) break;
      --t;
    }
    tail_ = t;
  }
}

double Hierarchy::sweep_omega(int k, int nu) const {
  if (!(cheb_lmax_ > 0.0) || nu < 1) return omega_;
  const double b = cheb_lmax_, a = cheb_lmax_ / cheb_ratio_;
  const double theta = 0.5 * (a + b) + 0.5 * (b - a) * std::cos((2 * k + 1) * M_PI / (2.0 * nu));
  return 1.0 / theta;
}

int Hierarchy::au_record(int which, hipStream_t s) {
  if (au_ev_.empty()) {
    au_ev_.assign(2 * kAuEvents, nullptr);
    for (auto& e : au_ev_) TMX_HIP(hipEventCreate(&e));
  }
  if (au_used_ >= kAuEvents) return 0;  // full: later launches are not timed until a read
  TMX_HIP(hipEventRecord(au_ev_[2 * au_used_ + which], s));
  if (which == 1) ++au_used_;
  return 0;
}

int Hierarchy::au_read(double* total_us, int* launches) {
  double t = 0.0;
  for (int i = 0; i < au_used_; ++i) {
    TMX_HIP(hipEventSynchronize(au_ev_[2 * i + 1]));
    float ms = 0.f;
    TMX_HIP(hipEventElapsedTime(&ms, au_ev_[2 * i], au_ev_[2 * i + 1]));
    t += 1000.0 * ms;
  }
  *total_us = t;
  *launches = au_used_;
  au_used_ = 0;
  return 0;
}

void Hierarchy::drop_graphs() {
  for (auto& g : graphs_)
    for (auto& e : g)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
}

Hierarchy::~Hierarchy() {
  if (side_) (void)hipStreamSynchronize(side_);
  drop_graphs();
  if (fork_ev_) (void)hipEventDestroy(fork_ev_);
  if (join_ev_) (void)hipEventDestroy(join_ev_);
  if (side_) (void)hipStreamDestroy(side_);
  for (auto& e : au_ev_)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : solve_ev_)
    if (e) (void)hipEventDestroy(e);
}

void SolveStats::complete() {
  if (pending_rr) {
    rel_residual = pending_bb > 0.0 ? std::sqrt(std::max(*pending_rr, 0.0) / pending_bb) : 0.0;
    pending_rr = nullptr;
  }
  if (pending_ev[0] && pending_ev[1]) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pending_ev[0], pending_ev[1]) == hipSuccess) solve_ms = ms;
    pending_ev[0] = pending_ev[1] = nullptr;
  }
}

void Hierarchy::set_rows(int rows64, int rows32) {
  drop_graphs();
  if (rows64 > 0) levels_[0].sh64 = SweepShape::make(levels_[0].nx, levels_[0].ny, rows64);
  if (rows32 > 0) levels_[0].sh32 = SweepShape::make(levels_[0].nx, levels_[0].ny, rows32);
}

int Hierarchy::alloc_vectors() {
  const size_t bs = bs_;
  size_t tmp_nodes = 0;
  for (size_t l = 0; l < levels_.size(); ++l) {
    Level& L = levels_[l];
    const size_t n = bs * L.nn;
    TMX_TRY(L.t64.alloc(n));
    TMX_TRY(L.r64.alloc(n));
    TMX_TRY(L.t32.alloc(n));
    TMX_TRY(L.r32.alloc(n));
    TMX_TRY(L.b64.alloc(n));
    TMX_TRY(L.x64.alloc(n));
    TMX_TRY(L.b32.alloc(n));
    TMX_TRY(L.x32.alloc(n));
    if (l > 0 && L.kind == kElem && phys_ == kElasticity) {
      TMX_TRY(L.cw64.alloc(2 * (size_t)L.nx * L.ny));
      TMX_TRY(L.cw32.alloc(2 * (size_t)L.nx * L.ny));
    }
    const bool last = (l + 1 == levels_.size());
    if (L.kind == kStencil || last) {
      TMX_TRY(L.st64.alloc(9 * bs * bs * L.nn));
      TMX_TRY(L.st32.alloc(9 * bs * bs * L.nn));
    }
    // element level followed by a stencil level needs its explicit stencil for the RAP
    if (L.kind == kElem && !last && levels_[l + 1].kind == kStencil)
      tmp_nodes = std::max(tmp_nodes, L.nn);
  }
  if (tmp_nodes) TMX_TRY(tmpst_.alloc(9 * bs * bs * tmp_nodes));
  if (phys_ == kElasticity) {
    TMX_TRY(cw0_.alloc(2 * (size_t)levels_[0].nx * levels_[0].ny));
    TMX_TRY(cw0_32_.alloc(2 * (size_t)levels_[0].nx * levels_[0].ny));
    TMX_TRY(xi32_.alloc(levels_[0].nn));
    TMX_TRY(xi64buf_.alloc(levels_[0].nn));
    xi64_ = xi64buf_.get();
  }
  const size_t n0 = n();
  TMX_TRY(x64_.alloc(n0));
  TMX_TRY(r64_.alloc(n0));
  TMX_TRY(z64_.alloc(n0));
  TMX_TRY(p64_.alloc(n0));
  TMX_TRY(q64_.alloc(n0));
  TMX_TRY(pb64_.alloc(n0));
  TMX_TRY(pb32_.alloc(n0));
  TMX_TRY(r32_.alloc(n0));
  TMX_TRY(z32_.alloc(n0));
  TMX_TRY(p32_.alloc(n0));
  TMX_TRY(q32_.alloc(n0));
  TMX_TRY(d32_.alloc(n0));
  TMX_TRY(rr64_.alloc(n0));
  if (phys_ == kElasticity)
    for (auto& w : mrw_) TMX_TRY(w.alloc(n0));
  return 0;
}

// Operator arguments of element level l.  The preconditioner's operators (V-cycle sweeps,
// transfers and the coarse hierarchy) use |factor| raised to at least m, so they stay SPD
// when a filtered design dips below zero and an odd penalty makes some factors negative (and
// the preconditioned operator then has eigenvalues near +-1 there); the operator that PCG /
// MINRES solve with (apply, residual) keeps the factors as they are.
template <typename T>
static ElastOpArgs<T> elast_args(const Hierarchy& h, int l, const T* xi, const T* cw,
                                 bool fp64, bool precond) {
  const Level& L = h.levels_[l];
  ElastOpArgs<T> a;
  a.sh = fp64 ? L.sh64 : L.sh32;
  a.fixed = h.fixed_;
  a.xi = l == 0 ? xi : nullptr;
  a.cellw = l == 0 ? nullptr : cw;
  if (l == 0 && precond && h.has_cw0_ && (h.fine_cw_ & (fp64 ? 2u : 1u))) {
    a.xi = nullptr;  // the preconditioner's weights of the finest cells, formed once
    a.cellw = fp64 ? reinterpret_cast<const T*>(h.cw0_.get())
                   : reinterpret_cast<const T*>(h.cw0_32_.get());
  }
  a.p = h.p_;
  a.m = (T)h.m_;
  a.cp = (T)h.cp_;
  a.wmin = precond ? (T)h.m_ : -std::numeric_limits<T>::max();
  a.k = ElastConst<T>::make(h.lam_, h.mu_, L.hx, L.hy);
  return a;
}

// explicit stencil of element level l (fp64) into dst
static int element_stencil(const Hierarchy& h, int l, double* dst, hipStream_t s) {
  const Level& L = h.levels_[l];
  if (h.phys_ == kElasticity) {
    ElastOpArgs<double> a = elast_args<double>(h, l, h.xi64_, L.cw64.get(), true, true);
    TMX_HIP(elast_to_stencil<double>(a, dst, s));
  } else {
    TMX_HIP(filter_to_stencil<double>(L.nx, L.ny, FilterConst<double>::make(h.eps_, L.hx, L.hy),
                                      dst, s));
  }
  return 0;
}

int Hierarchy::build_coarse_stencils(hipStream_t s) {
  const int last = (int)levels_.size() - 1;
  for (int l = 1; l <= last; ++l) {
    Level& L = levels_[l];
    Level& P = levels_[l - 1];
    if (L.kind == kStencil) {
      const double* src = nullptr;
      if (P.kind == kElem) {
        TMX_TRY(element_stencil(*this, l - 1, tmpst_.get(), s));
        src = tmpst_.get();
      } else {
        src = P.st64.get();
      }
      if (bs_ == 2)
        TMX_HIP(stencil_rap<2>(P.nx, P.ny, L.nx, L.ny, 0, fixed_, src, L.st64.get(), s,
                               L.st32.get()));
      else
        TMX_HIP(stencil_rap<1>(P.nx, P.ny, L.nx, L.ny, 0, fixed_, src, L.st64.get(), s,
                               L.st32.get()));
    } else if (l == last) {
      TMX_TRY(element_stencil(*this, l, L.st64.get(), s));
    }
  }
  if (last == 0) TMX_TRY(element_stencil(*this, 0, levels_[0].st64.get(), s));
  const Level& C = levels_[last];
  if (bs_ == 2)
    TMX_HIP(stencil_to_dense<2>(C.nx, C.ny, C.st64.get(), dense_.get(), s));
  else
    TMX_HIP(stencil_to_dense<1>(C.nx, C.ny, C.st64.get(), dense_.get(), s));
  TMX_HIP(dense_invert(cn_, dense_.get(), nullptr, s));
  return 0;
}

int Hierarchy::set_fine_operator(const double* xi, int p, double m, hipStream_t s) {
  if (p != p_ || m != m_) drop_graphs();  // p, m are launch arguments of captured kernels
  p_ = p;
  m_ = m;
  cp_ = (1.0 - m) * 2.0 / ((p + 1.0) * (p + 2.0));
  has_cw0_ = false;
  const Level& L0 = levels_[0];
  if (xi != xi64buf_.get())
    TMX_HIP(copy_split(L0.nn, xi, xi64buf_.get(), xi32_.get(), s));
  else
    TMX_HIP((convert<double, float>(L0.nn, xi64buf_.get(), xi32_.get(), s)));
  has_operator_ = true;
  return 0;
}

int Hierarchy::fine_cell_weights(hipStream_t s) {
  const Level& L0 = levels_[0];
  TMX_HIP(elast_cell_weights<double>(L0.nx, L0.ny, xi64_, p_, m_, cp_, m_, cw0_.get(), s,
                                     cw0_32_.get()));
  has_cw0_ = true;
  return 0;
}

bool Hierarchy::pair_post_ok(int l) const {
  return l >= 0 && (int)levels_.size() >= l + 2 && levels_[l + 1].nested_from_prev &&
         !(levels_[l].nx & 1) && !(levels_[l].ny & 1);
}

int Hierarchy::update_elasticity(const double* xi, int p, double m, hipStream_t s,
                                 bool overlap) {
  TMX_TRY(join_coarse(s));  // (the previous build reads the cell weights rewritten below)
  TMX_TRY(set_fine_operator(xi, p, m, s));
  const int last = (int)levels_.size() - 1;
  TMX_TRY(fine_cell_weights(s));
  int first = last;  // the first level that build_coarse_stencils writes
  for (int l = 1; l <= last; ++l) {
    Level& L = levels_[l];
    if (L.kind != kElem) {
      first = l;
      break;
    }
    const double* src = l == 1 ? cw0_.get() : levels_[l - 1].cw64.get();
    TMX_HIP(elast_restrict_weights<double>(L.nx, L.ny, src, L.cw64.get(), s, L.cw32.get()));
  }
  // The side stream runs the build while s runs the finest and cell-weight levels of the
  // first V-cycle.  Not while launches are recorded for a merged batch or captured into
  // graphs, where a second stream would escape them.
  if (overlap && overlap_setup_ && !use_graphs_ && !active_recorder() && first >= 1) {
    if (!side_) TMX_HIP(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    if (!fork_ev_) TMX_HIP(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
    if (!join_ev_) TMX_HIP(hipEventCreateWithFlags(&join_ev_, hipEventDisableTiming));
    TMX_HIP(hipEventRecord(fork_ev_, s));
    TMX_HIP(hipStreamWaitEvent(side_, fork_ev_, 0));
    const int rc = build_coarse_stencils(side_);
    TMX_HIP(hipEventRecord(join_ev_, side_));  // (joined also when a launch failed)
    coarse_pending_ = true;
    coarse_first_ = first;
    return rc;
  }
  return build_coarse_stencils(s);
}

int Hierarchy::join_coarse(hipStream_t s) {
  if (!coarse_pending_) return 0;
  coarse_pending_ = false;
  TMX_HIP(hipStreamWaitEvent(s, join_ev_, 0));
  return 0;
}

int Hierarchy::build_filter(hipStream_t s) { return build_coarse_stencils(s); }



// ------------------------------------------------------------------ level operators
template <typename T>
struct PrecBufs;
template <>
struct PrecBufs<double> {
  static double* t(Level& L) { return L.t64.get(); }
  static double* r(Level& L) { return L.r64.get(); }
  static double* b(Level& L) { return L.b64.get(); }
  static double* x(Level& L) { return L.x64.get(); }
  static const double* cw(const Level& L) { return L.cw64.get(); }
  static const double* st(const Level& L) { return L.st64.get(); }
  static const SweepShape& sh(const Level& L) { return L.sh64; }
};
template <>
struct PrecBufs<float> {
  static float* t(Level& L) { return L.t32.get(); }
  static float* r(Level& L) { return L.r32.get(); }
  static float* b(Level& L) { return L.b32.get(); }
  static float* x(Level& L) { return L.x32.get(); }
  static const float* cw(const Level& L) { return L.cw32.get(); }
  static const float* st(const Level& L) { return L.st32.get(); }
  static const SweepShape& sh(const Level& L) { return L.sh32; }
};

template <typename T>
static const T* xi_for(const Hierarchy& h);
template <>
const double* xi_for<double>(const Hierarchy& h) { return h.xi64_; }
template <>
const float* xi_for<float>(const Hierarchy& h) { return h.xi32_.get(); }

template <typename T>
static int level_op(Hierarchy& h, int l, int op, const T* x, const T* b, T* y, RedSlot* dot,
                    hipStream_t s, const SweepExtra<T>& ex = SweepExtra<T>(),
                    bool precond = true,
                    double omega = std::numeric_limits<double>::quiet_NaN()) {
  Level& L = h.levels_[l];
  const bool fp64 = sizeof(T) == 8;
  const T om = (T)(std::isnan(omega) ? h.omega_ : omega);  // NaN: the hierarchy's damping
  if (L.kind == kElem) {
    if (h.phys_ == kElasticity) {
      ElastOpArgs<T> a = elast_args<T>(h, l, xi_for<T>(h), PrecBufs<T>::cw(L), fp64, precond);
      TMX_HIP(elast_sweep<T>(a, op, x, b, y, om, dot, s, ex));
    } else {
      TMX_HIP(filter_sweep<T>(PrecBufs<T>::sh(L), FilterConst<T>::make(h.eps_, L.hx, L.hy), op, x,
                              b, y, om, dot, s, ex));
    }
  } else {
    if (dot || op > kSmoothZero) {
      set_error("internal: 9-point levels have no fused dot / transfer ops");
      return -1;
    }
    if (h.bs_ == 2)
      TMX_HIP((stencil_op<T, 2>(L.nx, L.ny, PrecBufs<T>::st(L), op, x, b, y, om, s)));
    else
      TMX_HIP((stencil_op<T, 1>(L.nx, L.ny, PrecBufs<T>::st(L), op, x, b, y, om, s)));
  }
  return 0;
}

template <typename T>
int Hierarchy::apply(const T* x, T* y, RedSlot* dot, hipStream_t s) {
  return level_op<T>(*this, 0, kApply, x, nullptr, y, dot, s, SweepExtra<T>(), false);
}

template <typename T>
int Hierarchy::residual(const T* x, const T* b, T* r, hipStream_t s) {
  return level_op<T>(*this, 0, kResidual, x, b, r, nullptr, s, SweepExtra<T>(), false);
}

template <typename T>
int Hierarchy::apply_update(const T* p_old, const T* z, T* q, T* p_new, const double* beta,
                            RedSlot* dot, hipStream_t s, const float* zlow, T* xk,
                            const double* alpha) {
  SweepExtra<T> ex;
  ex.pnew = p_new;
  ex.beta = beta;
  ex.zlow = zlow;
  ex.xk = xk;
  ex.alpha = alpha;
  return level_op<T>(*this, 0, kApplyUpdate, p_old, zlow ? nullptr : z, q, dot, s, ex, false);
}

template <typename T>
int Hierarchy::smooth0(const T* x, const T* b, T* y, hipStream_t s) {
  return level_op<T>(*this, 0, kSmooth, x, b, y, nullptr, s);
}

template <typename T>
int Hierarchy::sweep_pair0(int op1, int rows, const T* x, const T* b, T* y, double om1, double om2,
                           RedSlot* dot, hipStream_t s, int l) {
  if (phys_ != kElasticity || l < 0 || l >= (int)levels_.size() || levels_[l].kind != kElem ||
      rows < 1) {
    set_error("internal: paired sweeps need an elasticity element level");
    return -1;
  }
  const T* ec = nullptr;
  if (op1 == kProlongSmooth) {
    if (!pair_post_ok(l)) {
      set_error("internal: paired prolongation needs a nested level below");
      return -1;
    }
    ec = PrecBufs<T>::x(levels_[l + 1]);
  }
  const ElastOpArgs<T> a = elast_args<T>(*this, l, xi_for<T>(*this), PrecBufs<T>::cw(levels_[l]),
                                         sizeof(T) == 8, true);
  TMX_HIP(elast_sweep_pair<T>(a, op1, rows, x, b, y, (T)om1, (T)om2, dot, s, ec));
  return 0;
}

template <typename T>
int Hierarchy::filter_pipe0(bool down, int rows, hipStream_t s) {
  if (phys_ != kFilter || levels_.size() < 2 || !levels_[1].nested_from_prev) {
    set_error("internal: the filter pipelines need a nested level below the finest");
    return -1;
  }
  Level& L = levels_[0];
  Level& C = levels_[1];
  const int nu = down ? nu_pre_ : nu_post_;
  if (nu > kMaxPipeNu) {
    set_error("internal: too many sweeps for the filter pipelines");
    return -1;
  }
  T om[kMaxPipeNu];
  for (int k = 0; k < nu; ++k) om[k] = (T)sweep_omega(k, nu);
  const FilterConst<T> fk = FilterConst<T>::make(eps_, L.hx, L.hy);
  if (down)
    TMX_HIP(filter_pipe_down<T>(L.nx, L.ny, fk, nu, om, rows, PrecBufs<T>::b(L),
                                PrecBufs<T>::t(L), PrecBufs<T>::b(C), s));
  else
    TMX_HIP(filter_pipe_up<T>(L.nx, L.ny, fk, nu, om, rows, PrecBufs<T>::t(L), PrecBufs<T>::x(C),
                              PrecBufs<T>::b(L), PrecBufs<T>::x(L), nullptr, s));
  return 0;
}

template <typename T>
int Hierarchy::sweep_pair_rr0(int rows, const T* b, T* y, double om1, double om2, hipStream_t s,
                              int l) {
  if (phys_ != kElasticity || !pair_post_ok(l) || levels_[l].kind != kElem || rows < 1) {
    set_error("internal: fused restriction needs an elasticity element level over a nested one");
    return -1;
  }
  const ElastOpArgs<T> a = elast_args<T>(*this, l, xi_for<T>(*this), PrecBufs<T>::cw(levels_[l]),
                                         sizeof(T) == 8, true);
  TMX_HIP(elast_sweep_pair_rr<T>(a, rows, b, y, PrecBufs<T>::b(levels_[l + 1]), (T)om1, (T)om2,
                                 s));
  return 0;
}

template <typename T>
static int restrict_to(Hierarchy& h, int l, const T* rf, T* bc, hipStream_t s) {
  const Level& F = h.levels_[l];
  const Level& C = h.levels_[l + 1];
  if (C.nested_from_prev) {
    if (h.bs_ == 2) TMX_HIP((restrict_nested<T, 2>(C.nx, C.ny, h.fixed_, rf, bc, s)));
    else TMX_HIP((restrict_nested<T, 1>(C.nx, C.ny, h.fixed_, rf, bc, s)));
  } else {
    if (h.bs_ == 2) TMX_HIP((restrict_bilinear<T, 2>(F.nx, F.ny, C.nx, C.ny, h.fixed_, rf, bc, s)));
    else TMX_HIP((restrict_bilinear<T, 1>(F.nx, F.ny, C.nx, C.ny, h.fixed_, rf, bc, s)));
  }
  return 0;
}

template <typename T>
static int prolong_to(Hierarchy& h, int l, const T* ec, T* xf, hipStream_t s) {
  const Level& F = h.levels_[l];
  const Level& C = h.levels_[l + 1];
  if (C.nested_from_prev) {
    if (h.bs_ == 2) TMX_HIP((prolong_nested<T, 2>(C.nx, C.ny, h.fixed_, ec, xf, s)));
    else TMX_HIP((prolong_nested<T, 1>(C.nx, C.ny, h.fixed_, ec, xf, s)));
  } else {
    if (h.bs_ == 2) TMX_HIP((prolong_bilinear<T, 2>(F.nx, F.ny, C.nx, C.ny, h.fixed_, ec, xf, s)));
    else TMX_HIP((prolong_bilinear<T, 1>(F.nx, F.ny, C.nx, C.ny, h.fixed_, ec, xf, s)));
  }
  return 0;
}

template <typename T>
static int tail_launch(Hierarchy& h, const T* b, T* x, hipStream_t s) {
  const int last = (int)h.levels_.size() - 1;
  TailDesc<T> d;
  std::memset(&d, 0, sizeof(d));
  d.nlev = last - h.tail_ + 1;
  d.cn = h.cn_;
  d.nu_pre = h.nu_coarse_;
  d.nu_post = h.nu_coarse_;
  d.fixed = h.fixed_;
  for (int k = 0; k < kMaxNu; ++k) d.omega[k] = (T)h.sweep_omega(k, h.nu_coarse_);
  d.dense = h.dense_.get();
  d.b_in = b;
  d.x_out = x;
  size_t nn[kMaxTail];
  int off = 0;
  for (int l = h.tail_; l <= last; ++l) {
    Level& L = h.levels_[l];
    TailLevel<T>& t = d.lev[l - h.tail_];
    t.nx = L.nx;
    t.ny = L.ny;
    t.st = PrecBufs<T>::st(L);
    t.off = off;
    off += (l < last ? 3 : 2) * h.bs_ * (int)L.nn;
    nn[l - h.tail_] = L.nn;
  }
  const size_t lds = tail_lds_bytes(nn, d.nlev, h.bs_, sizeof(T));
  if (h.bs_ == 2) TMX_HIP((tail_vcycle<T, 2>(d, lds, s)));
  else TMX_HIP((tail_vcycle<T, 1>(d, lds, s)));
  return 0;
}

// Band-kernel description of level l (below it level l + 1) with the largest tile that fits
// the LDS budget and still gives about two workgroups per CU; small levels take the smallest
// tile, where latency rather than the halo's extra work decides.  False when the level runs
// launch per sweep instead.
template <typename T>
static bool band_level(const Hierarchy& h, int l, BandLevel<T>& B) {
  const Level& L = h.levels_[l];
  const Level& C = h.levels_[l + 1];
  const int first = sizeof(T) == 8 ? h.band_first_fp64_ : h.band_first_;
  if (first < 1) return false;
  if (l == 0 ? !(h.phys_ == kFilter && h.band_fine_filter_) : l < first) return false;
  std::memset(&B, 0, sizeof(B));
  B.nx = L.nx;
  B.ny = L.ny;
  B.nxc = C.nx;
  B.nyc = C.ny;
  B.nested = C.nested_from_prev;
  B.fixed = h.fixed_;
  if (L.kind == kElem) {
    if (h.phys_ == kElasticity) {
      if (l == 0) return false;  // the finest elasticity level carries the nodal SIMP factor
      B.src = kSrcCells;
      B.cw = PrecBufs<T>::cw(L);
      band_elastic_table<T>(B, L.hx, L.hy, h.lam_, h.mu_);
    } else {
      B.src = kSrcUnit;
      band_filter_table<T>(B, L.hx, L.hy, h.eps_, h.bs_);
    }
  } else {
    B.src = kSrcStencil;
    B.st = PrecBufs<T>::st(L);
  }
  B.nu_down = l == 0 ? h.nu_pre_ : h.nu_coarse_;
  B.nu_up = l == 0 ? h.nu_post_ : h.nu_coarse_;
  if (B.nu_down < 1 || B.nu_down > kMaxNu || B.nu_up < 1 || B.nu_up > kMaxNu) return false;
  for (int k = 0; k < B.nu_down; ++k) B.omega_down[k] = (T)h.sweep_omega(k, B.nu_down);
  for (int k = 0; k < B.nu_up; ++k) B.omega_up[k] = (T)h.sweep_omega(k, B.nu_up);
  if (h.band_wide_ == 1 && B.src == kSrcCells) {
    // one node per thread of a kBandThreadsWide workgroup: the largest even square tile whose
    // extended tile fits it (24 x 24 for three sweeps)
    const int nu = std::max(B.nu_down, B.nu_up);
    int t = 2;
    while (band_ext(t + 2, nu, true) * band_ext(t + 2, nu, true) <= (size_t)kBandThreadsWide) t += 2;
    if (band_ext(t, nu, true) * band_ext(t, nu, true) <= (size_t)kBandThreadsWide &&
        band_lds_bytes(t, t, nu, h.bs_, sizeof(T), B.src, true) <= kBandLdsBytes &&
        !(l == 0 && band_tiles(L.nx, L.ny, t, t) > kMaxPartials)) {
      B.tx = B.ty = t;
      B.wide = 1;
      return true;
    }
  }
  static const int cand[][2] = {{32, 32}, {32, 16}, {16, 16}, {16, 8}, {8, 8}};
  bool found = false;
  for (const auto& c : cand) {
    const int nu = std::max(B.nu_down, B.nu_up);
    if (band_lds_bytes(c[0], c[1], nu, h.bs_, sizeof(T), B.src, true) > kBandLdsBytes) continue;
    const int tiles = band_tiles(L.nx, L.ny, c[0], c[1]);
    if (l == 0 && tiles > kMaxPartials) continue;  // one r.z partial per tile
    B.tx = c[0];
    B.ty = c[1];
    found = true;
    if (tiles >= 480) break;
  }
  if (found && h.band_wide_ >= 2 && B.src == kSrcCells) B.wide = h.band_wide_;
  return found;
}

template <typename T>
int Hierarchy::vcycle(int l, const T* b, T* out, hipStream_t s, RedSlot* dot_bz) {
  const int last = (int)levels_.size() - 1;
  if (coarse_pending_ && l >= coarse_first_) TMX_TRY(join_coarse(s));
  if (l == last) {
    TMX_HIP(dense_apply<T>(cn_, dense_.get(), b, out, s));
    if (dot_bz) TMX_HIP(dot_t<T>((size_t)cn_, b, out, *dot_bz, s));
    return 0;
  }
  if (use_tail_ && l == tail_ && l > 0) return tail_launch<T>(*this, b, out, s);
  if (l == 0 && phys_ == kFilter && use_dct_ && !use_graphs_ &&
      DctPrecond<T>::supported(levels_[0].nx, levels_[0].ny)) {
    auto& dct = dct_slot<T>();
    if (!dct) {
      std::unique_ptr<DctPrecond<T>> d(new DctPrecond<T>);
      TMX_TRY(d->init(levels_[0].nx, levels_[0].ny, levels_[0].hx, levels_[0].hy, eps_, s));
      dct = std::move(d);
    }
    return dct->apply(b, out, dot_bz, s);
  }
  Level& L = levels_[l];
  Level& C = levels_[l + 1];
  const int pk = sizeof(T) == 8 ? 1 : 0;
  if (l == 0 && phys_ == kFilter && L.kind == kElem && C.nested_from_prev &&
      fpipe_rows_[pk][0] > 0 && fpipe_rows_[pk][1] > 0 && nu_pre_ <= kMaxPipeNu &&
      nu_post_ <= kMaxPipeNu) {
    T* t = PrecBufs<T>::t(L);
    T* cb = PrecBufs<T>::b(C);
    T* cx = PrecBufs<T>::x(C);
    T om_d[kMaxPipeNu], om_u[kMaxPipeNu];
    for (int k = 0; k < nu_pre_; ++k) om_d[k] = (T)sweep_omega(k, nu_pre_);
    for (int k = 0; k < nu_post_; ++k) om_u[k] = (T)sweep_omega(k, nu_post_);
    const FilterConst<T> fk = FilterConst<T>::make(eps_, L.hx, L.hy);
    TMX_HIP(filter_pipe_down<T>(L.nx, L.ny, fk, nu_pre_, om_d, fpipe_rows_[pk][0], b, t, cb, s));
    TMX_TRY(vcycle<T>(1, cb, cx, s));
    TMX_HIP(filter_pipe_up<T>(L.nx, L.ny, fk, nu_post_, om_u, fpipe_rows_[pk][1], t, cx, b, out,
                              dot_bz, s));
    return 0;
  }
  // elasticity element levels above pair_levels_ run the paired column-strip kernels
  const bool pair_level = phys_ == kElasticity && L.kind == kElem && l < pair_levels_;
  BandLevel<T> B;
  if (!pair_level && band_level<T>(*this, l, B)) {
    T* t = PrecBufs<T>::t(L);
    T* cb = PrecBufs<T>::b(C);
    T* cx = PrecBufs<T>::x(C);
    if (bs_ == 2) TMX_HIP((band_down<T, 2>(B, b, t, cb, s)));
    else TMX_HIP((band_down<T, 1>(B, b, t, cb, s)));
    TMX_TRY(vcycle<T>(l + 1, cb, cx, s));
    RedSlot* dot = l == 0 ? dot_bz : nullptr;
    if (bs_ == 2) TMX_HIP((band_up<T, 2>(B, t, cx, b, out, dot, s)));
    else TMX_HIP((band_up<T, 1>(B, t, cx, b, out, dot, s)));
    return 0;
  }
  T* buf[2] = {out, PrecBufs<T>::t(L)};
  const int nuc = pair_level && pair_nu_coarse_ > 0 ? pair_nu_coarse_ : nu_coarse_;
  const int npre = l == 0 ? nu_pre_ : nuc;
  const int npost = l == 0 ? nu_post_ : nuc;
  T* cb = PrecBufs<T>::b(C);
  T* cx = PrecBufs<T>::x(C);
  const bool fused_el = fused_ && L.kind == kElem && C.nested_from_prev;  // nested transfer
  // paired elasticity levels run their first two pre-sweeps, and prolongation with their first
  // two post-sweeps, as one launch each
  const int pi = sizeof(T) == 8 ? 1 : 0;
  const int rows_pre = l == 0 ? pair_rows_[pi] : pair_rows_coarse_[0];
  const int rows_post = l == 0 ? pair_rows_post_[pi] : pair_rows_coarse_[1];
  const int rows_rr = l == 0 ? pair_rr_rows_[pi] : pair_rows_coarse_[2];
  const bool pair_pre = pair_level && rows_pre > 0 && npre >= 2;
  const bool pair_post = pair_level && rows_post > 0 && fused_el && npost >= 2;
  const bool pair_rr = pair_pre && npre == 2 && fused_el && rows_rr > 0;
  const int launches = npre - (pair_pre ? 1 : 0) + npost - (pair_post ? 1 : 0);
  int w = (launches - 1) % 2;
  int kpre = 1;
  if (pair_rr) {  // writes the pre-smoothed iterate and the coarse right-hand side cb
    TMX_TRY(sweep_pair_rr0<T>(rows_rr, b, buf[w], sweep_omega(0, npre), sweep_omega(1, npre), s,
                              l));
    kpre = 2;
  } else if (pair_pre) {
    TMX_TRY(sweep_pair0<T>(kSmoothZero, rows_pre, nullptr, b, buf[w], sweep_omega(0, npre),
                           sweep_omega(1, npre), nullptr, s, l));
    kpre = 2;
  } else {
    TMX_TRY(level_op<T>(*this, l, kSmoothZero, nullptr, b, buf[w], nullptr, s, SweepExtra<T>(),
                        true, sweep_omega(0, npre)));
  }
  for (; kpre < npre; ++kpre) {
    TMX_TRY(level_op<T>(*this, l, kSmooth, buf[w], b, buf[w ^ 1], nullptr, s, SweepExtra<T>(),
                        true, sweep_omega(kpre, npre)));
    w ^= 1;
  }
  // (9-point levels keep separate residual / restriction / prolongation launches: recomputing
  // the fine residuals per coarse node costs more than the extra launch there)
  const bool fused_st = false;
  if (pair_rr) {
    // done by the pair
  } else if (fused_el) {
    TMX_TRY(level_op<T>(*this, l, kResidualRestrict, buf[w], b, cb, nullptr, s));
  } else if (fused_st) {
    if (bs_ == 2)
      TMX_HIP((stencil_resrestrict<T, 2>(L.nx, L.ny, C.nx, C.ny, fixed_, PrecBufs<T>::st(L),
                                         buf[w], b, cb, s)));
    else
      TMX_HIP((stencil_resrestrict<T, 1>(L.nx, L.ny, C.nx, C.ny, fixed_, PrecBufs<T>::st(L),
                                         buf[w], b, cb, s)));
  } else {
    T* r = PrecBufs<T>::r(L);
    TMX_TRY(level_op<T>(*this, l, kResidual, buf[w], b, r, nullptr, s));
    TMX_TRY(restrict_to<T>(*this, l, r, cb, s));
  }
  TMX_TRY(vcycle<T>(l + 1, cb, cx, s));
  int post = npost, j = 0;  // j: index of the next post-sweep in the smoothing pass
  RedSlot* last_dot = l == 0 ? dot_bz : nullptr;
  if (pair_post) {
    TMX_TRY(sweep_pair0<T>(kProlongSmooth, rows_post, buf[w], b, buf[w ^ 1],
                           sweep_omega(0, npost), sweep_omega(1, npost),
                           post == 2 ? last_dot : nullptr, s, l));
    j = 2;
    w ^= 1;
    post -= 2;
  } else if (fused_el) {
    SweepExtra<T> ex;
    ex.ec = cx;
    TMX_TRY(level_op<T>(*this, l, kProlongSmooth, buf[w], b, buf[w ^ 1],
                        post == 1 ? last_dot : nullptr, s, ex, true, sweep_omega(j++, npost)));
    w ^= 1;
    --post;
  } else if (fused_st) {
    if (bs_ == 2)
      TMX_HIP((stencil_prolong_smooth<T, 2>(L.nx, L.ny, C.nx, C.ny, fixed_, PrecBufs<T>::st(L),
                                            buf[w], cx, b, buf[w ^ 1], (T)omega_, s)));
    else
      TMX_HIP((stencil_prolong_smooth<T, 1>(L.nx, L.ny, C.nx, C.ny, fixed_, PrecBufs<T>::st(L),
                                            buf[w], cx, b, buf[w ^ 1], (T)omega_, s)));
    w ^= 1;
    --post;
  } else {
    TMX_TRY(prolong_to<T>(*this, l, cx, buf[w], s));
  }
  for (int k = 0; k < post; ++k) {
    TMX_TRY(level_op<T>(*this, l, kSmooth, buf[w], b, buf[w ^ 1],
                        k == post - 1 ? last_dot : nullptr, s, SweepExtra<T>(), true,
                        sweep_omega(j++, npost)));
    w ^= 1;
  }
  return 0;
}

int Hierarchy::estimate_lmax(Workspace& ws, hipStream_t s, double margin) {
  constexpr int kIters = 24;
  TMX_TRY(join_coarse(s));
  RedSlot r = ws.region(3);
  double lmax = 0.0;
  const int last = (int)levels_.size() - 1;
  for (int l = 0; l < last; ++l) {
    Level& L = levels_[l];
    const size_t n = (size_t)bs_ * L.nn;
    double* x = L.x64.get();
    double* y = L.t64.get();
    double* zero = L.b64.get();
    TMX_HIP(fill_zero_bytes(zero, sizeof(double) * n, s));
    TMX_HIP(fill_hash(n, x, s));
    double est = 0.0;
    for (int it = 0; it < kIters; ++it) {
      TMX_HIP(dot(n, x, x, r, ws.sc() + kScTmp, s));
      // y = x + D^-1 A x (one Jacobi sweep on b = 0 with step -1): eigenvalues 1 + lambda
      TMX_TRY(level_op<double>(*this, l, kSmooth, x, zero, y, nullptr, s, SweepExtra<double>(),
                               true, -1.0));
      TMX_HIP(dot(n, y, y, r, ws.sc() + kScTmp + 1, s));
      TMX_TRY(ws.read(s, kScTmp, 2));
      const double xx = ws.host[kScTmp], yy = ws.host[kScTmp + 1];
      if (!(xx > 0) || !std::isfinite(yy)) break;
      est = std::sqrt(yy / xx) - 1.0;
      TMX_HIP(scale_copy(n, 1.0 / std::sqrt(yy), y, x, s));
    }
    lmax = std::max(lmax, est);
    if (g_trace) std::fprintf(stderr, "[tmx] level %d lambda_max(D^-1 A) ~ %.4f\n", l, est);
  }
  cheb_lmax_ = lmax > 0.0 ? margin * lmax : 0.0;
  cheb_done_ = lmax > 0.0;  // (no estimate from a non-finite operator: the next one tries)
  drop_graphs();
  return 0;
}

// ------------------------------------------------------------------ Krylov
template <typename T>
struct KrylovBufs;
template <>
struct KrylovBufs<double> {
  static double* x(Hierarchy& h) { return h.xit_ ? h.xit_ : h.x64_.get(); }
  static double* r(Hierarchy& h) { return h.r64_.get(); }
  static double* z(Hierarchy& h) { return h.z64_.get(); }
  static double* p(Hierarchy& h) { return h.p64_.get(); }
  static double* q(Hierarchy& h) { return h.q64_.get(); }
  static double* pb(Hierarchy& h) { return h.pb64_.get(); }
};
template <>
struct KrylovBufs<float> {
  static float* x(Hierarchy& h) { return h.d32_.get(); }
  static float* r(Hierarchy& h) { return h.r32_.get(); }
  static float* z(Hierarchy& h) { return h.z32_.get(); }
  static float* p(Hierarchy& h) { return h.p32_.get(); }
  static float* q(Hierarchy& h) { return h.q32_.get(); }
  static float* pb(Hierarchy& h) { return h.pb32_.get(); }
};

template <typename T>
T* Hierarchy::x_buf() {
  return KrylovBufs<T>::x(*this);
}

// z = M r for Krylov vectors in T and the V-cycle in PC (partials of r.z to rz).  With an FP32
// V-cycle under FP64 Krylov vectors, r arrives already rounded in the float buffer r32_
// (written by the update kernel, or converted here at the start) and z stays in float (z32_),
// which the direction update reads directly.
template <typename T, typename PC>
static int enqueue_precondition(Hierarchy& h, RedSlot* rz, bool convert_r, hipStream_t s) {
  if constexpr (std::is_same<T, PC>::value) {
    TMX_TRY(h.vcycle<T>(0, KrylovBufs<T>::r(h), KrylovBufs<T>::z(h), s, rz));
  } else {
    static_assert(std::is_same<T, double>::value && std::is_same<PC, float>::value,
                  "mixed PCG is FP64 Krylov with an FP32 V-cycle");
    const size_t nv = h.n();
    if (convert_r) TMX_HIP((convert<double, float>(nv, h.r64_.get(), h.r32_.get(), s)));
    TMX_TRY(h.vcycle<float>(0, h.r32_.get(), h.z32_.get(), s, rz));
  }
  return 0;
}

// PCG start, given r = b already in place: z = M r, rz = r.z, beta = 0, p_old = 0
template <typename T, typename PC>
static int enqueue_pcg_init(Hierarchy& h, Workspace& ws, hipStream_t s) {
  const size_t nv = h.n();
  RedSlot rz = ws.region(2);
  TMX_TRY((enqueue_precondition<T, PC>(h, &rz, false, s)));  // float r written by pcg_begin
  TMX_HIP(fill_zero_bytes(KrylovBufs<T>::p(h), sizeof(T) * nv, s));
  TMX_HIP(pcg_start(rz, ws.sc(), kPcg, s));
  return 0;
}

// One PCG iteration, all scalars on the device, four launches plus the V-cycle:
//   p = z + beta p_old and q = A p (partials p.q); alpha; x += alpha p, r -= alpha q
//   (partials r.r); z = M r (partials r.z from the last sweep); rr, beta, stop flags.
// parity selects which of the two direction buffers holds p_old.
template <typename T, typename PC>
static int enqueue_pcg_iteration(Hierarchy& h, Workspace& ws, int parity, hipStream_t s) {
  const size_t nv = h.n();
  T* x = KrylovBufs<T>::x(h);
  T* r = KrylovBufs<T>::r(h);
  T* z = KrylovBufs<T>::z(h);
  T* q = KrylovBufs<T>::q(h);
  T* p_old = parity ? KrylovBufs<T>::pb(h) : KrylovBufs<T>::p(h);
  T* p_new = parity ? KrylovBufs<T>::p(h) : KrylovBufs<T>::pb(h);
  double* sc = ws.sc();
  RedSlot pq = ws.region(0), rr = ws.region(1), rz = ws.region(2);
  constexpr bool split = !std::is_same<T, PC>::value;
  // x += alpha p of the previous iteration rides along with the direction update (which reads
  // p_old anyway); this iteration's alpha is left in sc[alpha] for the next one, and the last
  // one is applied after the loop (pcg's flush)
  const bool timed = h.au_timing_ && !h.use_graphs_ && h.phys_ == kElasticity;
  if (timed) TMX_TRY(h.au_record(0, s));
  TMX_TRY(h.apply_update<T>(p_old, z, q, p_new, sc + kPcg.beta, &pq, s,
                            split ? h.z32_.get() : nullptr, x, sc + kPcg.alpha));
  if (timed) TMX_TRY(h.au_record(1, s));
  TMX_HIP(pcg_update_xr_fused<T>(nv, sc, kPcg, pq, nullptr, r, p_new, q, rr, s,
                                 split ? h.r32_.get() : nullptr));
  TMX_TRY((enqueue_precondition<T, PC>(h, &rz, false, s)));
  // the iteration graph of this parity reports into its own ring slot
  TMX_HIP(pcg_finalize(pq, rr, rz, sc, kPcg, s, ws.host_dev + Workspace::kScalars + 4 * parity));
  return 0;
}

template <typename T, typename PC>
static int capture(Hierarchy& h, Workspace& ws, int which, hipGraphExec_t* out, hipStream_t s) {
  hipGraph_t g = nullptr;
  TMX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int rc = which == 0 ? enqueue_pcg_init<T, PC>(h, ws, s)
                            : enqueue_pcg_iteration<T, PC>(h, ws, which - 1, s);
  const hipError_t e = hipStreamEndCapture(s, &g);
  if (rc != 0) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  TMX_HIP(e);
  const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  TMX_HIP(ei);
  return 0;
}

template <typename T, typename PC>
int Hierarchy::pcg_enqueue_begin(Workspace& ws, const T* b, double rtol, bool init,
                                 hipStream_t s) {
  RedSlot bbs = ws.region(3);
  constexpr bool split = !std::is_same<T, PC>::value;
  TMX_HIP(pcg_begin<T>(n(), b, KrylovBufs<T>::x(*this), KrylovBufs<T>::r(*this),
                       split ? r32_.get() : nullptr, bbs, s));
  TMX_HIP(pcg_begin_final(bbs, ws.sc(), kPcg, rtol * rtol, kScBB,
                          ws.host_dev + Workspace::kScalars + kRingBB, s));
  if (init) TMX_TRY((enqueue_pcg_init<T, PC>(*this, ws, s)));
  return 0;
}

template <typename T, typename PC>
int Hierarchy::pcg_enqueue_iteration(Workspace& ws, int parity, hipStream_t s) {
  return enqueue_pcg_iteration<T, PC>(*this, ws, parity, s);
}

template <typename T>
int Hierarchy::pcg_enqueue_end(Workspace& ws, int launched, hipStream_t s, bool gated) {
  T* p_last = ((launched - 1) & 1) ? KrylovBufs<T>::p(*this) : KrylovBufs<T>::pb(*this);
  TMX_HIP(axpy_dev<T>(n(), ws.sc() + kPcg.alpha, p_last, KrylovBufs<T>::x(*this), s,
                      gated ? ws.sc() + kPcg.flag : nullptr));
  return 0;
}

template <typename T, typename PC>
int Hierarchy::pcg(Workspace& ws, const T* b, double rtol, int max_iter, double* bnorm2,
                   int* iters, double* rel, hipStream_t s) {
  const size_t nv = n();
  const int prec = !std::is_same<T, PC>::value ? 2 : sizeof(T) == 8 ? 0 : 1;
  if (graphs_cw0_ != has_cw0_) drop_graphs();  // captured with the other finest-level source
  graphs_cw0_ = has_cw0_;
  if (use_graphs_)
    for (int w = 0; w < 3; ++w)
      if (!graphs_[prec][w]) TMX_TRY((capture<T, PC>(*this, ws, w, &graphs_[prec][w], s)));
  *iters = 0;
  *rel = 1.0;
  // r = b, x = 0 (and the float r of a mixed V-cycle) with |b|^2, the tolerance and the start
  // flag formed on the device: no host round trip before the first iteration
  volatile double* host_bb = ws.host + Workspace::kScalars + kRingBB;
  host_bb[0] = std::numeric_limits<double>::quiet_NaN();
  TMX_TRY((pcg_enqueue_begin<T, PC>(ws, b, rtol, !use_graphs_, s)));
  if (use_graphs_) TMX_HIP(hipGraphLaunch(graphs_[prec][0], s));
  // Iteration k runs graph 1 + (k-1)%2, whose last kernel writes (flag, r.r, iterations) and
  // then the sequence number k to ring slot (k-1)%2 in pinned host memory; the host polls that
  // number (no event packets between the graphs).  While the host waits for iteration k it
  // normally has iteration k+1 queued already, so the device never idles.  An iteration queued
  // behind the one that converges would run in full for nothing (its stop flag makes it change
  // nothing), so once the residual history predicts convergence within the next iteration the
  // host stops queueing ahead and pays one host round trip instead.  At most two iterations are
  // in flight, so a ring slot is rewritten (by iteration k+2) only after the host has read it.
  volatile double* ring = ws.host + Workspace::kScalars;
  auto slot = [&](int k) { return ring + 4 * ((k - 1) & 1); };
  // no PCG iteration of this stream is in flight: an earlier solve returned only after every
  // iteration it queued had reported
  ring[3] = ring[7] = 0.0;
  auto launch = [&](int k) -> hipError_t {
    if (use_graphs_) return hipGraphLaunch(graphs_[prec][1 + ((k - 1) & 1)], s);
    if (enqueue_pcg_iteration<T, PC>(*this, ws, (k - 1) & 1, s) != 0) return hipErrorUnknown;
    return hipSuccess;
  };
  // wait until iteration k has reported; a stream that drains without the report is an error,
  // never a hang
  auto wait_iter = [&](int k) -> int {
    volatile double* sl = slot(k);
    for (long spins = 1; sl[3] < (double)k; ++spins) {
      if ((spins & 1023) == 0) {
        const hipError_t q = hipStreamQuery(s);
        if (q == hipSuccess && sl[3] < (double)k) {
          set_error("internal: PCG iteration " + std::to_string(k) + " did not report");
          return -2;
        }
        if (q != hipSuccess && q != hipErrorNotReady) TMX_HIP(q);
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return 0;
  };
  double bb = -1.0, tol2 = 0.0;  // |b|^2 from the ring, read with the first report
  double flag = 0.0, rr_last = 0.0, ratio = 1.0;
  bool near = false;
  int launched = 0, done = 0;
  // The iteration whose x update was queued gated on its stop flag: when nothing is queued
  // behind the awaited iteration (convergence predicted, or the iteration limit), its update
  // is queued before the wait, so that a converged solve's end does not wait for the host's
  // round trip; unconverged, it changes nothing and the next iteration applies the update.
  int gated = 0;
  if (max_iter >= 1) {
    TMX_HIP(launch(1));
    launched = 1;
  }
  while (done < launched) {
    if (launched == done + 1 && launched < max_iter && !near) {
      TMX_HIP(launch(launched + 1));
      ++launched;
    }
    if (launched == done + 1 && gate_end_) {
      TMX_TRY(pcg_enqueue_end<T>(ws, launched, s, true));
      gated = launched;
    }
    TMX_TRY(wait_iter(done + 1));
    ++done;
    if (bb < 0.0) {
      bb = host_bb[0];
      tol2 = rtol * rtol * bb;
      rr_last = bb;
    }
    flag = slot(done)[0];
    if (flag != 0.0) break;
    const double rr = slot(done)[1];
    if (rr_last > 0.0) ratio = rr / rr_last;
    rr_last = rr;
    near = rr * std::min(ratio, 1.0) <= 4.0 * tol2;
    if (launched == done && done < max_iter) {
      TMX_HIP(launch(done + 1));
      ++launched;
    }
  }
  // the last iteration's deferred x += alpha p (alpha is 0 once stopped), queued before the
  // host waits for the last queued iteration's report, so the device does not idle meanwhile
  // (already queued when the gated update of the last iteration found its stop flag set)
  if (launched > 0 && !(gated == launched && done == launched && flag != 0.0))
    TMX_TRY(pcg_enqueue_end<T>(ws, launched, s));
  if (launched > done) TMX_TRY(wait_iter(launched));
  if (launched == 0) {  // no iteration reported |b|^2: read it
    TMX_TRY(ws.read(s, kScBB, 1));
    bb = ws.host[kScBB];
  } else if (bb < 0.0) {
    bb = host_bb[0];
  }
  *bnorm2 = bb;
  if (launched > 0) flag = slot(launched)[0];
  const double rrv = launched > 0 ? slot(launched)[1] : bb;
  *iters = launched > 0 ? (int)slot(launched)[2] : 0;
  *rel = bb > 0.0 ? std::sqrt(std::max(rrv, 0.0) / bb) : 0.0;
  if (g_trace)
    std::fprintf(stderr, "[tmx] pcg<%s> %s n=%zu: %d its, rel %.3e, flag %g\n",
                 prec == 2 ? "f64/f32" : prec == 0 ? "f64" : "f32", phys_ == kElasticity ? "elast" : "filter", nv,
                 *iters, *rel, flag);
  if (flag == 1.0) return 0;
  if (flag == 2.0) {
    set_error("PCG breakdown: p.Ap <= 0 (operator not positive definite)");
    return -4;
  }
  if (flag == 3.0) {
    set_error("non-finite right-hand side or residual in PCG");
    return -5;
  }
  set_error("PCG did not reach the tolerance within the iteration limit");
  return -3;
}

// Reliable-update PCG iteration: the FP32 Krylov iteration of enqueue_pcg_iteration<float>
// with the increments of the FP64 solution gathered in the FP32 buffer d32_ (deferred_x: this
// iteration adds the previous one's alpha p there; off after a replacement, which has added it
// already).  With `replace`, after the x/r update: x64 += d32 + alpha p, d32 = 0, then the
// recursive FP32 residual is replaced by the true one, r64 = b - A x64 with the FP64 operator
// (r64_), before the V-cycle; the stop test of that iteration sees |r64|^2.
static int enqueue_ru_iteration(Hierarchy& h, Workspace& ws, int parity, bool replace,
                                bool deferred_x, const double* b64, hipStream_t s) {
  const size_t nv = h.n();
  float* xs = h.d32_.get();
  float* r = h.r32_.get();
  float* z = h.z32_.get();
  float* q = h.q32_.get();
  float* p_old = parity ? h.pb32_.get() : h.p32_.get();
  float* p_new = parity ? h.p32_.get() : h.pb32_.get();
  double* sc = ws.sc();
  RedSlot pq = ws.region(0), rr = ws.region(1), rz = ws.region(2);
  const bool timed = h.au_timing_ && h.phys_ == kElasticity;
  if (timed) TMX_TRY(h.au_record(0, s));
  TMX_TRY(h.apply_update<float>(p_old, z, q, p_new, sc + kPcg.beta, &pq, s, nullptr,
                                deferred_x ? xs : nullptr, sc + kPcg.alpha));
  if (timed) TMX_TRY(h.au_record(1, s));
  TMX_HIP(pcg_update_xr_fused<float>(nv, sc, kPcg, pq, nullptr, r, p_new, q, rr, s, nullptr));
  if (replace) {
    TMX_HIP(ru_flush(nv, sc + kPcg.alpha, p_new, xs, h.x64_.get(), s));
    TMX_TRY(h.residual<double>(h.x64_.get(), b64, h.r64_.get(), s));
    TMX_HIP(ru_replace(nv, h.r64_.get(), r, rr, s));
  }
  TMX_TRY((enqueue_precondition<float, float>(h, &rz, false, s)));
  TMX_HIP(pcg_finalize(pq, rr, rz, sc, kPcg, s, ws.host_dev + Workspace::kScalars + 4 * parity));
  return 0;
}

int Hierarchy::pcg_ru(Workspace& ws, const double* b, double rtol, int max_iter, double* bnorm2,
                      int* iters, double* rel, hipStream_t s) {
  const size_t nv = n();
  double* sc = ws.sc();
  *iters = 0;
  *rel = 1.0;
  // Iteration k replaces its residual ru_every_ iterations after the last replacement, never
  // while a replacement is still in flight, and not once the recursive residual is predicted
  // (from the reported history: the last report times the recent reduction per iteration)
  // within 10x of the tolerance: the recursion finishes from there.  Replacing close to the
  // FP64 round-off floor of the true residual (about 5e-11 relative at N=400, 2e-10 at N=1000
  // on these designs) would restart the recursion from that floor and stall it.
  const int every = std::max(1, ru_every_);
  std::vector<char> replaced((size_t)std::max(1, max_iter) + 2, 0);
  auto replace_at = [&](int k) { return replaced[(size_t)k] != 0; };
  int last_rep = 0;
  bool rep_pending = false;
  // x64 = 0; the FP32 iteration starts from r = (float) b, d32 = 0, p = 0
  TMX_HIP(fill_zero_bytes(x64_.get(), sizeof(double) * nv, s));
  float* b32 = levels_[0].b32.get();
  TMX_HIP((convert<double, float>(nv, b, b32, s)));
  volatile double* host_bb = ws.host + Workspace::kScalars + kRingBB;
  host_bb[0] = std::numeric_limits<double>::quiet_NaN();
  RedSlot bbs = ws.region(3);
  TMX_HIP(pcg_begin<float>(nv, b32, d32_.get(), r32_.get(), nullptr, bbs, s));
  TMX_HIP(pcg_begin_final(bbs, sc, kPcg, rtol * rtol, kScBB,
                          ws.host_dev + Workspace::kScalars + kRingBB, s));
  TMX_TRY((enqueue_pcg_init<float, float>(*this, ws, s)));
  // the host loop of pcg(): iteration k reports into ring slot (k - 1) % 2, the host keeps
  // one iteration queued ahead until the residual history predicts convergence
  volatile double* ring = ws.host + Workspace::kScalars;
  auto slot = [&](int k) { return ring + 4 * ((k - 1) & 1); };
  ring[3] = ring[7] = 0.0;
  double bb = -1.0, tol2 = 0.0, flag = 0.0, rr_last = 0.0, ratio = 1.0;
  double rr_now = -1.0, per_it = 1.0;  // last reported |r|^2 and its reduction per iteration
  int rr_at = 0;                         // the iteration it was reported by
  auto launch = [&](int k) -> hipError_t {
    bool rep = !rep_pending && k - last_rep >= every;
    if (rep && rr_now >= 0.0 && bb > 0.0) {
      const double pred = rr_now * std::pow(std::min(per_it, 1.0), (double)(k - rr_at));
      rep = pred > 100.0 * tol2;
    }
    replaced[(size_t)k] = rep ? 1 : 0;
    if (rep) {
      last_rep = k;
      rep_pending = true;
    }
    const bool deferred = k == 1 || !replace_at(k - 1);
    if (enqueue_ru_iteration(*this, ws, (k - 1) & 1, replace_at(k), deferred, b, s) != 0)
      return hipErrorUnknown;
    return hipSuccess;
  };
  auto wait_iter = [&](int k) -> int {
    volatile double* sl = slot(k);
    for (long spins = 1; sl[3] < (double)k; ++spins) {
      if ((spins & 1023) == 0) {
        const hipError_t q = hipStreamQuery(s);
        if (q == hipSuccess && sl[3] < (double)k) {
          set_error("internal: PCG iteration " + std::to_string(k) + " did not report");
          return -2;
        }
        if (q != hipSuccess && q != hipErrorNotReady) TMX_HIP(q);
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return 0;
  };
  bool near = false;
  int launched = 0, done = 0;
  if (max_iter >= 1) {
    TMX_HIP(launch(1));
    launched = 1;
  }
  while (done < launched) {
    if (launched == done + 1 && launched < max_iter && !near) {
      TMX_HIP(launch(launched + 1));
      ++launched;
    }
    TMX_TRY(wait_iter(done + 1));
    ++done;
    if (bb < 0.0) {
      bb = host_bb[0];
      tol2 = rtol * rtol * bb;
      rr_last = bb;
    }
    flag = slot(done)[0];
    {
      const double rrd = slot(done)[1];
      if (rr_now > 0.0 && rrd > 0.0) per_it = std::pow(rrd / rr_now, 1.0 / (double)(done - rr_at));
      rr_now = rrd;
      rr_at = done;
      if (replace_at(done)) rep_pending = false;
    }
    if (g_trace)
      std::fprintf(stderr, "[tmx] pcg-ru it %d%s rel %.3e flag %g\n", done,
                   replace_at(done) ? " (replaced)" : "",
                   bb > 0.0 ? std::sqrt(std::max((double)slot(done)[1], 0.0) / bb) : 0.0, flag);
    if (flag != 0.0) break;
    const double rr = slot(done)[1];
    if (rr_last > 0.0) ratio = rr / rr_last;
    rr_last = rr;
    near = rr * std::min(ratio, 1.0) <= 4.0 * tol2;
    if (launched == done && done < max_iter) {
      TMX_HIP(launch(done + 1));
      ++launched;
    }
  }
  // the increments not yet in x64: d32 and, unless the last iteration replaced, its alpha p
  // (queued before the wait for the last report, as in pcg)
  if (launched > 0 && !replace_at(launched)) {
    float* p_last = ((launched - 1) & 1) ? p32_.get() : pb32_.get();
    TMX_HIP(ru_flush(nv, sc + kPcg.alpha, p_last, d32_.get(), x64_.get(), s));
  }
  if (launched > done) TMX_TRY(wait_iter(launched));
  if (launched == 0) {
    TMX_TRY(ws.read(s, kScBB, 1));
    bb = ws.host[kScBB];
  } else if (bb < 0.0) {
    bb = host_bb[0];
  }
  *bnorm2 = bb;
  if (launched > 0) flag = slot(launched)[0];
  const double rrv = launched > 0 ? slot(launched)[1] : bb;
  *iters = launched > 0 ? (int)slot(launched)[2] : 0;
  *rel = bb > 0.0 ? std::sqrt(std::max(rrv, 0.0) / bb) : 0.0;
  if (g_trace)
    std::fprintf(stderr, "[tmx] pcg-ru n=%zu: %d its, rel %.3e, flag %g\n", nv, *iters, *rel, flag);
  if (flag == 1.0) return 0;
  if (flag == 2.0) {
    set_error("PCG breakdown: p.Ap <= 0 (operator not positive definite)");
    return -4;
  }
  if (flag == 3.0) {
    set_error("non-finite right-hand side or residual in PCG");
    return -5;
  }
  set_error("PCG did not reach the tolerance within the iteration limit");
  return -3;
}

// MINRES scalar slots (after the PCG slots)
constexpr MrSlots kMr = {30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 46, 47};

int minres_core(Workspace& ws, size_t nv, const double* b, double* X, MinresBufs bf,
                const std::function<int(const double*, double*)>& apply,
                const std::function<int(const double*, double*, RedSlot*)>& precond,
                double rtol, int max_iter, double bnorm2, int* iters, double* rel,
                hipStream_t s, bool trace, const MinresApplyL1& apply_l1) {
  constexpr int kMaxRestarts = 8;
  double* sc = ws.sc();
  double *R1 = bf.r1, *R2 = bf.r2, *Y = bf.y, *Z = bf.z, *V = bf.v;
  double* W[3] = {bf.w0, bf.w1, bf.w2};
  RedSlot part = ws.region(0), rz = ws.region(2), gen = ws.region(3);
  volatile double* ring = ws.host + Workspace::kScalars + 8;
  double* ring_dev = ws.host_dev + Workspace::kScalars + 8;
  TMX_HIP(fill_zero_bytes(X, sizeof(double) * nv, s));
  int total = 0, rc = -3, stalled = 0;
  double relres = 1.0, best = 1.0;
  for (int restart = 0;; ++restart) {
    TMX_TRY(apply(X, Y));
    TMX_HIP(sub_vec(nv, b, Y, R1, s));  // r1 = b - A x
    TMX_HIP(dot(nv, R1, R1, gen, sc + kScTmp, s));
    TMX_TRY(ws.read(s, kScTmp, 1));
    relres = std::sqrt(std::max(ws.host[kScTmp], 0.0) / bnorm2);
    if (!std::isfinite(relres)) {
      set_error("non-finite residual in MINRES");
      return -5;
    }
    if (relres <= rtol) {
      rc = 0;
      break;
    }
    if (trace) std::fprintf(stderr, "[tmx] minres restart %d: |r|/|b| = %.3e after %d its\n",
                            restart, relres, total);
    // the true residual has a round-off floor: stop once two restarts gain less than 2x
    if (restart > 0) {
      stalled = relres > 0.5 * best ? stalled + 1 : 0;
      best = std::min(best, relres);
      if (stalled >= 2) break;
    }
    if (restart >= kMaxRestarts || total >= max_iter) break;
    // reduction asked of this run, measured in the preconditioner's norm
    const double tol = 0.5 * rtol / relres;
    TMX_TRY(precond(R1, Z, &rz));
    TMX_HIP(mr_start(rz, sc, kMr, tol, s));
    TMX_HIP(copy_vec<double>(nv, R1, R2, s));
    TMX_HIP(fill_zero_bytes(W[1], sizeof(double) * nv, s));
    TMX_HIP(fill_zero_bytes(W[2], sizeof(double) * nv, s));
    TMX_HIP(mr_scale(nv, sc, kMr, Z, V, s));
    // Iteration k's mr_step writes (flag, |r|/|b| estimate, iterations) and then the sequence
    // number k to ring slot (k-1)%2; the host keeps iteration k+1 queued while it waits for
    // iteration k's report, so the device does not idle between iterations.  An iteration
    // queued behind the one that stops changes no iterate (mr_step marks it skipped and
    // mr_update leaves x alone); the true residual below synchronises the stream, so no
    // report of this run is still in flight when the next run resets the ring.
    auto slot = [&](int k) { return ring + 4 * ((k - 1) & 1); };
    ring[3] = ring[7] = 0.0;
    auto launch = [&](int k) -> int {
      if (apply_l1) {
        TMX_TRY(apply_l1(V, R1, Y, &part, sc + kMr.beta, sc + kMr.oldb));
        TMX_HIP(mr_lanczos2(nv, sc, kMr, part, R2, Y, s));
      } else {
        TMX_TRY(apply(V, Y));
        TMX_HIP(mr_lanczos(nv, sc, kMr, V, R1, R2, Y, part, s));
      }
      std::swap(R1, R2);  // r1 <- r2
      std::swap(R2, Y);   // r2 <- y (the new Lanczos vector); y <- the old r1 (free)
      TMX_TRY(precond(R2, Z, &rz));
      TMX_HIP(mr_step(rz, sc, kMr, ring_dev + 4 * ((k - 1) & 1), s));
      TMX_HIP(mr_update(nv, sc, kMr, V, Z, W[0], W[1], W[2], X, s));
      std::rotate(W, W + 1, W + 3);
      return 0;
    };
    auto wait_iter = [&](int k) -> int {
      volatile double* sl = slot(k);
      for (long spins = 1; sl[3] < (double)k; ++spins) {
        if ((spins & 1023) == 0) {
          const hipError_t q = hipStreamQuery(s);
          if (q == hipSuccess && sl[3] < (double)k) {
            set_error("internal: MINRES iteration " + std::to_string(k) + " did not report");
            return -2;
          }
          if (q != hipSuccess && q != hipErrorNotReady) TMX_HIP(q);
        }
      }
      std::atomic_thread_fence(std::memory_order_acquire);
      return 0;
    };
    // An iteration queued behind the one that stops runs in full for nothing (a preconditioner
    // application), so once the residual estimates predict the stop within the next
    // iteration the host stops queueing ahead and pays one round trip instead.
    const int budget = max_iter - total;
    int launched = 0, done = 0;
    double flag = 0.0, est_last = 1.0, ratio = 1.0;
    bool near = false;
    while (done < budget) {
      while (launched < budget && launched < done + (near ? 1 : 2)) TMX_TRY(launch(++launched));
      TMX_TRY(wait_iter(done + 1));
      ++done;
      flag = slot(done)[0];
      if (flag != 0.0) break;
      const double est = slot(done)[1];
      ratio = est_last > 0.0 ? est / est_last : 1.0;
      est_last = est;
      near = est * std::min(ratio, 1.0) <= 2.0 * tol;
    }
    total += done;
    if (flag == 2.0) {
      set_error("MINRES breakdown: preconditioner not positive definite");
      return -4;
    }
    if (flag == 3.0) {
      set_error("non-finite value in MINRES");
      return -5;
    }
  }
  *iters = total;
  *rel = relres;
  if (trace)
    std::fprintf(stderr, "[tmx] minres n=%zu: %d its, rel %.3e, rc %d\n", nv, total, relres, rc);
  if (rc != 0) {
    char msg[160];
    std::snprintf(msg, sizeof(msg),
                  "MINRES stopped at relative residual %.3e above the tolerance %.3e (iteration "
                  "limit or round-off floor)", relres, rtol);
    set_error(msg);
  }
  return rc;
}

int Hierarchy::minres(Workspace& ws, const double* b, double rtol, int max_iter, double bnorm2,
                      int* iters, double* rel, hipStream_t s) {
  if (!mrw_[0].get()) {
    set_error("internal: MINRES buffers exist for elasticity only");
    return -1;
  }
  MinresBufs bf{r64_.get(), pb64_.get(), q64_.get(), z64_.get(), p64_.get(),
                mrw_[0].get(), mrw_[1].get(), mrw_[2].get()};
  auto apply_fn = [&](const double* x, double* y) { return apply<double>(x, y, nullptr, s); };
  auto pc_fn = [&](const double* r, double* z, RedSlot* rz) { return vcycle<double>(0, r, z, s, rz); };
  return minres_core(ws, n(), b, x64_.get(), bf, apply_fn, pc_fn, rtol, max_iter, bnorm2, iters,
                     rel, s, g_trace);
}

static double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

int Hierarchy::solve(Workspace& ws, const double* b, double* x, const SolveParams& prm,
                     SolveStats* st, hipStream_t s) {
  const size_t nv = n();
  auto t0 = std::chrono::steady_clock::now();
  // a coarse-operator build of update_elasticity(overlap) that no V-cycle joined (no
  // iteration ran) is ordered before whatever s runs after the solve
  struct JoinAtExit {
    Hierarchy* h;
    hipStream_t s;
    ~JoinAtExit() { (void)h->join_coarse(s); }
  } join_at_exit{this, s};
  RedSlot gen = ws.region(3);
  SolveStats local;
  SolveStats& S = st ? *st : local;
  S.levels = levels();
  S.iterations = 0;
  S.outer = 0;
  int rc = 0;
  if (prm.precision == kSolveFP64 || prm.precision == kSolveMixed ||
      prm.precision == kSolveMixedRU) {
    // FP64 and mixed PCG iterate in the caller's x itself (no copy at the end), unless
    // captured graphs hold x64_ or x is b; MINRES and the reliable-update PCG use x64_
    bool in_x = !use_graphs_ && x != b && prm.precision != kSolveMixedRU;
    auto krylov = [&](int* its, double* relv, double* bbv) {
      xit_ = in_x ? x : nullptr;
      const int r = prm.precision == kSolveFP64
                        ? pcg<double>(ws, b, prm.rtol, prm.max_iter, bbv, its, relv, s)
                    : prm.precision == kSolveMixed
                        ? pcg<double, float>(ws, b, prm.rtol, prm.max_iter, bbv, its, relv, s)
                        : pcg_ru(ws, b, prm.rtol, prm.max_iter, bbv, its, relv, s);
      xit_ = nullptr;
      return r;
    };
    const bool defer = !prm.sync;
    if (defer) {
      for (auto& e : solve_ev_)
        if (!e) TMX_HIP(hipEventCreate(&e));
      TMX_HIP(hipEventRecord(solve_ev_[0], s));
    }
    // |b|^2 is formed on the device inside pcg (b = 0 converges at once with x = 0; a
    // non-finite b stops with the non-finite code)
    int it = 0;
    double rel = 0, bb = 0;
    rc = krylov(&it, &rel, &bb);
    S.iterations = it;
    S.outer = 1;
    if (rc == -4 && cheb_lmax_ > 0.0) {
      // Breakdown with Chebyshev smoothing: either the bound is too low for this operator or
      // the operator is indefinite.  This solve retries with damped Jacobi; the bound comes
      // back afterwards, and is re-estimated at the next operator when Jacobi succeeded (a
      // bound that was too low), so one design does not change the smoother of later ones.
      const double keep = cheb_lmax_;
      cheb_lmax_ = 0.0;
      drop_graphs();
      int it2 = 0;
      rc = krylov(&it2, &rel, &bb);
      S.iterations += it2;
      it += it2;
      cheb_lmax_ = keep;
      if (rc != -4) cheb_done_ = false;
      drop_graphs();
    }
    if (rc == -4 && phys_ == kElasticity) {  // K(xi) not positive definite: MINRES
      int it2 = 0;
      rc = minres(ws, b, prm.rtol, std::max(1, prm.max_iter - it), bb, &it2, &rel, s);
      S.iterations += it2;
      S.outer = 2;
      in_x = false;
    }
    if (rc == 0 || rc == -3) {
      if (!in_x) TMX_HIP(copy_vec<double>(nv, x64_.get(), x, s));
      S.rel_residual = rel;
      if (prm.precision == kSolveMixedRU && rc == 0 && bb > 0.0) {
        // the FP32 recursion stops on its own residual, which drifts from the true one after
        // the last replacement: check the true residual and, while it is above the tolerance
        // and the last round more than halved it, solve for the remaining residual
        // (b' = b - A x, x += d) in another round
        double prev = 0.0;
        for (int round = 0; round < 3; ++round) {
          TMX_TRY(residual<double>(x, b, rr64_.get(), s));
          TMX_HIP(dot(nv, rr64_.get(), rr64_.get(), gen, ws.sc() + kScVerify, s));
          TMX_TRY(ws.read(s, kScVerify, 1));
          const double true_rel = std::sqrt(std::max(ws.host[kScVerify], 0.0) / bb);
          if (!std::isfinite(true_rel)) {
            set_error("non-finite residual in the reliable-update solve");
            return -5;
          }
          S.rel_residual = true_rel;
          if (true_rel <= prm.rtol || (round > 0 && true_rel > 0.5 * prev) ||
              S.iterations >= prm.max_iter)
            break;
          prev = true_rel;
          int it2 = 0;
          double rel2 = 0, bb2 = 0;
          const int rc2 = pcg_ru(ws, rr64_.get(), prm.rtol / true_rel,
                                 prm.max_iter - S.iterations, &bb2, &it2, &rel2, s);
          S.iterations += it2;
          S.outer += 1;
          if (rc2 != 0 && rc2 != -3) return rc2;
          TMX_HIP(add_vec(nv, x64_.get(), x, s));
        }
      } else if (prm.verify && bb > 0.0) {  // the true residual
        TMX_TRY(residual<double>(x, b, rr64_.get(), s));
        TMX_HIP(dot(nv, rr64_.get(), rr64_.get(), gen, ws.sc() + kScVerify, s,
                    defer ? ws.host_dev + kScVerify : nullptr));
        if (defer) {
          S.pending_rr = ws.host + kScVerify;
          S.pending_bb = bb;
        } else {
          TMX_TRY(ws.read(s, kScVerify, 1));
          S.rel_residual = std::sqrt(ws.host[kScVerify] / bb);
        }
      }
    }
    if (defer) {
      TMX_HIP(hipEventRecord(solve_ev_[1], s));
      S.pending_ev[0] = solve_ev_[0];
      S.pending_ev[1] = solve_ev_[1];
      return rc;
    }
  } else {
    TMX_HIP(dot(nv, b, b, gen, ws.sc() + kScBB, s));
    TMX_TRY(ws.read(s, kScBB, 1));
    const double bb = ws.host[kScBB];
    if (!std::isfinite(bb)) {
      set_error("non-finite right-hand side");
      return -5;
    }
    if (bb == 0.0) {
      TMX_HIP(fill_zero_bytes(x, sizeof(double) * nv, s));
      S.rel_residual = 0.0;
      return 0;
    }
    TMX_HIP(fill_zero_bytes(x, sizeof(double) * nv, s));
    TMX_HIP(copy_vec<double>(nv, b, rr64_.get(), s));
    double rr = bb;
    TMX_HIP(set_scalar(ws.sc(), kScRR64, bb, s));
    rc = -3;
    int budget = prm.max_iter;
    double best = 1.0;
    int stalled = 0;
    for (int outer = 0; outer < prm.max_outer; ++outer) {
      S.rel_residual = std::sqrt(rr / bb);
      if (S.rel_residual < prm.rtol) {
        rc = 0;
        break;
      }
      // the FP64 residual has a round-off floor; stop when two steps gain < 2x
      if (outer > 0) {
        stalled = (S.rel_residual > 0.5 * best) ? stalled + 1 : 0;
        best = std::min(best, S.rel_residual);
        if (stalled >= 2) break;
      }
      if (budget <= 0) break;
      float* rin = levels_[0].b32.get();
      TMX_HIP(scale_to<float>(nv, ws.sc(), kScRR64, rr64_.get(), rin, s));
      int it = 0;
      double rel = 0;
      const double inner = std::max(prm.inner_rtol, 0.5 * prm.rtol / S.rel_residual);
      double bbf = 0.0;
      int irc = pcg<float>(ws, rin, inner, budget, &bbf, &it, &rel, s);
      if (g_trace)
        std::fprintf(stderr, "[tmx] refine %d: |r|/|b| = %.3e, inner its %d rel %.3e rc %d\n",
                     outer, S.rel_residual, it, rel, irc);
      S.iterations += it;
      budget -= it;
      S.outer = outer + 1;
      if (irc == -4 && phys_ == kElasticity) {  // K(xi) not positive definite: MINRES
        int it2 = 0;
        double rel2 = 0;
        rc = minres(ws, b, prm.rtol, std::max(1, budget), bb, &it2, &rel2, s);
        S.iterations += it2;
        S.outer = outer + 2;
        if (rc != 0 && rc != -3) return rc;
        TMX_HIP(copy_vec<double>(nv, x64_.get(), x, s));
        S.rel_residual = rel2;
        TMX_HIP(hipStreamSynchronize(s));
        S.solve_ms = ms_since(t0);
        return rc;
      }
      if (irc != 0 && irc != -3) return irc;
      TMX_HIP(axpy_scaled<float>(nv, ws.sc(), kScRR64, d32_.get(), x, s));
      TMX_TRY(residual<double>(x, b, rr64_.get(), s));
      TMX_HIP(dot(nv, rr64_.get(), rr64_.get(), gen, ws.sc() + kScRR64, s));
      TMX_TRY(ws.read(s, kScRR64, 1));
      rr = ws.host[kScRR64];
      if (!std::isfinite(rr)) {
        set_error("non-finite residual in iterative refinement");
        return -5;
      }
    }
    if (rc == -3) {
      S.rel_residual = std::sqrt(rr / bb);
      if (S.rel_residual < prm.rtol) rc = 0;
      else {
        char msg[160];
        std::snprintf(msg, sizeof(msg), "mixed-precision refinement stopped at relative residual "
                      "%.3e above the tolerance %.3e", S.rel_residual, prm.rtol);
        set_error(msg);
      }
    }
  }
  TMX_HIP(hipStreamSynchronize(s));
  S.solve_ms = ms_since(t0);
  return rc;
}

template int Hierarchy::pcg_enqueue_begin<double, double>(Workspace&, const double*, double, bool,
                                                          hipStream_t);
template int Hierarchy::pcg_enqueue_begin<double, float>(Workspace&, const double*, double, bool,
                                                         hipStream_t);
template int Hierarchy::pcg_enqueue_iteration<double, double>(Workspace&, int, hipStream_t);
template int Hierarchy::pcg_enqueue_iteration<double, float>(Workspace&, int, hipStream_t);
template int Hierarchy::pcg_enqueue_end<double>(Workspace&, int, hipStream_t, bool);
template int Hierarchy::apply<double>(const double*, double*, RedSlot*, hipStream_t);
template int Hierarchy::apply_update<double>(const double*, const double*, double*, double*,
                                             const double*, RedSlot*, hipStream_t, const float*,
                                             double*, const double*);
template int Hierarchy::apply<float>(const float*, float*, RedSlot*, hipStream_t);
template double* Hierarchy::x_buf<double>();
template float* Hierarchy::x_buf<float>();
template int Hierarchy::residual<double>(const double*, const double*, double*, hipStream_t);
template int Hierarchy::smooth0<double>(const double*, const double*, double*, hipStream_t);
template int Hierarchy::smooth0<float>(const float*, const float*, float*, hipStream_t);
template int Hierarchy::sweep_pair0<double>(int, int, const double*, const double*, double*,
                                            double, double, RedSlot*, hipStream_t, int);
template int Hierarchy::sweep_pair0<float>(int, int, const float*, const float*, float*, double,
                                           double, RedSlot*, hipStream_t, int);
template int Hierarchy::filter_pipe0<double>(bool, int, hipStream_t);
template int Hierarchy::filter_pipe0<float>(bool, int, hipStream_t);
template int Hierarchy::sweep_pair_rr0<double>(int, const double*, double*, double, double,
                                               hipStream_t, int);
template int Hierarchy::sweep_pair_rr0<float>(int, const float*, float*, double, double,
                                              hipStream_t, int);
template int Hierarchy::vcycle<double>(int, const double*, double*, hipStream_t, RedSlot*);
template int Hierarchy::vcycle<float>(int, const float*, float*, hipStream_t, RedSlot*);

}  // namespace tmx
