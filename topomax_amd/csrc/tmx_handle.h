// The tmx_problem handle and the helpers that entry points in more than one source file
// share (tmx_api.hip, tmx_batch.hip).  Internal to libtmx.so.
#pragma once

#include "../../include/tmx.h"

#include <string>

#include "tmx_p2.h"
#include "tmx_solver.h"

namespace tmx {
const char* last_error_cstr();
}

using namespace tmx;

struct tmx_problem {
  tmx_problem_desc desc;
  hipStream_t stream = nullptr;  // internal non-blocking stream (all work, graph capture)
  hipStream_t user = nullptr;    // caller's stream (tmx_set_stream), joined by events
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipEvent_t ev_setup[2] = {nullptr, nullptr};  // brackets the operator rebuild (timing)
  int device = 0;  // HIP device current at creation; all of the handle's memory lives there
  int nx = 0, ny = 0;
  double hx = 0, hy = 0, lam = 0, mu = 0, m = 1e-6;
  unsigned fixed = 0;  // sides actually fixed after the near() test
  int p = 0;
  bool has_penalty = false;
  Hierarchy el, fl;
  Workspace ws;
  DevBuf<double> F, Fm, lumped;
  DevBuf<double> xi_s, u_s, b_s, w_s;     // scratch of the fused call
  DevBuf<double> tmp64;
  DevBuf<float> xi32, u32, y32;
  SolveParams el_prm, fl_prm;
  bool fl_rtol_set = false;  // filter tolerance given explicitly (else derived per solve)
  bool bench_cold = false;   // tmx_bench_kernel: evict the caches before each timed launch
  // P2 element (degree 2): lattice geometry, element weights of the operator (W2) and of the
  // preconditioner (W2pc, from max(xi, 0)), P2 block diagonal, loads and solver vectors
  int degree = 1;
  P2Geom g2{};
  DevBuf<double> W2, W2pc, D2, F2, F2m, p2v[11], r1c, e1c;
  // the mixed scheme's FP32 preconditioner: the P1 correction's vectors, the P2 sweeps'
  // weights and diagonal blocks, and their vectors (r, z and two iterates)
  DevBuf<float> r1c32, e1c32, W2pc32, D2_32, p2v32[4];
  // Merged batches (tmx_batch.hip): `version` changes whenever an entry point may have changed
  // the handle's solver state, so a batch re-records its launches; merge_ready[e][f] is set
  // once an objective step with elasticity scheme e and filter scheme f (kSolveFP64 or
  // kSolveMixed) has completed on this handle, which leaves every lazily built part (smoother
  // bound, filter preconditioner) in place, and cleared by any option change.
  unsigned long long version = 0;
  bool merge_ready[2][2] = {{false, false}, {false, false}};
  ~tmx_problem() {
    if (ev_in) (void)hipEventDestroy(ev_in);
    for (auto e : ev_setup)
      if (e) (void)hipEventDestroy(e);
    if (ev_out) (void)hipEventDestroy(ev_out);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// Orders an entry point's work after the caller's stream and the caller's stream after it.
class Scope {
 public:
  explicit Scope(tmx_problem* h) : h_(h) {
    if (hipEventRecord(h->ev_in, h->user) == hipSuccess) (void)hipStreamWaitEvent(h->stream, h->ev_in, 0);
  }
  ~Scope() {
    if (hipEventRecord(h_->ev_out, h_->stream) == hipSuccess)
      (void)hipStreamWaitEvent(h_->user, h_->ev_out, 0);
    ++h_->version;
  }

 private:
  tmx_problem* h_;
};


int invalid(const std::string& m);
int need_penalty(const tmx_problem* h);
// tmx_precision -> elasticity solve scheme (TMX_FP32 solves run as TMX_MIXED)
int solve_precision(int p);
// the filter solves in FP64 PCG, with an FP32 V-cycle unless FP64 is asked for
int filter_precision(int p);
SolveParams merge(const SolveParams& base, const tmx_solve_opts* o);
// the filter solve's parameters for a call with options o (tolerance derived from the
// elasticity tolerance unless set explicitly)
SolveParams filter_params(const tmx_problem* h, const tmx_solve_opts* o);
void export_stats(const SolveStats& s, tmx_solve_stats* o);
size_t nverts(const tmx_problem* h);
size_t ndofs(const tmx_problem* h);
const double* load_F(const tmx_problem* h);
// y = M x on the handle's stream
int mass_apply(tmx_problem* h, const double* x, double* y);
// b = the sensitivity of the compliance (P1 or P2) on the handle's stream
int sensitivity(tmx_problem* h, const double* xi, const double* u, double* b);
