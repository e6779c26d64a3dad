// extern "C" boundary of topomax_amd (declared in include/tmx.h).
#include "../../include/tmx.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "tmx_dct.h"
#include "tmx_handle.h"

static const double kDolfinEps = 3.0e-16;

int invalid(const std::string& m) {
  set_error(m);
  return TMX_ERR_INVALID;
}

extern "C" const char* tmx_last_error(void) { return last_error_cstr(); }
extern "C" int tmx_abi_version(void) { return TMX_ABI_VERSION; }

int solve_precision(int p) {
  return p == TMX_FP64 ? kSolveFP64 : p == TMX_MIXED_IR ? kSolveMixedIR
         : p == TMX_MIXED_RU ? kSolveMixedRU : kSolveMixed;
}
int filter_precision(int p) { return p == TMX_FP64 ? kSolveFP64 : kSolveMixed; }

static int create_impl(const tmx_problem_desc* d, tmx_problem* h) {
  if (d->nx < 1 || d->ny < 1) return invalid("nx and ny must be >= 1");
  if (!(d->width > 0) || !(d->height > 0)) return invalid("width and height must be > 0");
  if (!(d->young_modulus > 0)) return invalid("young_modulus must be > 0");
  if (!(d->poisson_ratio > -1.0 && d->poisson_ratio < 0.5))
    return invalid("poisson_ratio must be in (-1, 0.5)");
  if (!(d->penalty_minimum >= 0 && d->penalty_minimum < 1))
    return invalid("penalty_minimum must be in [0, 1)");
  if (!(d->filter_radius >= 0)) return invalid("filter_radius must be >= 0");
  if (d->n_tractions < 0 || d->n_tractions > TMX_MAX_TRACTIONS)
    return invalid("n_tractions out of range");
  for (int i = 0; i < d->n_tractions; ++i)
    if (d->tractions[i].side < 0 || d->tractions[i].side > 3)
      return invalid("Malformed side: " + std::to_string(d->tractions[i].side));
  if ((d->fixed_sides & ~15u) != 0) return invalid("Malformed side in fixed_sides");
  if (d->precision < 0 || d->precision > 4) return invalid("unknown precision");
  if (d->degree < 0 || d->degree > 2) return invalid("degree must be 1 or 2");
  h->desc = *d;
  h->nx = d->nx;
  h->ny = d->ny;
  // dolfin RectangleMesh vertex x = a + ab*ix, ab = (b - a)/nx
  const double ab = (d->width - 0.0) / d->nx, cd = (d->height - 0.0) / d->ny;
  h->hx = ab;
  h->hy = cd;
  h->m = d->penalty_minimum;
  // SidesDomain.inside (FEM_src/domains.py:18-38): a side is fixed iff its coordinate is
  // near (|x - t| < DOLFIN_EPS) the target
  unsigned fx = 0;
  if ((d->fixed_sides & TMX_FIX_LEFT) && std::fabs(0.0 - 0.0) < kDolfinEps) fx |= kLeft;
  if ((d->fixed_sides & TMX_FIX_RIGHT) && std::fabs(0.0 + ab * d->nx - d->width) < kDolfinEps)
    fx |= kRight;
  if ((d->fixed_sides & TMX_FIX_TOP) && std::fabs(0.0 + cd * d->ny - d->height) < kDolfinEps)
    fx |= kTop;
  if ((d->fixed_sides & TMX_FIX_BOTTOM) && std::fabs(0.0 - 0.0) < kDolfinEps) fx |= kBottom;
  h->fixed = fx;
  // Lame parameters as the reference computes them (FEM_src/elasisity_problem.py:98-99)
  h->mu = d->young_modulus / (2 * (1 + d->poisson_ratio));
  h->lam = h->mu * d->poisson_ratio / (0.5 - d->poisson_ratio);
  TMX_HIP(hipGetDevice(&h->device));
  TMX_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  TMX_HIP(hipEventCreateWithFlags(&h->ev_in, hipEventDisableTiming));
  TMX_HIP(hipEventCreateWithFlags(&h->ev_out, hipEventDisableTiming));
  TMX_HIP(hipEventCreate(&h->ev_setup[0]));
  TMX_HIP(hipEventCreate(&h->ev_setup[1]));
  hipStream_t s = h->stream;
  TMX_TRY(h->ws.init(s));
  TMX_TRY(h->el.init(kElasticity, h->nx, h->ny, ab, cd, fx, h->lam, h->mu, 0.0, 64, s));
  TMX_TRY(h->fl.init(kFilter, h->nx, h->ny, ab, cd, 0u, 0.0, 0.0, d->filter_radius, 128, s));
  TMX_TRY(h->fl.build_filter(s));
  h->el.nu_coarse_ = 3;  // with Chebyshev steps a third coarse sweep pays for itself
  // the filter's band kernels make extra sweeps cheap: three on every level measured fastest
  h->fl.nu_pre_ = h->fl.nu_post_ = h->fl.nu_coarse_ = 3;
  // FP64 element level 1 is faster launch per sweep: its band kernels are VALU-bound there
  h->el.band_first_fp64_ = 2;
  const size_t nv = (size_t)(h->nx + 1) * (h->ny + 1);
  TMX_TRY(h->F.alloc(2 * nv));
  TMX_TRY(h->Fm.alloc(2 * nv));
  TMX_TRY(h->lumped.alloc(nv));
  TMX_TRY(h->xi_s.alloc(nv));
  TMX_TRY(h->u_s.alloc(2 * nv));
  TMX_TRY(h->b_s.alloc(nv));
  TMX_TRY(h->w_s.alloc(nv));
  TMX_TRY(h->tmp64.alloc(2 * nv));
  TMX_TRY(h->xi32.alloc(nv));
  TMX_TRY(h->u32.alloc(2 * nv));
  TMX_TRY(h->y32.alloc(2 * nv));
  LoadDesc ld;
  std::memset(&ld, 0, sizeof(ld));
  ld.nx = h->nx;
  ld.ny = h->ny;
  ld.width = d->width;
  ld.height = d->height;
  ld.ab = ab;
  ld.cd = cd;
  ld.has_body = d->has_body_force ? 1 : 0;
  ld.bc[0] = d->body_force.center[0];
  ld.bc[1] = d->body_force.center[1];
  ld.br = d->body_force.radius;
  ld.bv[0] = d->body_force.value[0];
  ld.bv[1] = d->body_force.value[1];
  ld.ntr = d->n_tractions;
  for (int i = 0; i < d->n_tractions; ++i) {
    const tmx_traction& t = d->tractions[i];
    ld.tside[i] = t.side;
    ld.tlo[i] = t.center - t.length / 2;
    ld.thi[i] = t.center + t.length / 2;
    ld.tv[i][0] = t.value[0];
    ld.tv[i][1] = t.value[1];
  }
  TMX_HIP(elast_assemble_load(ld, h->F.get(), s));
  TMX_HIP(copy_vec<double>(2 * nv, h->F.get(), h->Fm.get(), s));
  TMX_HIP(mask_fixed(h->nx, h->ny, fx, h->Fm.get(), s));
  // lumped mass m_i = int phi_i = (M 1)_i
  TMX_HIP(fill(nv, 1.0, h->tmp64.get(), s));
  TMX_HIP(filter_sweep<double>(h->fl.levels_[0].sh64, FilterConst<double>::make(0.0, ab, cd),
                               kApply, h->tmp64.get(), nullptr, h->lumped.get(), 0.0, nullptr, s));
  TMX_HIP(hipStreamSynchronize(s));
  h->degree = d->degree == 2 ? 2 : 1;
  if (h->degree == 2) {
    h->g2 = P2Geom{h->nx, h->ny, ab, cd, h->lam, h->mu, fx};
    const size_t n2 = h->g2.nodes(), ncell = (size_t)h->nx * h->ny;
    TMX_TRY(h->W2.alloc(12 * ncell));
    TMX_TRY(h->W2pc.alloc(12 * ncell));
    TMX_TRY(h->D2.alloc(3 * n2));
    TMX_TRY(h->F2.alloc(2 * n2));
    TMX_TRY(h->F2m.alloc(2 * n2));
    for (auto& v : h->p2v) TMX_TRY(v.alloc(2 * n2));
    TMX_TRY(h->r1c.alloc(2 * nv));
    TMX_TRY(h->e1c.alloc(2 * nv));
    TMX_TRY(h->r1c32.alloc(2 * nv));
    TMX_TRY(h->e1c32.alloc(2 * nv));
    TMX_TRY(h->W2pc32.alloc(12 * ncell));
    TMX_TRY(h->D2_32.alloc(3 * n2));
    for (auto& v : h->p2v32) TMX_TRY(v.alloc(2 * n2));
    TMX_TRY(h->u_s.alloc(2 * n2));
    TMX_HIP(p2_assemble_load(ld, h->F2.get(), s));
    TMX_HIP(copy_vec<double>(2 * n2, h->F2.get(), h->F2m.get(), s));
    TMX_HIP(p2_mask_fixed(h->g2, h->F2m.get(), s));
    TMX_HIP(hipStreamSynchronize(s));
  }
  h->el_prm.rtol = 1e-10;
  h->el_prm.max_iter = 1000;
  h->el_prm.precision = solve_precision(d->precision);
  h->fl_prm.rtol = 1e-12;
  h->fl_prm.max_iter = 500;
  h->fl_prm.precision = filter_precision(d->precision);
  // the filter operator is fixed: bound lambda_max(D^-1 A) for its Chebyshev steps once
  TMX_TRY(h->fl.estimate_lmax(h->ws, s));
  TMX_HIP(hipStreamSynchronize(s));
  return TMX_OK;
}

extern "C" int tmx_problem_create(const tmx_problem_desc* desc, tmx_problem** out) {
  if (!desc || !out) return invalid("null argument");
  *out = nullptr;
  tmx_problem* h = new (std::nothrow) tmx_problem();
  if (!h) return invalid("out of host memory");
  const int rc = create_impl(desc, h);
  if (rc != TMX_OK) {
    delete h;
    return rc;
  }
  *out = h;
  return TMX_OK;
}

extern "C" void tmx_problem_destroy(tmx_problem* h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  delete h;
}

extern "C" int tmx_set_stream(tmx_problem* h, void* stream) {
  if (!h) return invalid("null handle");
  if (h->user != static_cast<hipStream_t>(stream)) ++h->version;
  h->user = static_cast<hipStream_t>(stream);
  return TMX_OK;
}

extern "C" int tmx_problem_info(const tmx_problem* h, int64_t* n_vertices, int64_t* n_dofs,
                                unsigned* fixed_sides_applied, int* levels) {
  if (!h) return invalid("null handle");
  const int64_t nv = (int64_t)(h->nx + 1) * (h->ny + 1);
  if (n_vertices) *n_vertices = nv;
  if (n_dofs) *n_dofs = h->degree == 2 ? 2 * (int64_t)h->g2.nodes() : 2 * nv;
  if (fixed_sides_applied) {
    unsigned f = 0;
    if (h->fixed & kLeft) f |= TMX_FIX_LEFT;
    if (h->fixed & kRight) f |= TMX_FIX_RIGHT;
    if (h->fixed & kTop) f |= TMX_FIX_TOP;
    if (h->fixed & kBottom) f |= TMX_FIX_BOTTOM;
    *fixed_sides_applied = f;
  }
  if (levels) *levels = h->el.levels();
  return TMX_OK;
}

extern "C" int tmx_set_penalization(tmx_problem* h, double p) {
  if (!h) return invalid("null handle");
  if (!(p >= 1.0 && p <= 8.0) || p != std::floor(p)) {
    set_error("penalization must be an integer in [1, 8] (closed-form exact integration)");
    return TMX_ERR_UNSUPPORTED;
  }
  h->p = (int)p;
  h->has_penalty = true;
  ++h->version;
  return TMX_OK;
}

int need_penalty(const tmx_problem* h) {
  if (!h->has_penalty) return invalid("You must set penalization before calling penalizer");
  return TMX_OK;
}

SolveParams merge(const SolveParams& base, const tmx_solve_opts* o) {
  SolveParams p = base;
  if (!o) return p;
  if (o->rtol > 0) p.rtol = o->rtol;
  if (o->max_iter > 0) p.max_iter = o->max_iter;
  if (o->precision >= 0) p.precision = solve_precision(o->precision);
  if (o->inner_rtol > 0) p.inner_rtol = o->inner_rtol;
  if (o->max_outer > 0) p.max_outer = o->max_outer;
  return p;
}

void export_stats(const SolveStats& s, tmx_solve_stats* o) {
  if (!o) return;
  o->iterations = s.iterations;
  o->outer_iterations = s.outer;
  o->rel_residual = s.rel_residual;
  o->setup_ms = s.setup_ms;
  o->solve_ms = s.solve_ms;
  o->levels = s.levels;
}

size_t nverts(const tmx_problem* h) { return (size_t)(h->nx + 1) * (h->ny + 1); }
// displacement unknowns: 2 per P1 vertex, or 2 per P2 lattice node
size_t ndofs(const tmx_problem* h) {
  return h->degree == 2 ? 2 * h->g2.nodes() : 2 * nverts(h);
}
const double* load_F(const tmx_problem* h) { return h->degree == 2 ? h->F2.get() : h->F.get(); }

SolveParams filter_params(const tmx_problem* h, const tmx_solve_opts* o) {
  SolveParams prm = h->fl_prm;
  if (o && o->precision >= 0) prm.precision = filter_precision(o->precision);
  prm.verify = false;  // filter statistics are not reported
  if (!h->fl_rtol_set) {
    // an order of magnitude below the elasticity solve's own tolerance, within [1e-12, 1e-9]:
    // the filtered design then adds nothing visible to the solve's error (at N = 1000 and a
    // 1e-8 solve, J and the gradient move by 1e-12 relative between a 1e-9 and a 1e-10 filter)
    const double el = (o && o->rtol > 0) ? o->rtol : h->el_prm.rtol;
    prm.rtol = std::min(1e-9, std::max(1e-12, 1e-1 * el));
  }
  return prm;
}

static int filter_rhs_solve(tmx_problem* h, const double* b, double* w,
                            const tmx_solve_opts* o = nullptr, bool sync = true) {
  SolveStats st;
  SolveParams prm = filter_params(h, o);
  prm.sync = sync;
  const int rc = h->fl.solve(h->ws, b, w, prm, &st, h->stream);
  if (rc == TMX_ERR_NONFINITE) TMX_HIP(fill(nverts(h), NAN, w, h->stream));
  return rc;
}

int mass_apply(tmx_problem* h, const double* x, double* y) {
  TMX_HIP(filter_sweep<double>(h->fl.levels_[0].sh64, FilterConst<double>::make(0.0, h->hx, h->hy),
                               kApply, x, nullptr, y, 0.0, nullptr, h->stream));
  return TMX_OK;
}

extern "C" int tmx_mass_apply(tmx_problem* h, const double* x, double* y) {
  if (!h || !x || !y) return invalid("null argument");
  Scope scope(h);
  return mass_apply(h, x, y);
}

extern "C" int tmx_filter_apply(tmx_problem* h, const double* rho, double* xi) {
  if (!h || !rho || !xi) return invalid("null argument");
  Scope scope(h);
  TMX_TRY(mass_apply(h, rho, h->b_s.get()));
  return filter_rhs_solve(h, h->b_s.get(), xi);
}

extern "C" int tmx_filter_solve(tmx_problem* h, const double* b, double* w) {
  if (!h || !b || !w) return invalid("null argument");
  Scope scope(h);
  return filter_rhs_solve(h, b, w);
}

// P2 solve of K2(xi) u = F2: MINRES (symmetric; K2 may be indefinite when the filtered design
// dips below zero) preconditioned by a symmetric two-level cycle: nu damped block-Jacobi
// sweeps on the P2 operator of max(xi, 0), the P1 multigrid V-cycle on P1 (the P1 space of
// the same mesh is a subspace of P2, and the P1 operator is the Galerkin restriction of the
// P2 one), nu sweeps back.  MINRES runs in FP64; the cycle in FP64, or under TMX_MIXED in FP32
// (the residual rounded to float on entry, the result widened on exit: weights, diagonal
// blocks, sweeps, transfers and the P1 V-cycle all in float, as the P1 mixed PCG runs its
// V-cycle).
static int p2_solve(tmx_problem* h, const double* xi, double* u, const SolveParams& prm,
                    SolveStats* st) {
  hipStream_t s = h->stream;
  const P2Geom& g = h->g2;
  const size_t n = 2 * g.nodes();
  constexpr int kNu = 2;
  constexpr double kOmega = 0.6;  // lambda_max(D^-1 K) of P2 block Jacobi is about 2.7
  const bool vc32 = prm.precision == kSolveMixed;
  TMX_HIP(p2_weights(g, xi, h->p, h->m, false, h->W2.get(), s));
  TMX_HIP(p2_weights(g, xi, h->p, h->m, true, h->W2pc.get(), s));
  TMX_HIP(p2_diag(g, h->W2pc.get(), h->D2.get(), s));
  if (vc32) {
    const size_t ncell = (size_t)g.nx * g.ny;
    TMX_HIP((convert<double, float>(12 * ncell, h->W2pc.get(), h->W2pc32.get(), s)));
    TMX_HIP((convert<double, float>(3 * g.nodes(), h->D2.get(), h->D2_32.get(), s)));
  }
  DevBuf<double>* P = h->p2v;
  double* X = P[8].get();
  double* ta = P[9].get();
  double* tb = P[10].get();
  MinresBufs bf{P[0].get(), P[1].get(), P[2].get(), P[3].get(), P[4].get(), P[5].get(),
                P[6].get(), P[7].get()};
  auto apply_fn = [&](const double* x, double* y) -> int {
    TMX_HIP(p2_apply(g, h->W2.get(), x, y, s));
    return 0;
  };
  // z = C r for the two-level cycle C in precision T (weights, diagonal blocks and vectors of
  // that precision); the float cycle's first sweep reads the double r (writing r rounded to
  // float for the later steps) and its last writes the double z
  auto cycle = [&](auto tag, const auto* Wp, const auto* Dp, const double* r, double* z,
                   auto* rt, auto* ca, auto* cb, auto* r1, auto* e1) -> int {
    using T = decltype(tag);
    constexpr bool f32 = sizeof(T) == 4;
    T* cur = ca;
    T* nxt = cb;
    if constexpr (f32) TMX_HIP(p2_smooth_first32(g, Wp, Dp, r, rt, cur, kOmega, s));
    else TMX_HIP(p2_smooth<T>(g, Wp, Dp, nullptr, rt, cur, kOmega, s));
    for (int k = 1; k < kNu; ++k) {
      TMX_HIP(p2_smooth<T>(g, Wp, Dp, cur, rt, nxt, kOmega, s));
      std::swap(cur, nxt);
    }
    TMX_HIP(p2_residual<T>(g, Wp, cur, rt, nxt, s));
    TMX_HIP(p2_restrict<T>(g, nxt, r1, s));
    TMX_TRY(h->el.vcycle<T>(0, r1, e1, s));
    TMX_HIP(p2_prolong_add<T>(g, e1, cur, s));
    for (int k = 0; k < kNu; ++k) {
      if (k == kNu - 1) {
        if constexpr (f32) TMX_HIP(p2_smooth_last32(g, Wp, Dp, cur, rt, z, kOmega, s));
        else TMX_HIP(p2_smooth<T>(g, Wp, Dp, cur, rt, z, kOmega, s));
      } else {
        TMX_HIP(p2_smooth<T>(g, Wp, Dp, cur, rt, nxt, kOmega, s));
        std::swap(cur, nxt);
      }
    }
    return 0;
  };
  auto pc_fn = [&](const double* r, double* z, RedSlot* rz) -> int {
    if (vc32)
      TMX_TRY(cycle(float{}, h->W2pc32.get(), h->D2_32.get(), r, z, h->p2v32[0].get(),
                    h->p2v32[2].get(), h->p2v32[3].get(), h->r1c32.get(), h->e1c32.get()));
    else
      TMX_TRY(cycle(double{}, h->W2pc.get(), h->D2.get(), r, z, const_cast<double*>(r), ta, tb,
                    h->r1c.get(), h->e1c.get()));
    TMX_HIP(dot_t<double>(n, r, z, *rz, s));
    return 0;
  };
  RedSlot gen = h->ws.region(3);
  TMX_HIP(dot(n, h->F2m.get(), h->F2m.get(), gen, h->ws.sc() + kScBB, s));
  TMX_TRY(h->ws.read(s, kScBB, 1));
  const double bb = h->ws.host[kScBB];
  st->levels = h->el.levels() + 1;
  st->outer = 1;
  if (!std::isfinite(bb)) {
    set_error("non-finite right-hand side");
    return TMX_ERR_NONFINITE;
  }
  if (bb == 0.0) {
    TMX_HIP(fill_zero_bytes(u, sizeof(double) * n, s));
    st->rel_residual = 0.0;
    return TMX_OK;
  }
  auto t0 = std::chrono::steady_clock::now();
  int it = 0;
  double rel = 0.0;
  // the operator launch also forms MINRES's first Lanczos update and the v.y partials
  MinresApplyL1 apply_l1 = nullptr;
  if (p2_apply_lanczos_ok(g))
    apply_l1 = [&](const double* v, const double* r1, double* y, RedSlot* part,
                   const double* beta, const double* oldb) -> int {
      TMX_HIP(p2_apply_lanczos(g, h->W2.get(), v, r1, y, beta, oldb, part->partials,
                               &part->nblocks, s));
      return 0;
    };
  const int rc = minres_core(h->ws, n, h->F2m.get(), X, bf, apply_fn, pc_fn, prm.rtol,
                             prm.max_iter, bb, &it, &rel, s, std::getenv("TMX_TRACE") != nullptr,
                             apply_l1);
  st->iterations = it;
  st->rel_residual = rel;
  if (rc == TMX_OK || rc == TMX_ERR_NOT_CONVERGED) TMX_HIP(copy_vec<double>(n, X, u, s));
  TMX_HIP(hipStreamSynchronize(s));
  st->solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

// sync = false (P1): returns with the solve's work queued; st->complete() and setup_ms (from
// ev_setup) are then the caller's, after it synchronises the stream
static int elasticity_solve(tmx_problem* h, const double* xi, double* u, const tmx_solve_opts* o,
                            SolveStats* st, bool sync = true) {
  TMX_TRY(need_penalty(h));
  TMX_HIP(hipEventRecord(h->ev_setup[0], h->stream));
  TMX_TRY(h->el.update_elasticity(xi, h->p, h->m, h->stream, h->degree != 2));
  // lambda_max(D^-1 A) of block Jacobi hardly depends on the design (the factors scale D and
  // A alike), so the first operator's estimate (with a margin) serves the later ones
  if (h->el.cheb_auto_ && !h->el.cheb_done_) TMX_TRY(h->el.estimate_lmax(h->ws, h->stream));
  TMX_HIP(hipEventRecord(h->ev_setup[1], h->stream));
  if (h->degree == 2) {
    const int rc = p2_solve(h, xi, u, merge(h->el_prm, o), st);
    float setup = 0.f;
    if (hipEventElapsedTime(&setup, h->ev_setup[0], h->ev_setup[1]) != hipSuccess) setup = -1.f;
    st->setup_ms = setup;
    if (rc == TMX_ERR_NONFINITE) TMX_HIP(fill(ndofs(h), NAN, u, h->stream));
    return rc;
  }
  SolveParams prm = merge(h->el_prm, o);
  prm.sync = sync;
  const int rc = h->el.solve(h->ws, h->Fm.get(), u, prm, st, h->stream);
  if (sync) {
    float setup = 0.f;  // the solve has synchronised the stream past both events
    if (hipEventElapsedTime(&setup, h->ev_setup[0], h->ev_setup[1]) != hipSuccess) setup = -1.f;
    st->setup_ms = setup;
  }
  if (rc == TMX_ERR_NONFINITE) TMX_HIP(fill(ndofs(h), NAN, u, h->stream));
  return rc;
}

extern "C" int tmx_elasticity_solve(tmx_problem* h, const double* xi, double* u,
                                    const tmx_solve_opts* opts, tmx_solve_stats* stats) {
  if (!h || !xi || !u) return invalid("null argument");
  Scope scope(h);
  SolveStats st;
  const int rc = elasticity_solve(h, xi, u, opts, &st);
  export_stats(st, stats);
  return rc;
}

static int read_dot(tmx_problem* h, size_t n, const double* a, const double* b, int slot,
                    double* out) {
  RedSlot r = h->ws.region(3);
  TMX_HIP(dot(n, a, b, r, h->ws.sc() + slot, h->stream));
  TMX_TRY(h->ws.read(h->stream, slot, 1));
  *out = h->ws.host[slot];
  return TMX_OK;
}

extern "C" int tmx_compliance(tmx_problem* h, const double* u, double* out) {
  if (!h || !u || !out) return invalid("null argument");
  Scope scope(h);
  return read_dot(h, ndofs(h), load_F(h), u, kScJ, out);
}

extern "C" int tmx_integrate(tmx_problem* h, const double* values, double* out) {
  if (!h || !values || !out) return invalid("null argument");
  Scope scope(h);
  return read_dot(h, nverts(h), h->lumped.get(), values, kScInt, out);
}

int sensitivity(tmx_problem* h, const double* xi, const double* u, double* b) {
  TMX_TRY(need_penalty(h));
  if (h->degree == 2) {
    TMX_HIP(p2_sensitivity(h->g2, xi, u, h->p, h->m, b, h->stream));
    return TMX_OK;
  }
  TMX_HIP(elast_sensitivity(h->el.levels_[0].sh64, xi, u, h->p, h->m, h->lam, h->mu, h->hx, h->hy,
                            b, h->stream));
  return TMX_OK;
}

extern "C" int tmx_sensitivity(tmx_problem* h, const double* xi, const double* u, double* b) {
  if (!h || !xi || !u || !b) return invalid("null argument");
  Scope scope(h);
  return sensitivity(h, xi, u, b);
}

extern "C" int tmx_objective_and_gradient(tmx_problem* h, const double* rho, double* J,
                                          double* xi, double* u, double* w, double* Mw,
                                          const tmx_solve_opts* opts, tmx_solve_stats* stats) {
  if (!h || !rho) return invalid("null argument");
  TMX_TRY(need_penalty(h));
  Scope scope(h);
  double* xiv = xi ? xi : h->xi_s.get();
  double* uv = u ? u : h->u_s.get();
  // The three solves and the reductions between them are queued back to back: no solve waits
  // for its last kernels, so the device does not idle at the phase changes.  The statistics
  // (true residual, timings) and J are read after the one synchronisation at the end.
  SolveStats st;
  bool st_pending = false;
  auto finish = [&](int code) -> int {
    const hipError_t e = stream_wait(h->stream);
    if (st_pending) {
      st.complete();
      float setup = 0.f;
      if (hipEventElapsedTime(&setup, h->ev_setup[0], h->ev_setup[1]) != hipSuccess) setup = -1.f;
      st.setup_ms = setup;
    }
    export_stats(st, stats);
    if (e != hipSuccess) TMX_HIP(e);
    return code;
  };
  TMX_TRY(mass_apply(h, rho, h->b_s.get()));
  int rc = filter_rhs_solve(h, h->b_s.get(), xiv, opts, false);
  if (rc != TMX_OK && rc != TMX_ERR_NONFINITE) return finish(rc);
  if (rc == TMX_OK) {
    st_pending = h->degree != 2;
    rc = elasticity_solve(h, xiv, uv, opts, &st, false);
  } else {
    TMX_HIP(fill(ndofs(h), NAN, uv, h->stream));
  }
  if (rc != TMX_OK && rc != TMX_ERR_NONFINITE) return finish(rc);
  // J = F.u reduced into a device scalar now, read after the step's final synchronisation
  RedSlot gen = h->ws.region(3);
  TMX_HIP(dot(ndofs(h), load_F(h), uv, gen, h->ws.sc() + kScJ, h->stream,
              h->ws.host_dev + kScJ));
  if (w || Mw) {
    double* wv = w ? w : h->w_s.get();
    TMX_TRY(sensitivity(h, xiv, uv, h->b_s.get()));
    const int rc2 = filter_rhs_solve(h, h->b_s.get(), wv, opts, false);
    if (rc2 != TMX_OK && rc2 != TMX_ERR_NONFINITE) return finish(rc2);
    if (Mw) TMX_TRY(mass_apply(h, wv, Mw));
    if (rc == TMX_OK) rc = rc2;
  }
  const int frc = finish(rc);
  if (frc != rc) return frc;
  if (J) *J = h->ws.host[kScJ];
  if (rc == TMX_OK && h->degree == 1) {
    // every lazily built part of this scheme pair exists now (merged batches rely on it)
    const SolveParams ep = merge(h->el_prm, opts), fp = filter_params(h, opts);
    const bool e_ok = ep.precision == kSolveFP64 || ep.precision == kSolveMixed;
    const bool f_ok = fp.precision == kSolveFP64 || fp.precision == kSolveMixed;
    if (e_ok && f_ok) h->merge_ready[ep.precision][fp.precision] = true;
  }
  return rc;
}

extern "C" int tmx_matvec(tmx_problem* h, const double* xi, const double* u, double* y,
                          int precision) {
  if (!h || !xi || !u || !y) return invalid("null argument");
  TMX_TRY(need_penalty(h));
  if (precision != TMX_FP64 && precision != TMX_FP32)
    return invalid("precision must be TMX_FP64 or TMX_FP32");
  Scope scope(h);
  const size_t nv = nverts(h);
  if (h->degree == 2) {
    if (precision != TMX_FP64) {
      set_error("the P2 operator is applied in FP64 only");
      return TMX_ERR_UNSUPPORTED;
    }
    TMX_HIP(p2_weights(h->g2, xi, h->p, h->m, false, h->W2.get(), h->stream));
    TMX_HIP(p2_apply(h->g2, h->W2.get(), u, y, h->stream));
    return TMX_OK;
  }
  TMX_TRY(h->el.set_fine_operator(xi, h->p, h->m, h->stream));
  if (precision == TMX_FP64) {
    TMX_TRY(h->el.apply<double>(u, y, nullptr, h->stream));
  } else {
    TMX_HIP((convert<double, float>(2 * nv, u, h->u32.get(), h->stream)));
    TMX_TRY(h->el.apply<float>(h->u32.get(), h->y32.get(), nullptr, h->stream));
    TMX_HIP((convert<float, double>(2 * nv, h->y32.get(), y, h->stream)));
  }
  return TMX_OK;
}

extern "C" int tmx_load_vector(tmx_problem* h, double* F) {
  if (!h || !F) return invalid("null argument");
  Scope scope(h);
  TMX_HIP(copy_vec<double>(ndofs(h), load_F(h), F, h->stream));
  return TMX_OK;
}

extern "C" int tmx_set_option(tmx_problem* h, int option, double value) {
  if (!h) return invalid("null handle");
  // an option may change which launches a step makes, or leave parts to build lazily
  ++h->version;
  for (auto& r : h->merge_ready) r[0] = r[1] = false;
  Hierarchy* hs[2] = {&h->el, &h->fl};
  switch (option) {
    case TMX_OPT_FUSED:
      for (auto* x : hs) x->fused_ = value != 0.0;
      break;
    case TMX_OPT_TAIL:
      for (auto* x : hs) {
        x->use_tail_ = value != 0.0;
        if (value > 1.0) {
          x->tail_nodes_ = (int)value;
          x->find_tail();
        }
      }
      break;
    case TMX_OPT_NU_PRE:
    case TMX_OPT_NU_POST:
    case TMX_OPT_NU_COARSE:
    case TMX_OPT_NU_FINE_FILTER:
    case TMX_OPT_NU_COARSE_FILTER:
      if (value < 1 || value > 8 || value != std::floor(value)) return invalid("sweeps must be 1..8");
      if (option == TMX_OPT_NU_PRE) h->el.nu_pre_ = (int)value;
      if (option == TMX_OPT_NU_POST) h->el.nu_post_ = (int)value;
      if (option == TMX_OPT_NU_COARSE) h->el.nu_coarse_ = (int)value;
      if (option == TMX_OPT_NU_FINE_FILTER) h->fl.nu_pre_ = h->fl.nu_post_ = (int)value;
      if (option == TMX_OPT_NU_COARSE_FILTER) h->fl.nu_coarse_ = (int)value;
      break;
    case TMX_OPT_FILTER_RTOL:
      if (!(value > 0 && value < 1)) return invalid("filter rtol must be in (0, 1)");
      h->fl_prm.rtol = value;
      h->fl_rtol_set = true;
      break;
    case TMX_OPT_CHEB_ELAST:
    case TMX_OPT_CHEB_FILTER:
      if (!(value >= 0 && value < 100)) return invalid("Chebyshev bound must be in [0, 100)");
      {
        Hierarchy& x = option == TMX_OPT_CHEB_ELAST ? h->el : h->fl;
        x.cheb_lmax_ = value;
        x.cheb_auto_ = false;
        x.cheb_done_ = true;
      }
      break;
    case TMX_OPT_CHEB_RATIO:
      if (!(value > 1 && value < 1000)) return invalid("Chebyshev ratio must be in (1, 1000)");
      for (auto* x : hs) x->cheb_ratio_ = value;
      break;
    case TMX_OPT_BAND:
      if (value < 0 || value > 64 || value != std::floor(value)) return invalid("band level must be 0..64");
      for (auto* x : hs) x->band_first_ = x->band_first_fp64_ = (int)value;
      break;
    case TMX_OPT_BAND_FINE_FILTER:
      h->fl.band_fine_filter_ = value != 0.0;
      break;
    case TMX_OPT_GRAPHS:
      for (auto* x : hs) x->use_graphs_ = value != 0.0;
      break;
    case TMX_OPT_FILTER_DCT:
      h->fl.drop_graphs();
      h->fl.use_dct_ = value != 0.0;
      break;
    case TMX_OPT_PAIR_LEVELS:
      if (value < 1 || value > 8 || value != std::floor(value))
        return invalid("paired levels must be 1..8");
      h->el.drop_graphs();
      h->el.pair_levels_ = (int)value;
      break;
    case TMX_OPT_PAIR_ROWS_COARSE:
      if (value < 0 || value > 256 || value != std::floor(value))
        return invalid("sweep-pair rows must be 0..256");
      h->el.drop_graphs();
      for (int& r : h->el.pair_rows_coarse_) r = (int)value;
      break;
    case TMX_OPT_BAND_WIDE:
      h->el.drop_graphs();
      if (value != 0.0 && value != 1.0 && value != 2.0 && value != 3.0)
        return invalid("band_wide must be 0, 1, 2 or 3");
      h->el.band_wide_ = (int)value;
      break;
    case TMX_OPT_BENCH_COLD:
      h->bench_cold = value != 0.0;
      break;
    case TMX_OPT_SETUP_OVERLAP:
      TMX_TRY(h->el.join_coarse(h->stream));
      h->el.overlap_setup_ = value != 0.0;
      break;
    case TMX_OPT_RU_EVERY:
      if (value < 1 || value > 1000 || value != std::floor(value))
        return invalid("ru_every must be an integer in [1, 1000]");
      h->el.ru_every_ = (int)value;
      break;
    case TMX_OPT_PAIR_NU_COARSE:
      if (value < 0 || value > 16 || value != std::floor(value))
        return invalid("sweeps must be 0..16");
      h->el.drop_graphs();
      h->el.pair_nu_coarse_ = (int)value;
      break;
    case TMX_OPT_FILTER_PIPE:
      if (value < 0 || value > 256 || value != std::floor(value))
        return invalid("filter pipeline rows must be 0..256");
      h->fl.drop_graphs();
      for (auto& r : h->fl.fpipe_rows_) r[0] = r[1] = (int)value;
      break;
    case TMX_OPT_SWEEP_PAIRS:
    case TMX_OPT_SWEEP_PAIRS_POST:
    case TMX_OPT_SWEEP_PAIR_RESTRICT:
      if (value < 0 || value > 256 || value != std::floor(value))
        return invalid("sweep-pair rows must be 0..256");
      h->el.drop_graphs();
      for (int i = 0; i < 2; ++i)
        (option == TMX_OPT_SWEEP_PAIRS        ? h->el.pair_rows_
         : option == TMX_OPT_SWEEP_PAIRS_POST ? h->el.pair_rows_post_
                                              : h->el.pair_rr_rows_)[i] = (int)value;
      break;
    case TMX_OPT_FINE_CELL_WEIGHTS:
      if (value < 0 || value > 3 || value != std::floor(value)) return invalid("mask must be 0..3");
      h->el.fine_cw_ = (unsigned)value;
      break;
    case TMX_OPT_OMEGA_ELAST:
    case TMX_OPT_OMEGA_FILTER:
      if (!(value > 0 && value < 2)) return invalid("omega must be in (0, 2)");
      (option == TMX_OPT_OMEGA_ELAST ? h->el : h->fl).omega_ = value;
      break;
    default:
      return invalid("unknown option");
  }
  for (auto* x : hs) x->drop_graphs();
  return TMX_OK;
}

extern "C" int tmx_precondition(tmx_problem* h, int which, int precision, const double* xi,
                                const double* r, double* z) {
  if (!h || !r || !z) return invalid("null argument");
  if (which != 0 && which != 1) return invalid("which must be 0 (elasticity) or 1 (filter)");
  if (precision != TMX_FP64 && precision != TMX_FP32) return invalid("bad precision");
  if (which == 0 && !xi) return invalid("xi is required for the elasticity preconditioner");
  if (which == 0 && h->degree == 2) {
    set_error("tmx_precondition applies the P1 elasticity V-cycle; not available for P2");
    return TMX_ERR_UNSUPPORTED;
  }
  if (which == 0) TMX_TRY(need_penalty(h));
  Scope scope(h);
  hipStream_t s = h->stream;
  Hierarchy& hy = which == 0 ? h->el : h->fl;
  if (which == 0) TMX_TRY(h->el.update_elasticity(xi, h->p, h->m, s));
  const size_t n = hy.n();
  if (precision == TMX_FP64) {
    TMX_HIP(copy_vec<double>(n, r, hy.r64_.get(), s));
    TMX_TRY(hy.vcycle<double>(0, hy.r64_.get(), hy.z64_.get(), s));
    TMX_HIP(copy_vec<double>(n, hy.z64_.get(), z, s));
  } else {
    TMX_HIP((convert<double, float>(n, r, hy.r32_.get(), s)));
    TMX_TRY(hy.vcycle<float>(0, hy.r32_.get(), hy.z32_.get(), s));
    TMX_HIP((convert<float, double>(n, hy.z32_.get(), z, s)));
  }
  return TMX_OK;
}

extern "C" int tmx_set_tiles(tmx_problem* h, int rows64, int rows32) {
  if (!h) return invalid("null handle");
  ++h->version;
  h->el.set_rows(rows64, rows32);
  h->fl.set_rows(rows64, rows32);
  return TMX_OK;
}

extern "C" int tmx_autotune(tmx_problem* h, int* rows_out, double* best_us_out) {
  if (!h) return invalid("null handle");
  Scope scope(h);
  // times the fine operator with xi = 0.5 and p = current (or 3)
  const size_t nv = nverts(h);
  const int p = h->has_penalty ? h->p : 3;
  hipStream_t s = h->stream;
  TMX_HIP(fill(nv, 0.5, h->xi_s.get(), s));
  TMX_TRY(h->el.set_fine_operator(h->xi_s.get(), p, h->m, s));
  TMX_HIP(fill(2 * nv, 1.0, h->u_s.get(), s));
  TMX_HIP((convert<double, float>(2 * nv, h->u_s.get(), h->u32.get(), s)));
  hipEvent_t e0, e1;
  TMX_HIP(hipEventCreate(&e0));
  TMX_HIP(hipEventCreate(&e1));
  const int cands[] = {2, 4, 8, 16, 32, 64};
  int best[2] = {0, 0};
  double best_us[2] = {1e30, 1e30};
  for (int prec = 0; prec < 2; ++prec) {
    for (int rows : cands) {
      if (prec == 0) h->el.set_rows(rows, 0);
      else h->el.set_rows(0, rows);
      const int reps = 20;
      for (int warm = 0; warm < 3; ++warm) {
        if (prec == 0) TMX_TRY(h->el.apply<double>(h->u_s.get(), h->tmp64.get(), nullptr, s));
        else TMX_TRY(h->el.apply<float>(h->u32.get(), h->y32.get(), nullptr, s));
      }
      TMX_HIP(hipEventRecord(e0, s));
      for (int r = 0; r < reps; ++r) {
        if (prec == 0) TMX_TRY(h->el.apply<double>(h->u_s.get(), h->tmp64.get(), nullptr, s));
        else TMX_TRY(h->el.apply<float>(h->u32.get(), h->y32.get(), nullptr, s));
      }
      TMX_HIP(hipEventRecord(e1, s));
      TMX_HIP(hipEventSynchronize(e1));
      float ms = 0;
      TMX_HIP(hipEventElapsedTime(&ms, e0, e1));
      const double us = 1000.0 * ms / reps;
      if (us < best_us[prec]) {
        best_us[prec] = us;
        best[prec] = rows;
      }
    }
  }
  // paired sweeps of the elasticity V-cycle: chunk height per kernel and precision (the launch
  // is latency-bound, so the best height is the smallest whose workgroups still fit the
  // device at once, which depends on the grid)
  TMX_TRY(h->el.fine_cell_weights(s));
  const int pair_cands[] = {2, 3, 4, 6, 8, 12, 16, 24, 32, 48};
  for (int prec = 0; prec < 2; ++prec) {
    for (int kind = 0; kind < 3; ++kind) {
      if (kind >= 1 && !h->el.pair_post_ok()) continue;
      int& slot = (kind == 0   ? h->el.pair_rows_
                   : kind == 1 ? h->el.pair_rows_post_
                               : h->el.pair_rr_rows_)[prec == 0 ? 1 : 0];
      if (slot <= 0) continue;  // switched off by the caller
      const int op1 = kind == 0 ? kSmoothZero : kProlongSmooth;
      auto launch = [&](int rows) -> int {
        if (kind == 2) {
          if (prec == 0)
            return h->el.sweep_pair_rr0<double>(rows, h->Fm.get(), h->tmp64.get(), 0.5, 0.5, s);
          return h->el.sweep_pair_rr0<float>(rows, h->y32.get(), h->el.q32_.get(), 0.5, 0.5, s);
        }
        if (prec == 0)
          return h->el.sweep_pair0<double>(op1, rows, h->u_s.get(), h->Fm.get(), h->tmp64.get(),
                                            0.5, 0.5, nullptr, s);
        return h->el.sweep_pair0<float>(op1, rows, h->u32.get(), h->y32.get(), h->el.q32_.get(),
                                        0.5, 0.5, nullptr, s);
      };
      double best_pair = 1e30;
      for (int rows : pair_cands) {
        const int reps = 20;
        for (int warm = 0; warm < 2; ++warm) TMX_TRY(launch(rows));
        TMX_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) TMX_TRY(launch(rows));
        TMX_HIP(hipEventRecord(e1, s));
        TMX_HIP(hipEventSynchronize(e1));
        float ms = 0;
        TMX_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_pair) {
          best_pair = ms;
          slot = rows;
        }
      }
    }
  }
  // the filter's finest-level pipelines: chunk height per direction and precision
  const Hierarchy& fl = h->fl;
  const bool fl_pipes = fl.levels_.size() >= 2 && fl.levels_[1].nested_from_prev &&
                        fl.nu_pre_ <= kMaxPipeNu && fl.nu_post_ <= kMaxPipeNu;
  const int pipe_cands[] = {4, 6, 8, 10, 12, 16, 20, 24};
  for (int prec = 0; prec < 2 && fl_pipes; ++prec) {
    for (int dir = 0; dir < 2; ++dir) {
      int& slot = h->fl.fpipe_rows_[prec == 0 ? 1 : 0][dir];
      if (slot <= 0) continue;  // switched off by the caller
      double best_pipe = 1e30;
      for (int rows : pipe_cands) {
        auto launch = [&]() -> int {
          return prec == 0 ? h->fl.filter_pipe0<double>(dir == 0, rows, s)
                           : h->fl.filter_pipe0<float>(dir == 0, rows, s);
        };
        const int reps = 20;
        for (int warm = 0; warm < 2; ++warm) TMX_TRY(launch());
        TMX_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) TMX_TRY(launch());
        TMX_HIP(hipEventRecord(e1, s));
        TMX_HIP(hipEventSynchronize(e1));
        float ms = 0;
        TMX_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_pipe) {
          best_pipe = ms;
          slot = rows;
        }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  h->el.set_rows(best[0], best[1]);
  h->fl.set_rows(best[0], best[1]);
  if (rows_out) {
    rows_out[0] = best[0];
    rows_out[1] = best[1];
  }
  if (best_us_out) {
    best_us_out[0] = best_us[0];
    best_us_out[1] = best_us[1];
  }
  return TMX_OK;
}

extern "C" int tmx_kernel_timing(tmx_problem* h, int kernel, int enable) {
  if (!h) return invalid("null handle");
  if (kernel != 5) return invalid("in-step timing covers kernel 5 (PCG direction update)");
  h->el.au_timing_ = enable != 0;
  h->el.au_used_ = 0;
  return TMX_OK;
}

extern "C" int tmx_kernel_timing_read(tmx_problem* h, int kernel, double* total_us,
                                      int* launches) {
  if (!h || !total_us || !launches) return invalid("null argument");
  if (kernel != 5) return invalid("in-step timing covers kernel 5 (PCG direction update)");
  Scope scope(h);
  return h->el.au_read(total_us, launches);
}

extern "C" int tmx_bench_kernel(tmx_problem* h, int kernel, int precision, int reps,
                                double* avg_us, double* bytes) {
  if (!h || !avg_us || reps < 1) return invalid("bad arguments");
  if (kernel < 0 || kernel > 6) return invalid("unknown kernel id");
  if (kernel == 5 && precision != TMX_FP64) return invalid("kernel 5 runs with FP64 vectors");
  if (precision != TMX_FP64 && precision != TMX_FP32) return invalid("bad precision");
  const bool elastic = kernel != 2;
  if (elastic && !h->el.has_operator_) return invalid("no elasticity operator yet: solve first");
  if (elastic && h->degree == 2) {
    set_error("tmx_bench_kernel times the P1 elasticity kernels; not available for P2");
    return TMX_ERR_UNSUPPORTED;
  }
  Scope scope(h);
  const size_t nv = nverts(h);
  hipStream_t s = h->stream;
  const bool f64 = precision == TMX_FP64;
  const double sz = f64 ? 8.0 : 4.0;
  // inputs: u / x = current displacement scratch, b = masked load
  TMX_HIP(fill(2 * nv, 1.0, h->u_s.get(), s));
  TMX_HIP((convert<double, float>(2 * nv, h->u_s.get(), h->u32.get(), s)));
  TMX_HIP((convert<double, float>(2 * nv, h->Fm.get(), h->y32.get(), s)));
  auto launch = [&]() -> int {
    if (kernel == 0) {
      if (f64) return h->el.apply<double>(h->u_s.get(), h->tmp64.get(), nullptr, s);
      return h->el.apply<float>(h->u32.get(), h->el.q32_.get(), nullptr, s);
    }
    if (kernel == 1) {
      if (f64) return h->el.smooth0<double>(h->u_s.get(), h->Fm.get(), h->tmp64.get(), s);
      return h->el.smooth0<float>(h->u32.get(), h->y32.get(), h->el.q32_.get(), s);
    }
    if (kernel == 5) {  // the mixed PCG's direction update + operator, on its own buffers
      Hierarchy& e = h->el;
      return e.apply_update<double>(e.p64_.get(), nullptr, e.q64_.get(), e.pb64_.get(),
                                    h->ws.sc() + kPcg.beta, nullptr, s, e.z32_.get(),
                                    e.x64_.get(), h->ws.sc() + kPcg.alpha);
    }
    if (kernel == 6) {  // zero-start pair with residual + restriction
      if (!h->el.pair_post_ok()) return invalid("kernel 6 needs even nx and ny");
      const int rows = h->el.pair_rr_rows_[f64 ? 1 : 0];
      if (f64)
        return h->el.sweep_pair_rr0<double>(rows, h->Fm.get(), h->tmp64.get(), 0.5, 0.5, s);
      return h->el.sweep_pair_rr0<float>(rows, h->y32.get(), h->el.q32_.get(), 0.5, 0.5, s);
    }
    if (kernel >= 3) {  // paired sweeps with the handle's chunk heights
      const int op1 = kernel == 3 ? kSmoothZero : kProlongSmooth;
      const int rows = (kernel == 3 ? h->el.pair_rows_ : h->el.pair_rows_post_)[f64 ? 1 : 0];
      if (f64)
        return h->el.sweep_pair0<double>(op1, rows, h->u_s.get(), h->Fm.get(), h->tmp64.get(),
                                         0.5, 0.5, nullptr, s);
      return h->el.sweep_pair0<float>(op1, rows, h->u32.get(), h->y32.get(), h->el.q32_.get(),
                                      0.5, 0.5, nullptr, s);
    }
    if (f64) return h->fl.apply<double>(h->u_s.get(), h->tmp64.get(), nullptr, s);
    return h->fl.apply<float>(h->u32.get(), h->el.q32_.get(), nullptr, s);
  };
  for (int i = 0; i < 3; ++i) TMX_TRY(launch());
  hipEvent_t e0, e1;
  TMX_HIP(hipEventCreate(&e0));
  TMX_HIP(hipEventCreate(&e1));
  float ms = 0;
  if (!h->bench_cold) {
    TMX_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i) TMX_TRY(launch());
    TMX_HIP(hipEventRecord(e1, s));
    TMX_HIP(hipEventSynchronize(e1));
    TMX_HIP(hipEventElapsedTime(&ms, e0, e1));
  } else {
    // each launch timed alone after reading 512 MiB, twice the 256 MB Infinity Cache, so
    // the kernel finds none of its inputs in the caches (as inside a solve, where the
    // V-cycle's traffic separates its launches); reads leave no dirty lines to write back
    // during the timed launch
    DevBuf<double> evict;
    const size_t ne = (size_t)64 << 20;
    TMX_TRY(evict.alloc(ne));
    TMX_HIP(fill(ne, 1.0, evict.get(), s));
    TMX_HIP(hipStreamSynchronize(s));
    RedSlot er = h->ws.region(3);
    for (int i = 0; i < reps; ++i) {
      TMX_HIP(dot(ne, evict.get(), evict.get(), er, h->ws.sc() + kScTmp, s));
      TMX_HIP(hipEventRecord(e0, s));
      TMX_TRY(launch());
      TMX_HIP(hipEventRecord(e1, s));
      TMX_HIP(hipEventSynchronize(e1));
      float one = 0;
      TMX_HIP(hipEventElapsedTime(&one, e0, e1));
      ms += one;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_us = 1000.0 * ms / reps;
  if (bytes) {
    // algorithmic bytes per node: apply u (2) + y (2) + xi (1); sweep adds b (2); filter x + y.
    // A sweep reading the preconditioner's per-cell weights reads 2 values per cell instead of
    // xi per node.
    // A pair reads b and writes y (zero start) or also reads x (prolongation; the coarse
    // correction is a quarter-size stream).
    const double per = kernel == 0 ? 5.0 : kernel == 1 ? 7.0 : kernel == 2 ? 2.0 :
                       (kernel == 3 || kernel == 6) ? 5.0 : 7.0;
    *bytes = per * sz * (double)nv;
    const bool cw = h->el.has_cw0_ && (h->el.fine_cw_ & (f64 ? 2u : 1u));
    if (cw && (kernel == 1 || kernel == 3 || kernel == 4 || kernel == 6))
      *bytes += 2.0 * sz * (double)h->nx * (double)h->ny - sz * (double)nv;
    // the coarse stream: correction read (4) or restricted residual written (6)
    if (kernel == 4 || kernel == 6)
      *bytes += 2.0 * sz * (double)(h->nx / 2 + 1) * (double)(h->ny / 2 + 1);
    // direction update + operator per vertex: p_old (2 doubles), z (2 floats), xi, x read and
    // written, p and q written: 96 B
    if (kernel == 5) *bytes = 96.0 * (double)nv;
  }
  return TMX_OK;
}

// ------------------------------------------------------------------ entropic mirror descent
// The design-space half of src/solver.py:149-302 on the device: the latent step, the volume
// integrals of the projection and the design change.  The scalar root-finding around
// tmx_emd_volume stays on the host.
extern "C" int tmx_emd_half_step(tmx_problem* h, const double* psi, const double* w, double step,
                                 double* half) {
  if (!h || !psi || !w || !half) return invalid("null argument");
  Scope scope(h);
  TMX_HIP(emd_half_step(nverts(h), psi, w, step, half, h->stream));
  return TMX_OK;
}

extern "C" int tmx_emd_volume(tmx_problem* h, const double* half, double c, double* vol,
                              double* dvol) {
  if (!h || !half || !vol || !dvol) return invalid("null argument");
  Scope scope(h);
  RedSlot r0 = h->ws.region(2), r1 = h->ws.region(3);
  TMX_HIP(emd_volume(nverts(h), half, c, h->lumped.get(), r0, r1, h->ws.sc() + kScInt,
                     h->ws.sc() + kScTmp, h->stream));
  TMX_TRY(h->ws.read(h->stream, kScInt, 2));
  *vol = h->ws.host[kScInt];
  *dvol = h->ws.host[kScTmp];
  return TMX_OK;
}

extern "C" int tmx_emd_apply(tmx_problem* h, const double* half, double c, double* psi,
                             double* rho) {
  if (!h || !half || !psi || !rho) return invalid("null argument");
  Scope scope(h);
  TMX_HIP(emd_apply(nverts(h), half, c, psi, rho, h->stream));
  return TMX_OK;
}

extern "C" int tmx_emd_change(tmx_problem* h, const double* rho, const double* psi_prev,
                              double* out) {
  if (!h || !rho || !psi_prev || !out) return invalid("null argument");
  Scope scope(h);
  RedSlot r = h->ws.region(3);
  TMX_HIP(emd_change(nverts(h), rho, psi_prev, h->lumped.get(), r, h->ws.sc() + kScInt,
                     h->stream));
  TMX_TRY(h->ws.read(h->stream, kScInt, 1));
  *out = std::sqrt(h->ws.host[kScInt]);
  return TMX_OK;
}

extern "C" int tmx_logit(tmx_problem* h, const double* rho, double* psi) {
  if (!h || !rho || !psi) return invalid("null argument");
  Scope scope(h);
  TMX_HIP(logit(nverts(h), rho, psi, h->stream));
  return TMX_OK;
}

// ------------------------------------------------------------------ design transfer
extern "C" int tmx_interpolate_design(tmx_problem* h, int coarse_nx, int coarse_ny,
                                      const double* rho_coarse, double* rho) {
  if (!h || !rho_coarse || !rho) return invalid("null argument");
  if (coarse_nx < 1 || coarse_ny < 1) return invalid("coarse nx and ny must be >= 1");
  Scope scope(h);
  TMX_HIP(interp_p1(coarse_nx, coarse_ny, h->desc.width / coarse_nx, h->desc.height / coarse_ny,
                    h->nx, h->ny, h->hx, h->hy, rho_coarse, rho, h->stream));
  return TMX_OK;
}
