/* topomax_amd C ABI: the elasticity hot path of topomax on one MI355X (gfx950).
 *
 * One handle = one ElasticityProblem on a structured W x H rectangle meshed with nx x ny
 * cells split by the 'right' diagonal (P1 displacement, CG1 design), with its Helmholtz
 * filter, SIMP penalizer, load vector and Dirichlet sides.  All array arguments are DEVICE
 * pointers (e.g. torch tensor data_ptr()) in vertex order v = iy*(nx+1) + ix; displacements
 * are node-interleaved (u[2v], u[2v+1]).  All work is enqueued on the handle's stream
 * (tmx_set_stream); functions returning a host scalar synchronise that stream.
 *
 * Return codes: TMX_OK, or a negative TMX_ERR_*; tmx_last_error() gives the message of the
 * last failure on the calling thread.  Misuse maps to the reference's ValueError, numerical
 * or HIP failures to RuntimeError (see the Python wrapper).
 *
 * Reference interfaces replaced (topomax repository, relative paths):
 *   ElasticityProblem  FEM_src/elasisity_problem.py:76-196
 *   HelmholtzFilter    FEM_src/filter.py:8-41
 *   SmartMumpsSolver   FEM_src/pde_solver.py:75-133
 *   ElasticPenalizer   src/penalizers.py:29-45
 *   Problem ABC        src/problem.py:7-23
 *   FEMSolver.integrate FEM_src/solver.py:81-84
 */
#ifndef TMX_H
#define TMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TMX_ABI_VERSION 2

enum tmx_status {
  TMX_OK = 0,
  TMX_ERR_INVALID = -1,       /* misuse: bad argument, call order, unset penalty */
  TMX_ERR_HIP = -2,           /* HIP runtime failure */
  TMX_ERR_NOT_CONVERGED = -3, /* iteration limit reached before the tolerance */
  TMX_ERR_BREAKDOWN = -4,     /* operator not SPD (e.g. zero stiffness) */
  TMX_ERR_NONFINITE = -5,     /* non-finite data; outputs are filled with NaN */
  TMX_ERR_UNSUPPORTED = -6    /* valid for the reference but not implemented here */
};

enum tmx_side { TMX_LEFT = 0, TMX_RIGHT = 1, TMX_TOP = 2, TMX_BOTTOM = 3 };
enum tmx_side_bits { TMX_FIX_LEFT = 1, TMX_FIX_RIGHT = 2, TMX_FIX_TOP = 4, TMX_FIX_BOTTOM = 8 };
/* Solve schemes.  TMX_FP64: FP64 multigrid-PCG.  TMX_MIXED: FP64 PCG (FP64 operator, residual
 * and iterates) preconditioned by an FP32 V-cycle.  TMX_MIXED_IR: FP32 multigrid-PCG inner
 * solves inside FP64 iterative refinement.  TMX_MIXED_RU: one FP32 multigrid-PCG (FP32
 * operator, Krylov vectors and V-cycle) with the iterate accumulated in FP64 and the FP32
 * residual replaced by the true FP64 one b - A x every few iterations (reliable updates; the
 * search direction is kept), finished when the FP64 residual meets the tolerance.  TMX_FP32
 * names single-precision kernels (tmx_matvec, tmx_precondition, tmx_bench_kernel); as a solve
 * precision it means TMX_MIXED.  Filter solves run in FP64 PCG with an FP32 V-cycle unless
 * TMX_FP64 is asked for. */
enum tmx_precision { TMX_FP64 = 0, TMX_MIXED = 1, TMX_FP32 = 2, TMX_MIXED_IR = 3,
                     TMX_MIXED_RU = 4 };

#define TMX_MAX_TRACTIONS 16

/* designs/definitions.py:150-166 Traction */
typedef struct tmx_traction {
  int side;          /* enum tmx_side */
  double center, length;
  double value[2];
} tmx_traction;

/* designs/definitions.py:137-146 Force (+ CircularRegion :90-97) */
typedef struct tmx_body_force {
  double center[2];
  double radius;
  double value[2];
} tmx_body_force;

/* ElasticityProblem.__init__ inputs: mesh (FEM_src/solver.py:38-47), DomainParameters and
 * ElasticityParameters (designs/definitions.py:170-223), penalizer minimum
 * (src/penalizers.py:34). */
typedef struct tmx_problem_desc {
  int nx, ny;                 /* cells: int(W*N), int(H*N) */
  double width, height;
  double young_modulus, poisson_ratio;
  double penalty_minimum;     /* m in r = m + x^p (1-m); the reference uses 1e-6 */
  double filter_radius;       /* Helmholtz epsilon, used unscaled */
  unsigned fixed_sides;       /* requested sides, TMX_FIX_* bits (near-test applied inside) */
  int n_tractions;
  tmx_traction tractions[TMX_MAX_TRACTIONS];
  int has_body_force;
  tmx_body_force body_force;
  int precision;              /* default tmx_precision for solves */
  int degree;                 /* displacement element: 1 (P1, default when 0) or 2 (P2, the
                                 reference's VectorElement("CG", triangle, 2),
                                 FEM_src/elasisity_problem.py:194-196); P2 solves run in FP64
                                 (MINRES preconditioned by P2 block Jacobi over the P1
                                 multigrid) and u has 2 (2nx+1)(2ny+1) entries */
} tmx_problem_desc;

typedef struct tmx_solve_opts {
  double rtol;                /* relative residual |b - A u| / |b| */
  int max_iter;               /* PCG iterations (inner, summed over refinements) */
  int precision;              /* -1: handle default, else tmx_precision */
  double inner_rtol;          /* mixed: reduction asked of each FP32 inner solve */
  int max_outer;              /* mixed: refinement steps */
} tmx_solve_opts;

typedef struct tmx_solve_stats {
  int iterations;             /* PCG iterations (all inner solves) */
  int outer_iterations;       /* refinement steps (mixed), 1 for FP64 */
  double rel_residual;        /* final |b - A u| / |b| */
  double setup_ms, solve_ms;  /* operator rebuild (device time on the solve's stream; with
                                 TMX_OPT_SETUP_OVERLAP the coarse part runs beside the solve
                                 and is not in it) / solve (host clock) */
  int levels;                 /* multigrid levels */
} tmx_solve_stats;

typedef struct tmx_problem tmx_problem;

const char* tmx_last_error(void);
int tmx_abi_version(void);

/* ElasticityProblem.__init__ (FEM_src/elasisity_problem.py:79-102): mesh, marker, BCs,
 * HelmholtzFilter (A assembled once, filter.py:15-38), load vector assembled once
 * (pde_solver.py:102-104).  All device workspace is allocated here. */
int tmx_problem_create(const tmx_problem_desc* desc, tmx_problem** out);
void tmx_problem_destroy(tmx_problem* h);
int tmx_set_stream(tmx_problem* h, void* hip_stream);
int tmx_problem_info(const tmx_problem* h, int64_t* n_vertices, int64_t* n_dofs,
                     unsigned* fixed_sides_applied, int* levels);

/* FEMProblem.set_penalization (FEM_src/problem.py:29-35) -> Penalizer.set_penalization
 * (src/penalizers.py:11-12).  Integer p in [1, 8]. */
int tmx_set_penalization(tmx_problem* h, double p);

/* HelmholtzFilter.apply with a CG1 input (filter.py:40-41): solve A xi = M rho. */
int tmx_filter_apply(tmx_problem* h, const double* rho, double* xi);
/* HelmholtzFilter.apply with a form input whose vector b is already assembled: A w = b. */
int tmx_filter_solve(tmx_problem* h, const double* b, double* w);
/* y = M x (CG1 mass); M w is the Euclidean gradient from the L2 representer w. */
int tmx_mass_apply(tmx_problem* h, const double* x, double* y);

/* ElasticityProblem.forward (:168-169) -> SmartMumpsSolver.solve(a_arg=xi): K(xi) u = F
 * with zero Dirichlet dofs.  opts/stats may be NULL. */
int tmx_elasticity_solve(tmx_problem* h, const double* xi, double* u, const tmx_solve_opts* opts,
                         tmx_solve_stats* stats);
/* compliance J = int u.f dx + int_ds u.t = F.u (:161-164) */
int tmx_compliance(tmx_problem* h, const double* u, double* out_host);
/* sensitivity rhs b_i = int -r'(xi) sigma0(u):eps(u) phi_i (:146-149) */
int tmx_sensitivity(tmx_problem* h, const double* xi, const double* u, double* b);
/* calculate_objective + calculate_objective_gradient (:134-166) fused:
 * xi = F(rho), u = K(xi)^-1 F, J = F.u, w = A^-1 b(xi, u), Mw = M w.
 * xi, u, w, Mw may be NULL when not wanted (xi/u use handle scratch then). */
int tmx_objective_and_gradient(tmx_problem* h, const double* rho, double* J_host, double* xi,
                               double* u, double* w, double* Mw, const tmx_solve_opts* opts,
                               tmx_solve_stats* stats);
/* y = A(xi) u for the Dirichlet-masked stiffness (identity on fixed dofs). precision:
 * TMX_FP64 or TMX_FP32 (inputs/outputs stay double). */
int tmx_matvec(tmx_problem* h, const double* xi, const double* u, double* y, int precision);
/* load vector F (2 n_v) as assembled at create time (before Dirichlet masking) */
int tmx_load_vector(tmx_problem* h, double* F);
/* FEMSolver.integrate (FEM_src/solver.py:81-84): int of the CG1 interpolant of values */
int tmx_integrate(tmx_problem* h, const double* values, double* out_host);
/* tmx_integrate without a handle and without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops).  The nx x ny 'right' mesh of the width x height rectangle is given
 * directly; the lumped mass is formed inside the kernel and the block partial sums use a
 * static 32 KiB device buffer that each GPU holds until the process ends, so at most
 * 8,388,608 vertices ((nx+1)(ny+1)) per call; larger meshes return TMX_ERR_INVALID.
 * values: n_v device doubles (vertex order); out: ONE device double, written on `stream`
 * (a hipStream_t, NULL for the default stream); the call does not synchronise.  Calls that
 * overlap on one GPU race on the shared buffer. */
int tmx_integrate_noalloc(int nx, int ny, double width, double height, const double* values,
                          double* out, void* stream);

/* Autotuner: times the FP64 and FP32 elasticity operator sweeps on this problem's fine grid
 * for each candidate row-chunk size and keeps the fastest; results are written to
 * rows_out[2] (fp64, fp32) and best_us_out[2].  It also picks the chunk heights of the paired
 * sweeps (TMX_OPT_SWEEP_PAIRS, _POST, _RESTRICT) per precision. */
int tmx_autotune(tmx_problem* h, int* rows_out, double* best_us_out);
/* Average duration (microseconds, HIP events on the handle's stream) of `reps` back-to-back
 * launches of one fine-grid kernel with the operator of the last solve:
 * kernel 0 = elasticity K(xi)u apply, 1 = elasticity block-Jacobi sweep, 2 = filter apply,
 * 3 = two elasticity sweeps from zero in one launch, 4 = prolongation and two sweeps in one
 * launch (3 and 4 with the handle's sweep-pair chunk heights), 6 = two sweeps from zero with
 * the residual restricted to the next level in one launch, 5 = the mixed PCG's direction
 * update with the FP64 operator application (p = z + beta p_old, q = K(xi) p, x += alpha p_old;
 * TMX_FP64 only; runs on the PCG's own vectors).
 * precision TMX_FP64 or TMX_FP32.  Writes the per-launch algorithmic bytes to *bytes. */
int tmx_bench_kernel(tmx_problem* h, int kernel, int precision, int reps, double* avg_us,
                     double* bytes);
/* In-step timing (measurement only; no counterpart in the reference).  With `enable` set,
 * every launch of kernel 5 of tmx_bench_kernel (the PCG direction update with the FP64
 * operator, one per elasticity PCG iteration) inside later solves is bracketed by HIP events
 * on the handle's stream; tmx_kernel_timing_read synchronises them and returns the summed
 * duration (microseconds) and the count of launches timed since the last read (at most 256;
 * later ones are not timed until a read).  Timing stays off under TMX_OPT_GRAPHS. */
int tmx_kernel_timing(tmx_problem* h, int kernel, int enable);
int tmx_kernel_timing_read(tmx_problem* h, int kernel, double* total_us, int* launches);
/* Solver options (both operator families).  TMX_OPT_FUSED: 1 (default) fuses residual with
 * restriction and prolongation with the next smoothing sweep on element levels; TMX_OPT_TAIL:
 * 0 off, 1 on (default), > 1 on with that many nodes as the largest level run by the
 * one-workgroup bottom-of-cycle kernel; TMX_OPT_NU_PRE / TMX_OPT_NU_POST: elasticity Jacobi
 * sweeps on the finest level (>= 1); TMX_OPT_OMEGA_ELAST / TMX_OPT_OMEGA_FILTER: Jacobi
 * damping. */
enum tmx_option {
  TMX_OPT_FUSED = 0, TMX_OPT_TAIL = 1, TMX_OPT_NU_PRE = 2, TMX_OPT_NU_POST = 3,
  TMX_OPT_OMEGA_ELAST = 4, TMX_OPT_OMEGA_FILTER = 5,
  TMX_OPT_NU_COARSE = 6,    /* pre- and post-sweeps on the coarser levels (PRE/POST: finest) */
  TMX_OPT_FILTER_RTOL = 7,  /* relative residual of the Helmholtz filter solves (default:
                               1e-1 x the elasticity rtol, clamped to [1e-12, 1e-9]) */
  TMX_OPT_NU_FINE_FILTER = 8,   /* filter: pre- and post-sweeps on the finest level */
  TMX_OPT_NU_COARSE_FILTER = 9, /* filter: pre- and post-sweeps on the coarser levels */
  TMX_OPT_CHEB_ELAST = 10,      /* > 0: Chebyshev step sizes for an upper bound lambda of
                                   lambda_max(D^-1 A) (elasticity); 0: damped Jacobi */
  TMX_OPT_CHEB_FILTER = 11,     /* the same for the filter */
  TMX_OPT_CHEB_RATIO = 12,      /* smoothing band [lambda / ratio, lambda] (both; 8) */
  TMX_OPT_BAND = 13,            /* first level (>= 1) of the tiled multi-sweep V-cycle kernels
                                   (both families and precisions; default 1, FP64 elasticity
                                   2), 0: one launch per sweep / transfer instead */
  TMX_OPT_BAND_FINE_FILTER = 14, /* 1 (default): the filter's finest level in the tiled
                                   kernels too */
  TMX_OPT_FINE_CELL_WEIGHTS = 15, /* elasticity finest-level V-cycle sweeps read per-cell
                                   weights formed once per operator instead of the nodal xi:
                                   bit 0 FP32 V-cycles (default on), bit 1 FP64 */
  TMX_OPT_GRAPHS = 16,          /* 1: PCG iterations replay captured hipGraphs; 0 (default):
                                   their kernels are launched one by one (measured faster) */
  TMX_OPT_SWEEP_PAIRS = 17,     /* > 0: the elasticity finest level runs its first two
                                   pre-sweeps as one launch, in node-row chunks of this height
                                   (default 8); 0: one launch per sweep */
  TMX_OPT_SWEEP_PAIRS_POST = 18, /* the same for prolongation with the first two post-sweeps
                                   (default 12 for FP32 V-cycles, 16 for FP64) */
  TMX_OPT_SWEEP_PAIR_RESTRICT = 19, /* > 0: with two pre-sweeps, the zero-start pair also forms
                                   the residual and restricts it to the next level, in chunks
                                   of this many rows (default 12); 0: a separate launch.
                                   Setting any of 17-19 applies to both precisions;
                                   tmx_autotune re-picks the heights of those left > 0 */
  TMX_OPT_FILTER_PIPE = 20,     /* > 0: the filter's finest level (even nx, ny; at most 4
                                   sweeps per pass) runs as two register column pipelines in
                                   node-row chunks of this height (default 12; tmx_autotune
                                   re-picks it per direction and precision); 0: the band
                                   kernels */
  TMX_OPT_FILTER_DCT = 21,      /* 1 (default): the filter's preconditioner (FP32 and FP64) is
                                   the exact DCT-I / tridiagonal solver of
                                   eps^2 K + (lumped mass) when ny is even and nx + 1 <= 3072
                                   (3-5 PCG iterations); 0: the multigrid V-cycle */
  TMX_OPT_PAIR_LEVELS = 22,     /* elasticity element levels 0 .. value-1 run the paired
                                   column-strip kernels, the levels below them the band kernels
                                   (default 1: the finest only); 1..8 */
  TMX_OPT_PAIR_ROWS_COARSE = 23, /* node-row chunk height of the paired kernels on the coarser
                                   paired levels (0: one launch per sweep there) */
  TMX_OPT_PAIR_NU_COARSE = 24,  /* sweeps each way on the coarser paired levels (0: the
                                   coarse-level count, TMX_OPT_NU_COARSE) */
  TMX_OPT_BAND_WIDE = 25,       /* the elasticity's cell-weight band levels: 0: 256-thread
                                   workgroups; 1: 1024 threads with one node per thread (square
                                   tiles, e.g. 24 x 24), each node's row built once per launch;
                                   2 (default): 512 threads on the standard tiles; 3: 1024 */
  TMX_OPT_BENCH_COLD = 26,      /* 1: tmx_bench_kernel times each launch alone after reading
                                   512 MiB (no inputs left in the caches); 0: back to back */
  TMX_OPT_RU_EVERY = 27,        /* TMX_MIXED_RU: iterations between residual replacements
                                   (default 4; none once the residual nears the tolerance) */
  TMX_OPT_SETUP_OVERLAP = 28    /* 1 (default): an elasticity solve builds the coarse operators
                                   (Galerkin levels, coarsest inverse) on an internal stream
                                   while its first V-cycle smooths the finer levels; 0: in
                                   order on the solve's stream */
};
int tmx_set_option(tmx_problem* h, int option, double value);
/* One multigrid V-cycle z = M^-1 r (the PCG preconditioner).  which: 0 = elasticity with the
 * operator of xi (device, n_v), 1 = Helmholtz filter (xi ignored).  precision TMX_FP64 or
 * TMX_FP32 (r, z stay double). */
int tmx_precondition(tmx_problem* h, int which, int precision, const double* xi,
                     const double* r, double* z);
/* Sweep row-chunk override (fp64, fp32); <= 0 keeps the current value. */
int tmx_set_tiles(tmx_problem* h, int rows_fp64, int rows_fp32);

/* Batched objective + gradient over many independent problems (uneven sizes, penalties and
 * loads; each problem keeps its own penalty, tiles and operators).  No counterpart in the
 * reference (it solves one design).
 *
 * TMX_BATCH_MERGED (the default): the designs that qualify run as one merged step: one
 * kernel launch per PCG phase and V-cycle stage covers every such design, the designs'
 * iterations in lockstep, so the launches per step do not grow with the number of designs.
 * A design qualifies once an objective step with the same precision has completed on its
 * handle (P1 element, TMX_FP64 or TMX_MIXED / TMX_FP32, no graph replay, same device as the
 * first problem).  The others, and designs whose merged solve stops without converging, run
 * as single-design calls on worker threads.  Results are bitwise equal to single-design calls.
 * TMX_BATCH_ALONE: every design runs as a single-design call, each on its own stream, driven
 * by `workers` host threads (default min(16, cores), at least 4), largest first.
 *
 * The handles must not be used by other threads during the call.  rho[i], w[i] are device
 * pointers (n_v of design i); J_host[i] receives the compliance (NaN for a non-finite design).
 * stats may be NULL or point to `count` entries. */
typedef struct tmx_batch tmx_batch;
enum { TMX_BATCH_MERGED = 0, TMX_BATCH_ALONE = 1 };
typedef struct tmx_batch_report {
  int merged_designs;    /* designs whose results came from the merged step */
  int solo_designs;      /* designs run as single-design calls */
  int launches;          /* kernel launches of the merged step */
  int iterations[3];     /* merged PCG iterations: filter, elasticity, adjoint filter */
  double record_ms;      /* host time spent recording and merging (0: the plan was reused) */
} tmx_batch_report;
int tmx_batch_create(tmx_problem* const* problems, int count, tmx_batch** out);
void tmx_batch_destroy(tmx_batch* b);
int tmx_batch_set_workers(tmx_batch* b, int workers);
int tmx_batch_set_mode(tmx_batch* b, int mode);
int tmx_batch_objective_and_gradient(tmx_batch* b, const double* const* rho, double* J_host,
                                     double* const* w, const tmx_solve_opts* opts,
                                     tmx_solve_stats* stats);
/* what the last tmx_batch_objective_and_gradient call did */
int tmx_batch_last_report(const tmx_batch* b, tmx_batch_report* out);

/* Entropic mirror descent, design-space half (src/solver.py:149-302), on n_v nodal device
 * vectors; integrals are of the CG1 interpolant (lumped mass, as FEMSolver.integrate).
 *   tmx_emd_half_step: half = psi - step * w                          (Solver.step, :188-193)
 *   tmx_emd_volume:    vol = int expit(half + c), dvol = int expit'(half + c)
 *                      (the function and derivative of Solver.project's Newton solve, :158-162)
 *   tmx_emd_apply:     psi = half + c, rho = expit(psi)               (:176, :262)
 *   tmx_emd_change:    out = sqrt(int (rho - expit(psi_prev))^2)      (:286-288)
 *   tmx_logit:         psi = log(rho / (1 - rho))                     (:227) */
int tmx_emd_half_step(tmx_problem* h, const double* psi, const double* w, double step,
                      double* half);
int tmx_emd_volume(tmx_problem* h, const double* half, double c, double* vol_host,
                   double* dvol_host);
int tmx_emd_apply(tmx_problem* h, const double* half, double c, double* psi, double* rho);
int tmx_emd_change(tmx_problem* h, const double* rho, const double* psi_prev, double* out_host);
int tmx_logit(tmx_problem* h, const double* rho, double* psi);

/* The design-consistency re-evaluation (misc/get_consistent_objectives.py:110-156) moves a
 * design found on a coarser mesh to this problem's mesh: df.interpolate of a CG1 function on
 * the coarse 'right' mesh (coarse_nx x coarse_ny cells, same rectangle) onto this mesh's
 * CG1 space.  rho_coarse: (coarse_nx+1)(coarse_ny+1) device values; rho: n_v. */
int tmx_interpolate_design(tmx_problem* h, int coarse_nx, int coarse_ny,
                           const double* rho_coarse, double* rho);

#ifdef __cplusplus
}
#endif

#endif /* TMX_H */
