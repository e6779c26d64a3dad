"""Drop-in counterpart of topomax's FEM elasticity hot path on one MI355X.

Same call surface as the reference:

* ``ElasticityProblem(mesh, control_space, domain_parameters, elasticity_parameters)``
  (``FEM_src/elasisity_problem.py:76-196``) with ``set_penalization``,
  ``calculate_objective``, ``calculate_objective_gradient`` and ``forward``
  (the ``Problem`` ABC, ``src/problem.py:7-23``);
* ``HelmholtzFilter(epsilon, function_space).apply(f)`` (``FEM_src/filter.py:8-41``);
* ``ElasticPenalizer`` (``src/penalizers.py:29-45``).

Functions are torch float64 tensors on the GPU in vertex order v = iy*(nx+1) + ix (the
reference's arrays are in dolfin dof order, a permutation of this); displacements are
node-interleaved (2 n_v).  All numerics run in libtmx.so through ctypes; there is no CPU
fallback.
"""

from __future__ import annotations

import ctypes
import math

import torch

from . import _lib
from .designs import DomainParameters, ElasticityParameters


# solve schemes: FP64 multigrid-PCG; FP64 PCG with an FP32 V-cycle preconditioner; FP32
# multigrid-PCG inner solves inside FP64 iterative refinement
PRECISIONS = {"fp64": _lib.TMX_FP64, "mixed": _lib.TMX_MIXED, "mixed-ir": _lib.TMX_MIXED_IR,
              "mixed-ru": _lib.TMX_MIXED_RU}


def _device(device=None) -> torch.device:
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    device = torch.device(device)
    if device.type != "cuda":
        raise ValueError("topomax_amd runs on a GPU device only")
    return device


def _ptr(t: torch.Tensor):
    return ctypes.c_void_p(t.data_ptr())


class StructuredMesh:
    """``RectangleMesh((0,0), (W,H), nx, ny)`` with the default 'right' diagonal
    (``FEM_src/solver.py:38-47``); vertex x = 0 + (W/nx) ix, y = 0 + (H/ny) iy."""

    def __init__(self, width: float, height: float, nx: int, ny: int):
        if nx < 1 or ny < 1:
            raise ValueError("nx and ny must be >= 1")
        self.width, self.height = float(width), float(height)
        self.nx, self.ny = int(nx), int(ny)

    @classmethod
    def from_N(cls, width: float, height: float, full_N: int):
        """N per unit length as ``Solver.__init__`` computes it (``src/solver.py:65-67``)."""
        n = int(full_N / min(width, height))
        return cls(width, height, int(width * n), int(height * n))

    @property
    def n_vertices(self) -> int:
        return (self.nx + 1) * (self.ny + 1)

    @property
    def n_cells(self) -> int:
        return 2 * self.nx * self.ny

    def coordinates(self, device="cpu") -> torch.Tensor:
        ab = (self.width - 0.0) / self.nx
        cd = (self.height - 0.0) / self.ny
        xs = 0.0 + ab * torch.arange(self.nx + 1, dtype=torch.float64, device=device)
        ys = 0.0 + cd * torch.arange(self.ny + 1, dtype=torch.float64, device=device)
        yy, xx = torch.meshgrid(ys, xs, indexing="ij")
        return torch.stack([xx.reshape(-1), yy.reshape(-1)], dim=1)


class ControlSpace:
    """CG1 scalar space on a StructuredMesh (``FunctionSpace(mesh, "CG", 1)``)."""

    def __init__(self, mesh: StructuredMesh, device=None):
        self.mesh = mesh
        self.device = _device(device)

    def dim(self) -> int:
        return self.mesh.n_vertices

    def function(self, value: float = 0.0) -> torch.Tensor:
        return torch.full((self.dim(),), float(value), dtype=torch.float64, device=self.device)


class ElasticPenalizer:
    """SIMP r(x) = m + x^p (1-m), r'(x) = p x^(p-1) (1-m) (``src/penalizers.py:29-45``)."""

    def __init__(self):
        self.penalization = None
        self.minimum = 1e-6

    def set_penalization(self, penalization: float):
        self.penalization = penalization

    def assert_has_penalization(self):
        if self.penalization is None:
            raise ValueError("You must set penalization before calling penalizer")
        return self.penalization

    def __call__(self, rho):
        p, m = self.assert_has_penalization(), self.minimum
        return m + rho ** p * (1 - m)

    def derivative(self, rho):
        p, m = self.assert_has_penalization(), self.minimum
        return p * rho ** (p - 1) * (1 - m)


class _Handle:
    """Owns one tmx_problem."""

    def __init__(self, desc: _lib.ProblemDesc, device: torch.device):
        self.lib = _lib.load()
        self.device = device
        self.ptr = ctypes.c_void_p()
        with torch.cuda.device(device):
            _lib.check(self.lib.tmx_problem_create(ctypes.byref(desc), ctypes.byref(self.ptr)))
        nv, nd = ctypes.c_int64(), ctypes.c_int64()
        fixed, levels = ctypes.c_uint(), ctypes.c_int()
        _lib.check(self.lib.tmx_problem_info(self.ptr, ctypes.byref(nv), ctypes.byref(nd),
                                             ctypes.byref(fixed), ctypes.byref(levels)))
        self.n_vertices, self.n_dofs = nv.value, nd.value
        self.fixed_sides_applied = fixed.value
        self.levels = levels.value

    def stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.lib.tmx_set_stream(self.ptr, ctypes.c_void_p(s)))

    def __del__(self):
        lib = getattr(self, "lib", None)
        ptr = getattr(self, "ptr", None)
        if lib is not None and ptr is not None and ptr.value:
            lib.tmx_problem_destroy(ptr)
            self.ptr = ctypes.c_void_p()


def _make_desc(mesh: StructuredMesh, params: ElasticityParameters | None, filter_radius: float,
               precision: str, degree: int = 1) -> _lib.ProblemDesc:
    d = _lib.ProblemDesc()
    d.nx, d.ny = mesh.nx, mesh.ny
    d.width, d.height = mesh.width, mesh.height
    d.young_modulus = params.young_modulus if params else 1.0
    d.poisson_ratio = params.poisson_ratio if params else 0.3
    d.penalty_minimum = 1e-6
    d.filter_radius = float(filter_radius)
    fixed = 0
    if params:
        for side in params.fixed_sides:
            fixed |= _lib.FIX_BITS[side.label]
    d.fixed_sides = fixed
    tractions = (params.tractions or []) if params else []
    if len(tractions) > _lib.TMX_MAX_TRACTIONS:
        raise ValueError(f"at most {_lib.TMX_MAX_TRACTIONS} tractions are supported")
    d.n_tractions = len(tractions)
    for i, t in enumerate(tractions):
        d.tractions[i].side = _lib.SIDE_CODES[t.side.label]
        d.tractions[i].center = t.center
        d.tractions[i].length = t.length
        d.tractions[i].value[0], d.tractions[i].value[1] = t.value
    if params and params.body_force is not None:
        bf = params.body_force
        d.has_body_force = 1
        d.body_force.center[0], d.body_force.center[1] = bf.region.center
        d.body_force.radius = bf.region.radius
        d.body_force.value[0], d.body_force.value[1] = bf.value
    d.precision = PRECISIONS[precision]
    if degree not in (1, 2):
        raise ValueError("degree must be 1 (P1) or 2 (P2)")
    d.degree = degree
    return d


def _as_field(t, n: int, device: torch.device, name: str) -> torch.Tensor:
    t = torch.as_tensor(t, dtype=torch.float64, device=device)
    if t.numel() != n:
        raise ValueError(f"{name} has {t.numel()} values, expected {n}")
    return t.reshape(n).contiguous()


def solve_opts(rtol=None, max_iter=None, precision=None, inner_rtol=None,
               max_outer=None) -> _lib.SolveOpts:
    o = _lib.SolveOpts()
    o.rtol = rtol or 0.0
    o.max_iter = max_iter or 0
    o.precision = -1 if precision is None else PRECISIONS[precision]
    o.inner_rtol = inner_rtol or 0.0
    o.max_outer = max_outer or 0
    return o


class HelmholtzFilter:
    """-eps^2 Lap xi + xi = rho with natural BCs on CG1 (``FEM_src/filter.py:8-41``).

    ``apply(f)`` takes a CG1 function (nodal tensor) and solves A xi = M f;
    ``solve(b)`` takes an assembled right-hand side b_i = int g phi_i (the reference's
    form input) and solves A w = b.  The operator is built once, as in the reference.
    """

    def __init__(self, epsilon: float, function_space: ControlSpace, *, _handle=None):
        self.epsilon = float(epsilon)
        self.function_space = function_space
        if _handle is None:
            desc = _make_desc(function_space.mesh, None, self.epsilon, "fp64")
            _handle = _Handle(desc, function_space.device)
        self._h = _handle

    def apply(self, input_function) -> torch.Tensor:
        h = self._h
        rho = _as_field(input_function, h.n_vertices, h.device, "input_function")
        xi = torch.empty_like(rho)
        h.stream()
        rc = h.lib.tmx_filter_apply(h.ptr, _ptr(rho), _ptr(xi))
        if rc != _lib.TMX_ERR_NONFINITE:  # non-finite input: NaN output, as a direct solve
            _lib.check(rc)
        return xi

    def solve(self, rhs) -> torch.Tensor:
        h = self._h
        b = _as_field(rhs, h.n_vertices, h.device, "rhs")
        w = torch.empty_like(b)
        h.stream()
        rc = h.lib.tmx_filter_solve(h.ptr, _ptr(b), _ptr(w))
        if rc != _lib.TMX_ERR_NONFINITE:
            _lib.check(rc)
        return w

    def mass(self, x) -> torch.Tensor:
        """M x with the CG1 mass matrix."""
        h = self._h
        x = _as_field(x, h.n_vertices, h.device, "x")
        y = torch.empty_like(x)
        h.stream()
        _lib.check(h.lib.tmx_mass_apply(h.ptr, _ptr(x), _ptr(y)))
        return y


class ElasticityProblem:
    """Elastic compliance topology optimization problem (``FEM_src/elasisity_problem.py:76``).

    Extra keyword arguments (not in the reference): ``precision`` ("fp64"; "mixed": FP64 PCG
    with an FP32 multigrid preconditioner; "mixed-ir": FP32 multigrid-PCG inside FP64
    iterative refinement), ``rtol`` (relative residual of the state solve; the reference
    solves directly), ``degree`` (displacement element: 1 = P1, the default, or 2 = P2, the
    reference's element, solved in FP64), ``device``.
    """

    def __init__(self, mesh: StructuredMesh, control_space: ControlSpace,
                 domain_parameters: DomainParameters, elasticity_parameters: ElasticityParameters,
                 *, precision: str = "fp64", rtol: float = 1e-10, degree: int = 1, device=None):
        self.parameters = elasticity_parameters
        self.domain_size = (domain_parameters.width, domain_parameters.height)
        if (mesh.width, mesh.height) != (float(domain_parameters.width),
                                         float(domain_parameters.height)):
            raise ValueError("mesh and domain parameters disagree on the domain size")
        self.mesh = mesh
        self.control_space = control_space
        self.device = _device(device if device is not None else control_space.device)
        self.precision = precision
        self.rtol = rtol
        self.degree = degree
        desc = _make_desc(mesh, elasticity_parameters, elasticity_parameters.filter_radius,
                          precision, degree)
        self._h = _Handle(desc, self.device)
        self.filter = HelmholtzFilter(elasticity_parameters.filter_radius, control_space,
                                      _handle=self._h)
        self.Young_modulus = elasticity_parameters.young_modulus
        self.Poisson_ratio = elasticity_parameters.poisson_ratio
        self.penalizer = ElasticPenalizer()
        self.lamé_mu = self.Young_modulus / (2 * (1 + self.Poisson_ratio))
        self.lamé_lda = self.lamé_mu * self.Poisson_ratio / (0.5 - self.Poisson_ratio)
        self.u = None
        self.filtered_rho = None
        self.last_stats = None

    # --------------------------------------------------------------- Problem ABC
    def set_penalization(self, penalization: float):
        """``FEM_src/problem.py:29-35``."""
        if self.penalizer is None:
            raise ValueError("Classes deriving from Problem must "
                             "set a penalizer in their initializer")
        _lib.check(self._h.lib.tmx_set_penalization(self._h.ptr, float(penalization)))
        self.penalizer.set_penalization(penalization)

    def forward(self, rho, *, rtol: float | None = None, precision: str | None = None):
        """Solve K(rho) u = F (``forward`` :168-169); ``rho`` is the already-filtered
        density, as in the reference.  Returns u (2 n_v, node-interleaved)."""
        self.penalizer.assert_has_penalization()
        h = self._h
        xi = _as_field(rho, h.n_vertices, h.device, "rho")
        u = torch.empty(h.n_dofs, dtype=torch.float64, device=h.device)
        stats = _lib.SolveStats()
        opts = solve_opts(rtol=rtol or self.rtol, precision=precision)
        h.stream()
        rc = h.lib.tmx_elasticity_solve(h.ptr, _ptr(xi), _ptr(u), ctypes.byref(opts),
                                        ctypes.byref(stats))
        self.last_stats = stats.as_dict()
        if rc == _lib.TMX_ERR_NONFINITE:
            return u  # NaN-filled, like a direct solve of non-finite data
        _lib.check(rc)
        return u

    def calculate_objective(self, rho) -> float:
        """J(rho) = int u.f dx + int_ds u.t (``calculate_objective`` :152-166)."""
        self.penalizer.assert_has_penalization()
        self.filtered_rho = self.filter.apply(rho)
        self.u = self.forward(self.filtered_rho)
        if not bool(torch.isfinite(self.u).all()):
            return math.nan
        return self.compliance(self.u)

    def calculate_objective_gradient(self) -> torch.Tensor:
        """Filtered sensitivity w = A^{-1} b, b_i = int -r'(xi) sigma0(u):eps(u) phi_i
        (``calculate_objective_gradient`` :134-150).  ``M w`` is dJ/drho."""
        if self.filtered_rho is None or self.u is None:
            raise ValueError("You must call calculate_objective "
                             "before calling calculate_objective_gradient")
        b = self.sensitivity(self.filtered_rho, self.u)
        return self.filter.solve(b)

    # --------------------------------------------------------------- building blocks
    def compliance(self, u) -> float:
        h = self._h
        u = _as_field(u, h.n_dofs, h.device, "u")
        out = ctypes.c_double()
        h.stream()
        _lib.check(h.lib.tmx_compliance(h.ptr, _ptr(u), ctypes.byref(out)))
        return float(out.value)

    def sensitivity(self, xi, u) -> torch.Tensor:
        h = self._h
        xi = _as_field(xi, h.n_vertices, h.device, "xi")
        u = _as_field(u, h.n_dofs, h.device, "u")
        b = torch.empty_like(xi)
        h.stream()
        _lib.check(h.lib.tmx_sensitivity(h.ptr, _ptr(xi), _ptr(u), _ptr(b)))
        return b

    def objective_and_gradient(self, rho, *, rtol=None, precision=None, want_xi_u=False):
        """Fused filter -> solve -> compliance -> sensitivity -> filter adjoint.
        Returns (J, w, Mw) and caches xi, u like ``calculate_objective``."""
        self.penalizer.assert_has_penalization()
        h = self._h
        rho = _as_field(rho, h.n_vertices, h.device, "rho")
        xi = torch.empty_like(rho)
        u = torch.empty(h.n_dofs, dtype=torch.float64, device=h.device)
        w = torch.empty_like(rho)
        mw = torch.empty_like(rho)
        J = ctypes.c_double()
        stats = _lib.SolveStats()
        opts = solve_opts(rtol=rtol or self.rtol, precision=precision)
        h.stream()
        rc = h.lib.tmx_objective_and_gradient(h.ptr, _ptr(rho), ctypes.byref(J), _ptr(xi),
                                              _ptr(u), _ptr(w), _ptr(mw), ctypes.byref(opts),
                                              ctypes.byref(stats))
        self.last_stats = stats.as_dict()
        if rc != _lib.TMX_ERR_NONFINITE:
            _lib.check(rc)
        self.filtered_rho, self.u = xi, u
        value = float(J.value) if rc == _lib.TMX_OK else math.nan
        return value, w, mw

    def matvec(self, xi, u, precision: str = "fp64") -> torch.Tensor:
        h = self._h
        xi = _as_field(xi, h.n_vertices, h.device, "xi")
        u = _as_field(u, h.n_dofs, h.device, "u")
        y = torch.empty_like(u)
        prec = {"fp64": _lib.TMX_FP64, "fp32": _lib.TMX_FP32}[precision]
        h.stream()
        _lib.check(h.lib.tmx_matvec(h.ptr, _ptr(xi), _ptr(u), _ptr(y), prec))
        return y

    def load_vector(self) -> torch.Tensor:
        h = self._h
        F = torch.empty(h.n_dofs, dtype=torch.float64, device=h.device)
        h.stream()
        _lib.check(h.lib.tmx_load_vector(h.ptr, _ptr(F)))
        return F

    def interpolate_design(self, rho_coarse, coarse_mesh: StructuredMesh) -> torch.Tensor:
        """This mesh's CG1 interpolant of a design given on a coarser mesh of the same domain
        (``df.interpolate(small_rho, control_space)``, ``misc/get_consistent_objectives.py:
        132-134``)."""
        h = self._h
        if (coarse_mesh.width, coarse_mesh.height) != (self.mesh.width, self.mesh.height):
            raise ValueError("the coarse mesh covers a different domain")
        rc = _as_field(rho_coarse, coarse_mesh.n_vertices, h.device, "rho_coarse")
        out = torch.empty(h.n_vertices, dtype=torch.float64, device=h.device)
        h.stream()
        _lib.check(h.lib.tmx_interpolate_design(h.ptr, coarse_mesh.nx, coarse_mesh.ny, _ptr(rc),
                                                _ptr(out)))
        return out

    def integrate(self, values) -> float:
        """``FEMSolver.integrate`` (``FEM_src/solver.py:81-84``)."""
        h = self._h
        v = _as_field(values, h.n_vertices, h.device, "values")
        out = ctypes.c_double()
        h.stream()
        _lib.check(h.lib.tmx_integrate(h.ptr, _ptr(v), ctypes.byref(out)))
        return float(out.value)

    def autotune(self):
        rows = (ctypes.c_int * 2)()
        us = (ctypes.c_double * 2)()
        self._h.stream()
        _lib.check(self._h.lib.tmx_autotune(self._h.ptr, rows, us))
        return {"rows_fp64": rows[0], "rows_fp32": rows[1], "us_fp64": us[0], "us_fp32": us[1]}

    def set_option(self, name: str, value):
        """Solver knobs: fused, tail, nu_pre, nu_post, nu_coarse (elasticity), nu_fine_filter,
        nu_coarse_filter, omega_elast, omega_filter, filter_rtol, cheb_elast, cheb_filter,
        cheb_ratio, band, band_fine_filter, fine_cell_weights, graphs, sweep_pairs,
        sweep_pairs_post, sweep_pair_restrict, filter_pipe, filter_dct, pair_levels,
        pair_rows_coarse, pair_nu_coarse, band_wide, bench_cold, ru_every, setup_overlap."""
        opt = {"fused": _lib.TMX_OPT_FUSED, "tail": _lib.TMX_OPT_TAIL,
               "nu_pre": _lib.TMX_OPT_NU_PRE, "nu_post": _lib.TMX_OPT_NU_POST,
               "nu_coarse": _lib.TMX_OPT_NU_COARSE,
               "omega_elast": _lib.TMX_OPT_OMEGA_ELAST,
               "omega_filter": _lib.TMX_OPT_OMEGA_FILTER,
               "filter_rtol": _lib.TMX_OPT_FILTER_RTOL,
               "nu_fine_filter": _lib.TMX_OPT_NU_FINE_FILTER,
               "nu_coarse_filter": _lib.TMX_OPT_NU_COARSE_FILTER,
               "cheb_elast": _lib.TMX_OPT_CHEB_ELAST,
               "cheb_filter": _lib.TMX_OPT_CHEB_FILTER,
               "cheb_ratio": _lib.TMX_OPT_CHEB_RATIO,
               "band": _lib.TMX_OPT_BAND,
               "band_fine_filter": _lib.TMX_OPT_BAND_FINE_FILTER,
               "fine_cell_weights": _lib.TMX_OPT_FINE_CELL_WEIGHTS,
               "graphs": _lib.TMX_OPT_GRAPHS,
               "sweep_pairs": _lib.TMX_OPT_SWEEP_PAIRS,
               "sweep_pairs_post": _lib.TMX_OPT_SWEEP_PAIRS_POST,
               "sweep_pair_restrict": _lib.TMX_OPT_SWEEP_PAIR_RESTRICT,
               "filter_pipe": _lib.TMX_OPT_FILTER_PIPE,
               "filter_dct": _lib.TMX_OPT_FILTER_DCT,
               "pair_levels": _lib.TMX_OPT_PAIR_LEVELS,
               "pair_rows_coarse": _lib.TMX_OPT_PAIR_ROWS_COARSE,
               "pair_nu_coarse": _lib.TMX_OPT_PAIR_NU_COARSE,
               "band_wide": _lib.TMX_OPT_BAND_WIDE,
               "bench_cold": _lib.TMX_OPT_BENCH_COLD,
               "ru_every": _lib.TMX_OPT_RU_EVERY,
               "setup_overlap": _lib.TMX_OPT_SETUP_OVERLAP}[name]
        _lib.check(self._h.lib.tmx_set_option(self._h.ptr, opt, float(value)))

    def precondition(self, r, which: str = "elasticity", xi=None, precision: str = "fp64"):
        """One multigrid V-cycle z = M^-1 r of the elasticity (operator of xi) or filter
        hierarchy."""
        h = self._h
        w = {"elasticity": 0, "filter": 1}[which]
        n = h.n_dofs if w == 0 else h.n_vertices
        r = _as_field(r, n, h.device, "r")
        z = torch.empty_like(r)
        xi_p = None
        if xi is not None:
            xi_t = _as_field(xi, h.n_vertices, h.device, "xi")
            xi_p = _ptr(xi_t)
        prec = {"fp64": _lib.TMX_FP64, "fp32": _lib.TMX_FP32}[precision]
        h.stream()
        _lib.check(h.lib.tmx_precondition(h.ptr, w, prec, xi_p, _ptr(r), _ptr(z)))
        return z

    def set_tiles(self, rows_fp64: int = 0, rows_fp32: int = 0):
        _lib.check(self._h.lib.tmx_set_tiles(self._h.ptr, int(rows_fp64), int(rows_fp32)))

    @property
    def fixed_sides_applied(self) -> int:
        return self._h.fixed_sides_applied

    @property
    def levels(self) -> int:
        return self._h.levels


def problem_from_design(design_file_or_dict, N: int, *, precision="fp64", rtol=1e-10,
                        degree: int = 1, device=None):
    """Mesh + control space + ElasticityProblem the way ``FEMSolver`` builds them
    (``FEM_src/solver.py:16-30,38-55``)."""
    from .designs import parse_design, parse_design_dict
    if isinstance(design_file_or_dict, dict):
        domain, params = parse_design_dict(design_file_or_dict)
    else:
        domain, params = parse_design(design_file_or_dict)
    mesh = StructuredMesh.from_N(domain.width, domain.height, N)
    space = ControlSpace(mesh, device)
    return ElasticityProblem(mesh, space, domain, params, precision=precision, rtol=rtol,
                             degree=degree, device=device), domain


class ProblemBatch:
    """Objective + gradient of many independent designs in one call (``tmx_batch_*``).

    The problems may differ in size, loads, supports and penalty; each keeps its own
    hierarchy.  ``mode="merged"`` (the default): the designs that qualify (an objective step in
    the same precision has run on them before, P1 element, FP64 or mixed) run as one merged
    step, one kernel launch per PCG phase and V-cycle stage covering all of them; the others
    run as single-design calls on ``workers`` host threads (default min(16, cores), at least
    4) meanwhile.  ``mode="alone"``: every design runs as a single-design call on the worker
    threads, largest first.  Either way the results are identical, bit for bit, to calling
    each problem's :meth:`ElasticityProblem.objective_and_gradient` alone;
    :attr:`last_report` tells what the last call did."""

    MODES = {"merged": _lib.TMX_BATCH_MERGED, "alone": _lib.TMX_BATCH_ALONE}

    def __init__(self, problems, workers: int | None = None, mode: str = "merged"):
        self.problems = list(problems)
        if not self.problems:
            raise ValueError("a batch needs at least one problem")
        if mode not in self.MODES:
            raise ValueError(f"mode must be one of {sorted(self.MODES)}")
        self.lib = _lib.load()
        ptrs = (ctypes.c_void_p * len(self.problems))(*[p._h.ptr.value for p in self.problems])
        self.ptr = ctypes.c_void_p()
        _lib.check(self.lib.tmx_batch_create(ptrs, len(self.problems), ctypes.byref(self.ptr)))
        if workers is not None:
            _lib.check(self.lib.tmx_batch_set_workers(self.ptr, int(workers)))
        _lib.check(self.lib.tmx_batch_set_mode(self.ptr, self.MODES[mode]))
        self.mode = mode
        self.last_stats = []
        self.last_report = {}

    def objective_and_gradient(self, rhos, *, rtol=None, precision=None):
        """Returns (J list, w list); J is NaN for a design with non-finite entries."""
        n = len(self.problems)
        if len(rhos) != n:
            raise ValueError(f"expected {n} designs, got {len(rhos)}")
        rs, ws = [], []
        for p, r in zip(self.problems, rhos):
            p.penalizer.assert_has_penalization()
            h = p._h
            rs.append(_as_field(r, h.n_vertices, h.device, "rho"))
            ws.append(torch.empty_like(rs[-1]))
            h.stream()
        for p in self.problems:
            torch.cuda.current_stream(p._h.device).synchronize()
        J = (ctypes.c_double * n)()
        stats = (_lib.SolveStats * n)()
        rp = (ctypes.c_void_p * n)(*[r.data_ptr() for r in rs])
        wp = (ctypes.c_void_p * n)(*[w.data_ptr() for w in ws])
        rtols = {p.rtol for p in self.problems}
        if rtol is None and len(rtols) != 1:
            raise ValueError("problems of a batch with different rtol need an explicit rtol")
        opts = solve_opts(rtol=rtol or rtols.pop(), precision=precision)
        _lib.check(self.lib.tmx_batch_objective_and_gradient(self.ptr, rp, J, wp,
                                                             ctypes.byref(opts), stats))
        self.last_stats = [s.as_dict() for s in stats]
        rep = _lib.BatchReport()
        _lib.check(self.lib.tmx_batch_last_report(self.ptr, ctypes.byref(rep)))
        self.last_report = rep.as_dict()
        return [float(j) for j in J], ws

    def __del__(self):
        lib = getattr(self, "lib", None)
        ptr = getattr(self, "ptr", None)
        if lib is not None and ptr is not None and ptr.value:
            lib.tmx_batch_destroy(ptr)
            self.ptr = ctypes.c_void_p()
