"""Entropic mirror descent (EMD) over the GPU elasticity path.

Restates topomax's optimiser loop ``Solver`` (``src/solver.py:42-302``) with every design-space
vector (rho, psi, the half step, the gradient) resident on the GPU:

* ``logit`` / ``expit`` / ``expit_diff``: ``src/solver.py:26-39``;
* the volume projection ``project`` (``:149-186``): Newton's method on
  ``int expit(half + c) dx = V`` from c = 0 with tol 1e-12, then Brent's method on a doubling
  bracket (``smart_brentq``, ``src/utils.py:138-157``) if Newton fails.  Each evaluation of the
  integral and its derivative is one device reduction (``tmx_emd_volume``); the scalar root
  finding uses SciPy's ``newton`` and ``brentq``, the same routines the reference calls;
* ``step`` (``:188-194``), ``tolerance`` (``:196-199``), ``step_size_at_iter`` (``:201-206``);
* ``solve`` (``:208-302``): per penalty, the objective history and the exit rules (NaN
  objective, objective increasing by 2x, 50 iterations without improvement, design change
  below the tolerance, 1000 iterations).

The mesh rule is the reference's: N elements per unit of the shorter side
(``src/solver.py:65-67``).  Outputs (``save_iteration`` / ``save_result``) follow
``src/solver.py:304-349`` and ``FEM_src/utils.py:47-69`` when an output folder is given.
"""

from __future__ import annotations

import copyreg
import ctypes
import io
import math
import os
import pickle
import sys
import time
import types
from dataclasses import dataclass

import numpy as np
import torch
from scipy import optimize

from . import _lib
from .designs import parse_design, parse_design_dict
from .elasticity_problem import ControlSpace, ElasticityProblem, StructuredMesh, _ptr


def smart_brentq(f, initial_radius: float, max_radius: float):
    """Brent's method on [-r, r], doubling r from ``initial_radius`` until f changes sign
    (``src/utils.py:138-157``); ValueError once r exceeds ``max_radius``."""
    r = initial_radius
    while True:
        if r > max_radius:
            raise ValueError("f(-max_radius) and f(max_radius) must have different signs!")
        try:
            return optimize.brentq(f, -r, r, full_output=True)
        except ValueError:
            r *= 2


@dataclass
class PenaltyResult:
    penalty: float
    objectives: list
    times: list
    iteration: int
    exit_condition: str


class EMDSolver:
    """``Solver`` + ``FEMSolver`` for elasticity designs on the GPU.

    ``design`` is a design file path or its parsed JSON dict; ``N`` is the reference's mesh
    parameter.  ``data_path`` enables the reference's per-iteration and per-penalty output
    files under ``{data_path}/GPU/{design}/data``; ``dof_order`` (the vertex of each CG1 dof)
    makes the saved design vectors dof-ordered like the reference's (``save_rho``)."""

    def __init__(self, N: int, design, *, data_path: str | None = None, skip_multiple: int = 1,
                 precision: str = "fp64", rtol: float = 1e-12, filter_rtol: float | None = None,
                 max_iterations: int = 1000, degree: int = 1, device=None,
                 verbose: bool = False, dof_order=None):
        if isinstance(design, dict):
            self.parameters, problem_parameters = parse_design_dict(design)
            self.design_str = "design"
        else:
            self.parameters, problem_parameters = parse_design(design)
            self.design_str = os.path.splitext(os.path.basename(design))[0]
        self.width, self.height = self.parameters.width, self.parameters.height
        self.N, self.full_N = mesh_parameter(N, self.width, self.height)
        self.volume = self.width * self.height * self.parameters.volume_fraction
        self.step_size = self.parameters.fem_step_size
        self.max_iterations = max_iterations
        self.skip_multiple = skip_multiple
        self.verbose = verbose
        # vertex of each dof of the saved design vectors (None: vertex order), see save_rho
        self.dof_order = dof_order
        self.output_folder = (None if data_path is None
                              else f"{data_path}/GPU/{self.design_str}/data")

        self.mesh = StructuredMesh(self.width, self.height, int(self.width * self.N),
                                   int(self.height * self.N))
        self.control_space = ControlSpace(self.mesh, device)
        self.problem = ElasticityProblem(self.mesh, self.control_space, self.parameters,
                                         problem_parameters, precision=precision, rtol=rtol,
                                         degree=degree)
        if filter_rtol is not None:
            self.problem.set_option("filter_rtol", filter_rtol)
        self.rho = self.control_space.function(self.parameters.volume_fraction)
        self.penalty_formatter = _penalty_formatter(self.parameters.penalties)
        h = self.problem._h
        self._h = h
        n = h.n_vertices
        dev = self.control_space.device
        self._half = torch.empty(n, dtype=torch.float64, device=dev)
        self._vol, self._dvol = ctypes.c_double(), ctypes.c_double()

    # ---------------------------------------------------------------- device helpers
    def _call(self, name, *args):
        h = self._h
        h.stream()
        _lib.check(getattr(h.lib, name)(h.ptr, *args))

    def integrate(self, values) -> float:
        return self.problem.integrate(values)

    def _volume(self, c: float):
        self._call("tmx_emd_volume", _ptr(self._half), float(c), ctypes.byref(self._vol),
                   ctypes.byref(self._dvol))
        return self._vol.value, self._dvol.value

    def project(self, half_step: torch.Tensor, volume: float) -> float:
        """The constant c with int expit(half_step + c) = volume (``src/solver.py:149-186``)."""
        if half_step.data_ptr() != self._half.data_ptr():
            self._half.copy_(half_step)
        cache = {}

        def values(c):
            c = float(c)
            if c not in cache:
                cache[c] = self._volume(c)
            return cache[c]

        def error(c):
            return values(c)[0] - volume

        def error_derivative(c):
            return values(c)[1]

        try:
            c, result = optimize.newton(error, 0, error_derivative, tol=1e-12, full_output=True)
            if result.converged:
                return float(c)
        except RuntimeError:
            pass
        c, result = smart_brentq(error, 2, 2000)
        if not result.converged:
            raise ValueError("Projection failed to converge")
        return float(c)

    def step(self, previous_psi: torch.Tensor, step_size: float, psi_out: torch.Tensor):
        """One mirror-descent step (``src/solver.py:188-194``): psi_out and self.rho are
        updated in place."""
        gradient = self.problem.calculate_objective_gradient()
        self._call("tmx_emd_half_step", _ptr(previous_psi), _ptr(gradient), float(step_size),
                   _ptr(self._half))
        c = self.project(self._half, self.volume)
        self._call("tmx_emd_apply", _ptr(self._half), c, _ptr(psi_out), _ptr(self.rho))

    @staticmethod
    def tolerance(k: int) -> float:
        itol, ntol = 1e-2, 1e-5
        return min(25 * (k + 1) * ntol, itol)

    def step_size_at_iter(self, k: int) -> float:
        if len(self.parameters.penalties) > 1:
            return self.step_size * min((k + 1), 10)
        return self.step_size * (k + 1)

    # ---------------------------------------------------------------- the loop
    def solve(self) -> list:
        """Run EMD over every penalty of the design (``src/solver.py:208-302``)."""
        max_iterations = self.max_iterations
        objective_increasing_factor = 2
        max_iterations_without_improvement = 50
        n = self._h.n_vertices
        dev = self.control_space.device
        psi = torch.empty(n, dtype=torch.float64, device=dev)
        previous_psi = torch.empty_like(psi)
        self._call("tmx_logit", _ptr(self.rho), _ptr(psi))
        change = ctypes.c_double()
        results = []
        for penalty in self.parameters.penalties:
            self.problem.set_penalization(penalty)
            t0 = time.perf_counter()
            objectives = [self.problem.calculate_objective(self.rho)]
            times = [time.perf_counter() - t0]
            k = 0
            exit_condition = ""
            for k in range(max_iterations):
                if self.verbose:
                    print(f"p={penalty} k={k} objective={objectives[-1]:.12g}", flush=True)
                if k % self.skip_multiple == 0:
                    self.save_iteration(objectives[-1], k, penalty)
                t0 = time.perf_counter()
                previous_psi.copy_(psi)
                try:
                    self.step(previous_psi, self.step_size_at_iter(k), psi)
                except ValueError as e:
                    exit_condition = str(e)
                    break
                objectives.append(self.problem.calculate_objective(self.rho))
                times.append(time.perf_counter() - t0)
                if np.isnan(objectives[-1]):
                    exit_condition = "Objective is NaN"
                    break
                min_index = int(np.argmin(objectives))
                min_objective = objectives[min_index]
                if objectives[-1] > objective_increasing_factor * min_objective:
                    exit_condition = "Objective is increasing"
                    break
                if k >= min_index + max_iterations_without_improvement:
                    exit_condition = "Objective is not decreasing"
                    break
                self._call("tmx_emd_change", _ptr(self.rho), _ptr(previous_psi),
                           ctypes.byref(change))
                if change.value < self.tolerance(k):
                    exit_condition = "Convergence treshold reached"
                    break
            else:
                exit_condition = "Iteration did not converge"
            self.save_iteration(objectives[-1], k + 1, penalty)
            self.save_result(objectives, times, penalty, exit_condition)
            results.append(PenaltyResult(penalty, objectives, times, k + 1, exit_condition))
        return results

    # ---------------------------------------------------------------- outputs
    def _file_root(self, penalty, suffix):
        return f"{self.output_folder}/N={self.full_N}_p={self.penalty_formatter(penalty)}_{suffix}"

    def save_rho(self, file_root: str, dof_order=None) -> str:
        """``FEM_src/utils.py:47-69`` layout: a pickled dict with N (elements per unit length,
        as ``mesh_to_N`` gives it), domain_size, problem and the nodal vector.

        Without ``dof_order`` the vector is in vertex order (v = iy*(nx+1)+ix), marked by
        ``"order": "vertex"``.  With it (dof_order[k] = the vertex of dof k, e.g. dolfin's
        ``dof_to_vertex_map`` of the CG1 space) the vector is written in that dof order with
        no marker, as the reference's ``save_function`` writes it, so ``load_function`` and
        ``plot.py`` read it unchanged.  ``self.dof_order`` is the default."""
        rho_file = file_root + "_rho.dat"
        data = design_record(self.rho.detach().cpu().numpy(), self.N, self.width, self.height,
                             self.dof_order if dof_order is None else dof_order)
        with open(rho_file, "wb") as fh:
            pickle.dump(data, fh)
        return os.path.basename(rho_file)

    def save_iteration(self, objective: float, k: int, penalty: float):
        if self.output_folder is None:
            return
        file_root = self._file_root(penalty, f"k={k}")
        os.makedirs(os.path.dirname(file_root), exist_ok=True)
        rho_basename = self.save_rho(file_root)
        state = {"domain_size": (self.width, self.height), "objective": objective,
                 "iteration": k, "rho_file": rho_basename, "penalty": penalty}
        _dump_reference_record(f"{file_root}.dat", "IterationData", state)

    def save_result(self, objectives, times, penalty, exit_condition):
        if self.output_folder is None:
            return
        file_root = self._file_root(penalty, "result")
        os.makedirs(os.path.dirname(file_root), exist_ok=True)
        min_index = int(np.argmin(objectives))
        state = {"exit_condition": exit_condition, "min_objective": objectives[min_index],
                 "objectives": list(objectives), "iterations": len(objectives),
                 "min_index": min_index, "times": list(times)}
        _dump_reference_record(f"{file_root}.dat", "SolverResult", state)


def mesh_parameter(N: int, width: float, height: float):
    """(elements per unit length, the N of the file names) for the solver's mesh parameter N:
    N elements along the shorter side (``src/solver.py:65-67``)."""
    n_unit = int(N / min(width, height))
    return n_unit, int(n_unit * min(width, height))


def design_record(vec, n_unit: int, width: float, height: float, dof_order=None) -> dict:
    """The dict ``save_function`` pickles for a CG1 design (``FEM_src/utils.py:47-69``): N is
    the elements per unit length (``mesh_to_N``).  ``vec`` is in vertex order; with
    ``dof_order`` (dof_order[k] = the vertex of dof k) the vector is written in that order,
    otherwise it stays in vertex order and the dict is marked ``"order": "vertex"``."""
    vec = np.asarray(vec, dtype=np.float64)
    data = {"N": int(n_unit), "domain_size": (float(width), float(height)), "problem": "design"}
    if dof_order is None:
        data.update(vector=vec, order="vertex")
        return data
    perm = np.asarray(dof_order, dtype=np.int64)
    if perm.shape != vec.shape or not np.array_equal(np.sort(perm), np.arange(vec.size)):
        raise ValueError("dof_order is not a permutation of the mesh vertices")
    data["vector"] = vec[perm]
    return data


def _penalty_formatter(penalties):
    """Zero-padded penalty strings of equal width (``src/solver.py:128-147``)."""
    padding = []
    for p in penalties:
        left, _, right = str(p).partition(".")
        padding.append((len(left), len(right) if right else 0))
    left_pad = max(v[0] for v in padding)
    right_pad = max(v[1] for v in padding)

    def fmt(p):
        left = len(str(p).split(".", maxsplit=1)[0])
        return f"{'0' * (left_pad - left)}{p:.{right_pad}f}"

    return fmt


def _dump_reference_record(path: str, class_name: str, state: dict):
    """Pickle ``state`` as an instance of the reference's dataclass ``src.utils.<class_name>``
    (the objects ``src/solver.py:304-349`` writes), without importing the reference: the
    class reference is written by name, and the reference's plotting code resolves it when it
    loads the file."""
    placeholder_module = types.ModuleType("src.utils")
    cls = type(class_name, (), {"__module__": "src.utils", "__qualname__": class_name})
    setattr(placeholder_module, class_name, cls)
    obj = cls.__new__(cls)
    obj.__dict__.update(state)

    class _Pickler(pickle.Pickler):
        def reducer_override(self, o):
            if o is obj:
                return (copyreg.__newobj__, (cls,), dict(state))
            return NotImplemented

    saved = {k: sys.modules.get(k) for k in ("src", "src.utils")}
    real = saved["src.utils"]
    buf = io.BytesIO()
    try:
        if real is None or not hasattr(real, class_name):
            pkg = saved["src"] or types.ModuleType("src")
            sys.modules["src"] = pkg
            sys.modules["src.utils"] = placeholder_module
            _Pickler(buf, protocol=pickle.DEFAULT_PROTOCOL).dump(obj)
        else:  # running inside topomax: use its class directly
            real_cls = getattr(real, class_name)
            pickle.dump(real_cls(**state), buf)
    finally:
        for k, v in saved.items():
            if v is None:
                sys.modules.pop(k, None)
            else:
                sys.modules[k] = v
    with open(path, "wb") as fh:
        fh.write(buf.getvalue())
