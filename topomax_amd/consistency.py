"""Design-consistency re-evaluation on the GPU path.

Restates ``misc/get_consistent_objectives.py:110-145``: a design found on a coarse mesh
(parameter N) is interpolated onto the big_N = 320 mesh of the same design file
(``df.interpolate`` of the CG1 design onto the fine CG1 space) and its objective is evaluated
there with the design's largest penalty.
"""

from __future__ import annotations

import io
import pickle

import numpy as np

from .elasticity_problem import StructuredMesh, problem_from_design

# The only globals a design file may name: numpy's array and scalar reconstructors (module
# paths of numpy 1.x and 2.x).  Anything else is refused before it is imported or called.
_DESIGN_GLOBALS = frozenset({
    ("numpy.core.multiarray", "_reconstruct"), ("numpy._core.multiarray", "_reconstruct"),
    ("numpy.core.multiarray", "scalar"), ("numpy._core.multiarray", "scalar"),
    ("numpy", "ndarray"), ("numpy", "dtype"),
})


class _DesignUnpickler(pickle.Unpickler):
    def find_class(self, module, name):
        if (module, name) in _DESIGN_GLOBALS:
            return super().find_class(module, name)
        raise pickle.UnpicklingError(
            f"design file names {module}.{name}; only numpy arrays and scalars are accepted")


def read_design_dict(path: str) -> dict:
    """The ``FEM_src/utils.py:47-69`` dict of a design file (N, domain_size, problem, vector),
    read with an unpickler that accepts numpy arrays and scalars and nothing else."""
    with open(path, "rb") as fh:
        data = _DesignUnpickler(io.BytesIO(fh.read())).load()
    if not isinstance(data, dict) or not {"N", "domain_size", "vector"} <= set(data):
        raise ValueError(f"{path} is not a design file (a dict with N, domain_size, vector)")
    return data


def load_design(path: str, dof_order=None):
    """Read a design file: ``(N, domain_size, vector)`` with N the elements per unit length
    (``FEM_src/utils.py:29-37``) and the nodal vector in vertex order v = iy*(nx+1)+ix.

    Files written by ``EMDSolver.save_rho`` without a permutation are marked
    ``"order": "vertex"``.  Any other file (the reference's ``save_function`` writes dolfin's
    CG1 dof order) needs ``dof_order``: dof_order[k] is the vertex of dof k (dolfin's
    ``dof_to_vertex_map``), as ``EMDSolver.save_rho`` takes it."""
    data = read_design_dict(path)
    vec = np.asarray(data["vector"], dtype=np.float64)
    if data.get("order") != "vertex":
        if dof_order is None:
            raise ValueError("design vector is in dof order: pass dof_order (the vertex of "
                             "each dof) to map it to vertex order")
        perm = np.asarray(dof_order, dtype=np.int64)
        if perm.shape != vec.shape or not np.array_equal(np.sort(perm), np.arange(vec.size)):
            raise ValueError("dof_order is not a permutation of the design's vertices")
        out = np.empty_like(vec)
        out[perm] = vec
        vec = out
    return int(data["N"]), tuple(float(v) for v in data["domain_size"]), vec


def design_mesh(N: int, domain_size) -> StructuredMesh:
    """The mesh ``load_function`` rebuilds for a design file: int(w N) x int(h N) cells
    (``FEM_src/utils.py:79-90``), N per unit length."""
    w, h = domain_size
    return StructuredMesh(w, h, int(w * N), int(h * N))


def consistent_objective(design, N: int, rho_coarse, *, big_N: int = 320, degree: int = 1,
                         precision: str = "fp64", rtol: float = 1e-10, device=None,
                         coarse_mesh: StructuredMesh | None = None):
    """Objective of the coarse design (nodal values in vertex order) on the big_N mesh with the
    largest penalty of the design.  The coarse mesh is the solver's for mesh parameter N
    (``src/solver.py:65-67``) unless ``coarse_mesh`` is given (``design_mesh`` of a file read
    by ``load_design``).  Returns (objective, fine design)."""
    problem, domain = problem_from_design(design, big_N, precision=precision, rtol=rtol,
                                          degree=degree, device=device)
    problem.set_penalization(max(domain.penalties))
    coarse = coarse_mesh or StructuredMesh.from_N(domain.width, domain.height, N)
    rho = problem.interpolate_design(rho_coarse, coarse)
    return problem.calculate_objective(rho), rho
