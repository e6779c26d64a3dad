"""topomax_amd: topomax's elasticity hot path (Helmholtz filter -> SIMP P1 elasticity
solve -> compliance -> filtered sensitivity) on one MI355X, behind the reference's
ElasticityProblem / HelmholtzFilter call surface."""

from .designs import (DomainParameters, ElasticityParameters, Force, CircularRegion, Side,
                      Traction, parse_design, parse_design_dict)
from .elasticity_problem import (ControlSpace, ElasticPenalizer, ElasticityProblem,
                                 HelmholtzFilter, ProblemBatch, StructuredMesh,
                                 problem_from_design, solve_opts)
from .autograd import compliance

__all__ = [
    "DomainParameters", "ElasticityParameters", "Force", "CircularRegion", "Side", "Traction",
    "parse_design", "parse_design_dict", "ControlSpace", "ElasticPenalizer", "ElasticityProblem",
    "HelmholtzFilter", "ProblemBatch", "StructuredMesh", "problem_from_design", "solve_opts", "compliance",
]
