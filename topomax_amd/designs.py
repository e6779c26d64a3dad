"""Design-file schema: the input side of the ElasticityProblem boundary.

Mirrors ``designs/definitions.py`` (Side :22-46, CircularRegion :90-97, Force :137-146,
Traction :150-166, ElasticityParameters :170-200, DomainParameters :204-223) and
``designs/design_parser.py:8-30`` for the elasticity designs, with the same error
behaviour (``ValueError`` on malformed sides, objectives, problems or tuples).
"""

from __future__ import annotations

import json
from dataclasses import dataclass
from enum import Enum, auto


def to_2_tuple(ray):
    if len(ray) != 2:
        raise ValueError("Got array that should have had 2 elements, "
                         f"but had {len(ray)} instead: '{ray}'")
    return (ray[0], ray[1])


class Side(Enum):
    LEFT = auto()
    RIGHT = auto()
    TOP = auto()
    BOTTOM = auto()

    @classmethod
    def get_all(cls):
        return [cls.LEFT, cls.RIGHT, cls.TOP, cls.BOTTOM]

    @classmethod
    def from_string(cls, string: str):
        table = {"Left": cls.LEFT, "Right": cls.RIGHT, "Top": cls.TOP, "Bottom": cls.BOTTOM}
        if string in table:
            return table[string]
        raise ValueError(f"Malformed side: '{string}'\nLegal sides "
                         "are: 'Left', ''Right', 'Top' or 'Bottom'")

    @property
    def label(self) -> str:
        return {Side.LEFT: "Left", Side.RIGHT: "Right", Side.TOP: "Top",
                Side.BOTTOM: "Bottom"}[self]


class ProblemType(Enum):
    FLUID = auto()
    ELASTICITY = auto()

    @classmethod
    def from_string(cls, problem: str):
        if problem == "Fluid":
            return cls.FLUID
        if problem == "Elasticity":
            return cls.ELASTICITY
        raise ValueError(f"Malformed problem: '{problem}'\nLegal problems "
                         "are: 'Fluid' and 'Elasticity'")


@dataclass
class CircularRegion:
    center: tuple
    radius: float

    @classmethod
    def from_dict(cls, d: dict):
        return cls(to_2_tuple(d["center"]), d["radius"])


@dataclass
class Force:
    region: CircularRegion
    value: tuple

    @classmethod
    def from_dict(cls, d: dict):
        return cls(CircularRegion.from_dict(d["region"]), to_2_tuple(d["value"]))


@dataclass
class Traction:
    side: Side
    center: float
    length: float
    value: tuple

    @classmethod
    def from_dict(cls, d: dict):
        return cls(Side.from_string(d["side"]), d["center"], d["length"],
                   to_2_tuple(d["value"]))

    def to_tuple(self):
        return (self.side, self.center, self.length, self.value)


@dataclass
class ElasticityParameters:
    fixed_sides: list
    body_force: Force | None
    tractions: list | None
    filter_radius: float
    young_modulus: float
    poisson_ratio: float

    @classmethod
    def from_dict(cls, d: dict):
        fixed = [Side.from_string(s) for s in d["fixed_sides"]]
        body = Force.from_dict(d["body_force"]) if d.get("body_force") is not None else None
        tractions = None
        if d.get("tractions") is not None:
            tractions = [Traction.from_dict(t) for t in d["tractions"]]
        return cls(fixed, body, tractions, d["filter_radius"], d["young_modulus"],
                   d["poisson_ratio"])


@dataclass
class DomainParameters:
    width: float
    height: float
    problem: ProblemType
    fem_step_size: float
    dem_step_size: float
    penalties: list
    volume_fraction: float

    @classmethod
    def from_dict(cls, problem: str, d: dict):
        return cls(d["width"], d["height"], ProblemType.from_string(problem),
                   d["fem_step_size"], d["dem_step_size"], d["penalties"],
                   d["volume_fraction"])


def parse_design_dict(design: dict):
    keys = list(design.keys())
    if len(keys) > 1:
        raise ValueError(f"Malformed design: got more than one key: {keys}")
    key = keys[0]
    domain = DomainParameters.from_dict(key, design[key]["domain_parameters"])
    if domain.problem != ProblemType.ELASTICITY:
        raise ValueError("topomax_amd implements the elasticity problem only")
    return domain, ElasticityParameters.from_dict(design[key]["problem_parameters"])


def parse_design(filename: str):
    """``designs/design_parser.py:8-30`` (elasticity designs)."""
    with open(filename, "rb") as fh:
        return parse_design_dict(json.load(fh))
