"""`Layer`: one dielectric layer of the stack (src/Layers.jl:51-87).

Permittivity and permeability are stored complex as eps*(1 - j tan_delta)
(src/Layers.jl:80-81); the width is kept in meters.  The per-mode arrays
(P, M, N, beta, gamma, Y, c, tvec) are filled by `modes.choose_layer_modes` and
`modes.setup_modes` exactly where the reference fills them.
"""
from dataclasses import dataclass, field
from typing import List

import numpy as np

from .constants import to_meters

TE = 1
TM = 2


@dataclass(eq=False)
class Layer:
    name: str = "Layer"
    epsr: complex = 1.0 + 0.0j
    mur: complex = 1.0 + 0.0j
    width: float = 0.0          # meters
    user_width: float = 0.0     # in user units
    units: str = "mm"
    P: List[int] = field(default_factory=list)
    M: List[int] = field(default_factory=list)
    N: List[int] = field(default_factory=list)
    beta1: np.ndarray = field(default_factory=lambda: np.zeros(2))
    beta2: np.ndarray = field(default_factory=lambda: np.zeros(2))
    beta: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))
    gamma: np.ndarray = field(default_factory=lambda: np.zeros(0, complex))
    Y: np.ndarray = field(default_factory=lambda: np.zeros(0, complex))
    c: np.ndarray = field(default_factory=lambda: np.zeros(0, complex))
    tvec: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))

    @staticmethod
    def make(width: float = 0.0, units: str = "mm", epsr: float = 1.0, tandel: float = 0.0,
             mur: float = 1.0, mtandel: float = 0.0, name: str = "Layer") -> "Layer":
        """Keyword constructor with the reference's defaults (src/Layers.jl:73-86)."""
        ceps = epsr * complex(1.0, -tandel)
        cmu = mur * complex(1.0, -mtandel)
        return Layer(name=name, epsr=ceps, mur=cmu, width=to_meters(width, units),
                     user_width=float(width), units=units)

    def electrical_equal(self, other: "Layer") -> bool:
        return (self.epsr == other.epsr and self.mur == other.mur and
                self.width == other.width and list(self.P) == list(other.P) and
                list(self.M) == list(other.M) and list(self.N) == list(other.N))

    def copy(self) -> "Layer":
        import copy
        return copy.deepcopy(self)
