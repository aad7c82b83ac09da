"""Beam and Lattice containers, mirroring src/lattice.jl:33-132, 323-360."""

from __future__ import annotations

import math
from typing import Iterable, List

import numpy as np

from .elements import AbstractElement, TimeVaryingElement, get_length

__all__ = ["Beam", "beti", "Lattice", "total_length", "spos", "findelem", "materialize_lattice",
           "M_ELECTRON", "M_PROTON"]

# src/lattice.jl:42-43
M_ELECTRON = 0.51099895069e6
M_PROTON = 938.27208816e6


class Beam:
    """Beam(energy; mass=M_ELECTRON, charge=-1.0) -- src/lattice.jl:33-54.

    ``energy`` is treated as kinetic energy, as in the reference: gamma = (E + m) / m.
    """

    def __init__(self, energy: float, *, mass: float = M_ELECTRON, charge: float = -1.0):
        energy = float(energy)
        mass = float(mass)
        gamma = (energy + mass) / mass
        beta = math.sqrt(1.0 - 1.0 / (gamma * gamma))
        self.energy = energy
        self.mass = mass
        self.charge = float(charge)
        self.gamma = gamma
        self.beta = beta

    def __repr__(self) -> str:
        return f"Beam(energy={self.energy!r}, mass={self.mass!r}, gamma={self.gamma!r}, beta={self.beta!r})"


def beti(beam: Beam) -> float:
    """1/beta (src/lattice.jl:61)."""
    return 1.0 / beam.beta


class Lattice:
    """Lattice(elements; name=:LATTICE) -- src/lattice.jl:87-105."""

    def __init__(self, elements: Iterable[AbstractElement], *, name: str = "LATTICE"):
        self.elements: List[AbstractElement] = list(elements)
        self.name = str(name)

    def __len__(self) -> int:
        return len(self.elements)

    def __getitem__(self, i):
        return self.elements[i]

    def __iter__(self):
        return iter(self.elements)


def total_length(lat: Lattice) -> float:
    """Sum of element lengths (src/lattice.jl:113-120)."""
    L = 0.0
    for e in lat.elements:
        L += float(get_length(e))
    return L


def spos(lat: Lattice) -> np.ndarray:
    """s-position of each element boundary (src/lattice.jl:329-336)."""
    s = np.zeros(len(lat) + 1)
    for i, e in enumerate(lat.elements):
        s[i + 1] = s[i] + float(get_length(e))
    return s


def findelem(lat: Lattice, name: str) -> List[int]:
    """1-based indices of the elements named ``name`` (src/lattice.jl:351-360)."""
    out = []
    for i, e in enumerate(lat.elements):
        n = e.base.name if isinstance(e, TimeVaryingElement) else getattr(e, "name", None)
        if n == str(name):
            out.append(i + 1)
    return out


def materialize_lattice(lat: Lattice, *, time: float = 0.0, turn: int = 0) -> Lattice:
    """Static lattice with time-varying elements resolved (src/lattice.jl:128-132)."""
    elems = [e.materialize(time, turn) if isinstance(e, TimeVaryingElement) else e for e in lat.elements]
    return Lattice(elems, name=lat.name)
