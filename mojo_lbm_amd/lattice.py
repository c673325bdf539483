"""Velocity sets (src/lbm/lattice.mojo): directions in the reference order with
opposite directions adjacent, weights, opposite map and stress index pairs."""
from __future__ import annotations

from fractions import Fraction

import numpy as np


class Lattice:
    """Lattice[D, Q]: directions (Q x D ints), weights, opposite_indices, stress_indices."""

    def __init__(self, D: int, Q: int, directions, weights):
        self.D, self.Q = D, Q
        self.directions = np.asarray(directions, dtype=np.int32).reshape(Q, D)
        self.weights = np.asarray([float(w) for w in weights], dtype=np.float64)
        self.float_directions = self.directions.astype(np.float64)
        self.stress_indices = get_stress_indices(D)
        self.n_stress_components = len(self.stress_indices)
        # _get_opposite_indices (lattice.mojo:76-91)
        self.opposite_indices = np.zeros(Q, dtype=np.int32)
        for i in range(Q):
            for k in range(Q):
                if np.array_equal(self.directions[k], -self.directions[i]):
                    self.opposite_indices[i] = k
                    break

    def is_valid_for_esoteric_pull(self) -> bool:
        """Every direction after the rest one is followed by its opposite (lattice.mojo:93-113)."""
        for q in range(1, self.Q - 1, 2):
            if (self.directions[q] + self.directions[q + 1]).sum() != 0:
                return False
        return True

    def __repr__(self) -> str:
        return f"Lattice(D{self.D}Q{self.Q})"


_FACES = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
_EDGES = [(1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1),
          (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1)]
_CORNERS = [(1, 1, 1), (-1, -1, -1), (1, 1, -1), (-1, -1, 1), (1, -1, 1), (-1, 1, -1), (-1, 1, 1), (1, -1, -1)]


def get_D2Q9() -> Lattice:
    """get_D2Q9, lattice.mojo:302-351."""
    dirs = [(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1), (-1, 1), (1, -1)]
    w = [Fraction(4, 9)] + [Fraction(1, 9)] * 4 + [Fraction(1, 36)] * 4
    return Lattice(2, 9, dirs, w)


def get_D3Q19() -> Lattice:
    """get_D3Q19, lattice.mojo:221-299."""
    dirs = [(0, 0, 0)] + _FACES + _EDGES
    w = [Fraction(1, 3)] + [Fraction(1, 18)] * 6 + [Fraction(1, 36)] * 12
    return Lattice(3, 19, dirs, w)


def get_D3Q27() -> Lattice:
    """get_D3Q27, lattice.mojo:122-218."""
    dirs = [(0, 0, 0)] + _FACES + _EDGES + _CORNERS
    w = [Fraction(8, 27)] + [Fraction(2, 27)] * 6 + [Fraction(1, 54)] * 12 + [Fraction(1, 216)] * 8
    return Lattice(3, 27, dirs, w)


def get_stress_indices(D: int):
    """get_stress_indices, lattice.mojo:354-398."""
    if D == 1:
        return [(0, 0)]
    if D == 2:
        return [(0, 0), (0, 1), (1, 1)]
    if D == 3:
        return [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    raise ValueError("D must be 1, 2 or 3")
