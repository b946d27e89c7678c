"""Lattice-ring enumeration order.

Ring r holds the 8r lattice indices (m, n) with max(|m|, |n|) == r, visited in the
order of src/Rings.jl:9-63: the m=r column (n ascending), the m=-r column
(n ascending), then the n=-r row and the n=r row (m ascending, corners excluded).
The order is the summation order of the spatial sums and the spectral tables, so it
fixes their rounding.
"""
from typing import List, Tuple

import numpy as np


def ring(r: int) -> List[Tuple[int, int]]:
    if r < 0:
        raise ValueError("ring must be non-negative")
    if r == 0:
        return [(0, 0)]
    out = [(r, n) for n in range(-r, r + 1)]
    out += [(-r, n) for n in range(-r, r + 1)]
    out += [(m, -r) for m in range(-r + 1, r)]
    out += [(m, r) for m in range(-r + 1, r)]
    return out


def ring_array(rmax: int) -> np.ndarray:
    """All (m, n) of rings 0..rmax concatenated in summation order, shape (K, 2) int32."""
    pts = []
    for r in range(rmax + 1):
        pts.extend(ring(r))
    return np.asarray(pts, dtype=np.int32)
