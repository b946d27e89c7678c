"""Synthetic workloads of the per-frequency MoM sweeps (strata + frequencies + steering).

patch_free_space: PEC 5 mm square patch in a 10 mm cell, free space,
    rectstrip Nx = Ny = 14 (392 faces, 560 unknowns), 21 frequencies 2..30 GHz,
    normal incidence.
ring_on_slab: PEC annulus a = 3.0 mm, b = 3.6 mm in a 9 mm cell on a 0.787 mm
    eps_r = 2.2, tan_delta = 0.0009 slab; theta = 30 deg, phi in {0, 45} deg;
    51 frequencies 5..25 GHz.  Structured polar mesh (nr x nphi quads, 2 faces each).
patch_two_layer: 12 mm cell, square patch on eps_r = 3.0 (1.0 mm) + eps_r = 1.1
    foam (3.0 mm); 12 frequencies 8..12 GHz at normal incidence.
"""
import numpy as np

from .layers import Layer
from .sheets import rectstrip, annular_ring


def patch_free_space(n: int = 14, nfreq: int = 21, sheet_class: str = "J"):
    sh = rectstrip(Lx=5, Ly=5, Px=10, Py=10, Nx=n, Ny=n, units="mm", sheet_class=sheet_class)
    strata = [Layer.make(), sh, Layer.make()]
    freqs = np.linspace(2.0, 30.0, nfreq)
    steering = {"theta": [0.0], "phi": [0.0]}
    return strata, freqs, steering


def ring_on_slab(nr: int = 6, nphi: int = 168, nfreq: int = 51, phis=(0.0, 45.0)):
    sh = annular_ring(a=3.0, b=3.6, P=9.0, nr=nr, nphi=nphi, units="mm")
    strata = [Layer.make(), sh, Layer.make(width=0.787, units="mm", epsr=2.2, tandel=0.0009),
              Layer.make()]
    freqs = np.linspace(5.0, 25.0, nfreq)
    steering = {"phi": list(phis), "theta": [30.0]}
    return strata, freqs, steering


def patch_two_layer(n: int = 39, nfreq: int = 12, L: float = 8.0):
    sh = rectstrip(Lx=L, Ly=L, Px=12, Py=12, Nx=n, Ny=n, units="mm")
    strata = [Layer.make(), sh, Layer.make(width=1.0, units="mm", epsr=3.0),
              Layer.make(width=3.0, units="mm", epsr=1.1), Layer.make()]
    freqs = np.linspace(8.0, 12.0, nfreq)
    steering = {"theta": [0.0], "phi": [0.0]}
    return strata, freqs, steering


def sheet_points(strata, freqs, steering):
    """(#points, #faces, #unknowns) of a single-sheet workload."""
    from .rwg import setup_rwg
    sh = [s for s in strata if not isinstance(s, Layer)][0]
    rw = setup_rwg(sh)
    nst = 1
    for v in steering.values():
        nst *= len(np.atleast_1d(v))
    return nst * len(np.atleast_1d(freqs)), sh.nface, rw.nbf
