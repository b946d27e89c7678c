"""Synthetic workloads of the per-frequency MoM sweeps (strata + frequencies + steering).

patch_free_space: PEC 5 mm square patch in a 10 mm cell, free space,
    rectstrip Nx = Ny = 14 (392 faces, 560 unknowns), 21 frequencies 2..30 GHz,
    normal incidence.
ring_on_slab: PEC annulus a = 3.0 mm, b = 3.6 mm in a 9 mm cell on a 0.787 mm
    eps_r = 2.2, tan_delta = 0.0009 slab; theta = 30 deg, phi in {0, 45} deg;
    51 frequencies 5..25 GHz.  Structured polar mesh (nr x nphi quads, 2 faces each).
patch_two_layer: 12 mm cell, square patch on eps_r = 3.0 (1.0 mm) + eps_r = 1.1
    foam (3.0 mm); 12 frequencies 8..12 GHz at normal incidence.
meander_polarizer: four PEC meanderlines (6 x 10 mm cell, 0.3 mm strips, meander
    heights drawn from a seeded U(2, 3) mm), eps_r = 1.07 foam spacers 5 mm thick,
    eps_r = 2.8 skins 0.25 mm thick; 101 frequencies 8..14 GHz at theta 0 and 40
    deg, phi = 0.  Analysed with dbmin = 68 the foam layers carry 34 modes (the
    reference's selection keeps symmetric mode groups: 30 or 34 around 32).
cross_batch: independent single-sheet problems in free space, alternately a solid
    Jerusalem cross and a loaded cross drawn from a seed: period U(8, 15) mm, arm
    length (half the tip-to-tip span) U(0.3, 0.45) P, strip width U(0.03, 0.08) P,
    800..4000 requested triangles; Jerusalem: arm and cap width = strip, cap
    length = arm length; loaded cross: arm width 4 strips, metal strip = strip.
    64 frequencies 1..30 GHz at theta 0, 15, 30, 45 deg, phi = 0.
"""
import numpy as np

from .layers import Layer
from .sheets import rectstrip, annular_ring, meander, jerusalemcross, loadedcross


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


def meander_amplitudes(seed: int = 0):
    return np.random.default_rng(seed).uniform(2.0, 3.0, 4)


def meander_polarizer(ntri: int = 8000, nfreq: int = 101, seed: int = 0, thetas=(0.0, 40.0),
                      fufp: bool = True, hs=None):
    """fufp=False keeps every face pair its own class (needed for geometry
    gradients); hs overrides the seeded amplitudes (mm)."""
    hs = meander_amplitudes(seed) if hs is None else hs
    sheets = [meander(a=6.0, b=10.0, h=float(h), w1=0.3, w2=0.3, ntri=ntri, units="mm",
                      fufp=fufp) for h in hs]
    L = Layer.make
    foam = lambda: L(width=5.0, units="mm", epsr=1.07)
    skin = lambda: L(width=0.25, units="mm", epsr=2.8)
    strata = [L(), skin(), sheets[0], foam(), sheets[1], foam(), sheets[2], foam(), sheets[3],
              skin(), L()]
    freqs = np.linspace(8.0, 14.0, nfreq)
    steering = {"phi": [0.0], "theta": list(thetas)}
    return strata, freqs, steering


def cross_batch(nprob: int = 16, nfreq: int = 64, seed: int = 0, thetas=(0.0, 15.0, 30.0, 45.0),
                ntri_range=(800, 4000)):
    """List of (strata, freqs, steering, params) for the batched cross problems."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nprob):
        P = float(rng.uniform(8.0, 15.0))
        arm = float(rng.uniform(0.3, 0.45)) * P
        strip = float(rng.uniform(0.03, 0.08)) * P
        ntri = int(rng.integers(ntri_range[0], ntri_range[1] + 1))
        if i % 2 == 0:
            sh = jerusalemcross(P=P, L1=2 * arm, L2=strip, A=arm, B=strip, w=strip, ntri=ntri,
                                units="mm")
            par = dict(kind="jerusalemcross", P=P, L1=2 * arm, L2=strip, A=arm, B=strip, w=strip)
        else:
            sh = loadedcross(s1=[P, 0.0], s2=[0.0, P], L1=2 * arm, L2=4 * strip, w=strip,
                             ntri=ntri, units="mm")
            par = dict(kind="loadedcross", P=P, L1=2 * arm, L2=4 * strip, w=strip)
        par["ntri"] = ntri
        strata = [Layer.make(), sh, Layer.make()]
        freqs = np.linspace(1.0, 30.0, nfreq)
        steering = {"phi": [0.0], "theta": list(thetas)}
        out.append((strata, freqs, steering, par))
    return out


def sheet_points(strata, freqs, steering):
    """(#points, #faces, #unknowns) of a single-sheet workload."""
    from .rwg import setup_rwg
    sh = [s for s in strata if not isinstance(s, Layer)][0]
    rw = setup_rwg(sh)
    nst = 1
    for v in steering.values():
        nst *= len(np.atleast_1d(v))
    return nst * len(np.atleast_1d(freqs)), sh.nface, rw.nbf
