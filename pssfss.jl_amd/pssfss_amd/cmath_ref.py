"""Scalar complex helpers with the reference's branch conventions."""
import cmath


def mysqrt(z: complex) -> complex:
    """sqrt, with a pure negative-imaginary result flipped to +j (src/PGF.jl:116-120)."""
    a = cmath.sqrt(complex(z))
    if a.real == 0.0 and a.imag < 0.0:
        a = -a
    return a


def sinhcosh(x: complex):
    """src/GSMs.jl:893-901."""
    ex = cmath.exp(x)
    emx = 1 / ex
    return 0.5 * (ex - emx), 0.5 * (ex + emx)


def cis(t: float) -> complex:
    return complex(cmath.cos(t).real, cmath.sin(t).real)
