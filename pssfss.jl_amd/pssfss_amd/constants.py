"""Physical constants and tolerances of the MoM path.

Values follow src/Constants.jl:6-28 of PSSFSS.jl exactly (mu0 is the CODATA-2018
value, not 4*pi*1e-7), because every impedance entry scales with them.
"""
import math

MU0 = 1.25663706212 * 1e-6          # src/Constants.jl:6
C0 = 299792458.0                    # src/Constants.jl:9
EPS0 = 1.0 / (MU0 * C0 ** 2)        # src/Constants.jl:12
ETA0 = math.sqrt(MU0 / EPS0)        # src/Constants.jl:15
TWOPI = 2.0 * math.pi
FOURPI = 4.0 * math.pi
TOL = 1e-4                          # src/Constants.jl:19 (J/K cache key tolerance on u)
DBMIN = 30.0                        # src/Constants.jl:22
MIN_ELENGTH = TWOPI * 0.1           # src/Constants.jl:25

# Length units: value = units per meter (Unitful's exact rationals rounded once).
UNITS_PER_METER = {
    "m": 1.0,
    "cm": 100.0,
    "mm": 1000.0,
    "inch": 10000.0 / 254.0,
    "mil": 10000000.0 / 254.0,
}


def units_per_meter(units: str) -> float:
    try:
        return UNITS_PER_METER[units]
    except KeyError:
        raise ValueError(f"unknown length unit {units!r}") from None


METERS_PER_UNIT = {
    "m": 1.0,
    "cm": 0.01,
    "mm": 0.001,
    "inch": 0.0254,
    "mil": 2.54e-05,
}


def to_meters(value: float, units: str) -> float:
    """Length in `units` converted to meters (Unitful's ustrip(u"m", value*units))."""
    if units not in METERS_PER_UNIT:
        raise ValueError(f"unknown length unit {units!r}")
    return float(value) * METERS_PER_UNIT[units]
