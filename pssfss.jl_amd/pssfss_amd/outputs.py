"""Result records and the output functions of src/Outputs.jl.

`Result` mirrors src/Outputs.jl:25-38 (one per analysed point); `getsijmn` and the
polarisation basis changes `sourcemat` / `obsmat` follow src/Outputs.jl:73-200
(theory chapter 8); the named outputs (S..DB, AR..DB, DIPD.., DIL.., FGHZ, ...)
follow src/Outputs.jl:203-351 and `outputs(...)` stands in for the `@outputs`
macro (src/Outputs.jl:381-399), taking the same names as strings.
"""
from __future__ import annotations

import cmath
import enum
import math
import re
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Sequence, Tuple, Union

import numpy as np

from .constants import C0, TWOPI
from .gsm import GSM


class TEorTM(enum.IntEnum):
    """Mode type enumeration of src/Layers.jl:13."""
    TE = 1
    TM = 2


class HorV(enum.IntEnum):
    H = 1
    V = 2


class RorL(enum.IntEnum):
    R = 1
    L = 2


TE, TM = TEorTM.TE, TEorTM.TM
H, V = HorV.H, HorV.V
R, L = RorL.R, RorL.L


@dataclass
class Result:
    """One analysed point (src/Outputs.jl:25-38).  Wavevectors in rad/m."""
    gsm: GSM
    steering: Dict[str, float]
    beta00: np.ndarray
    FGHz: float
    epsr_in: complex = 1.0
    mur_in: complex = 1.0
    beta1_in: np.ndarray = field(default_factory=lambda: np.zeros(2))
    beta2_in: np.ndarray = field(default_factory=lambda: np.zeros(2))
    epsr_out: complex = 1.0
    mur_out: complex = 1.0
    beta1_out: np.ndarray = field(default_factory=lambda: np.zeros(2))
    beta2_out: np.ndarray = field(default_factory=lambda: np.zeros(2))

    # dictionary-style access used by the sweep drivers
    def __getitem__(self, key):
        alias = {"steer": "steering"}
        return getattr(self, alias.get(key, key))


def make_result(gsm: GSM, steering, beta00, fghz, layers) -> Result:
    """The Result of src/PSSFSS.jl:309-311 from the outer layers of the stack."""
    lin, lout = layers[0], layers[-1]
    return Result(gsm, dict(steering), np.asarray(beta00, float), float(fghz),
                  complex(lin.epsr), complex(lin.mur), np.asarray(lin.beta1, float),
                  np.asarray(lin.beta2, float), complex(lout.epsr), complex(lout.mur),
                  np.asarray(lout.beta1, float), np.asarray(lout.beta2, float))


def _block(o: Result, i: int, j: int) -> np.ndarray:
    return {(1, 1): o.gsm.s11, (1, 2): o.gsm.s12, (2, 1): o.gsm.s21, (2, 2): o.gsm.s22}[(i, j)]


def _sind(x):
    return math.sin(math.radians(x))


def _cosd(x):
    return math.cos(math.radians(x))


def theta_phi(o: Result) -> Tuple[float, float]:
    """Steering angles in degrees (src/Outputs.jl:286-304)."""
    b2 = float(o.beta00 @ o.beta00)
    if b2 == 0:
        return 0.0, float(o.steering.get("phi", 0.0))
    k2 = (TWOPI * o.FGHz * 1e9 / C0) ** 2 * (o.epsr_in * o.mur_in).real
    if b2 > k2:
        raise ValueError("Cut-off dominant mode")
    kz = math.sqrt(k2 - b2)
    th = math.degrees(math.acos(kz / math.sqrt(k2)))
    ph = math.degrees(math.atan2(o.beta00[1], o.beta00[0]))
    return th, ph


def hv_hat(theta: float, phi: float) -> Tuple[np.ndarray, np.ndarray]:
    """Ludwig-3 unit vectors (src/Outputs.jl:312-320)."""
    st, ct = _sind(theta), _cosd(theta)
    sp, cp = _sind(phi), _cosd(phi)
    th = np.array([ct * cp, ct * sp, -st])
    ph = np.array([-sp, cp, 0.0])
    return th * cp - ph * sp, th * sp + ph * cp


def _snell(o: Result, th1: float) -> float:
    n1 = math.sqrt(o.epsr_in.real * o.mur_in.real)
    n2 = math.sqrt(o.epsr_out.real * o.mur_out.real)
    return math.degrees(math.asin(n1 / n2 * _sind(th1)))


def _t12(o: Result):
    b = float(np.linalg.norm(o.beta00))
    bh = np.array([1.0, 0.0]) if b == 0 else o.beta00 / b
    return np.array([-bh[1], bh[0]]), bh          # zhat x beta_hat, beta_hat


def sourcemat(j: int, n: Union[HorV, RorL], o: Result) -> np.ndarray:
    """Right-multiplying basis change TE/TM -> H/V or R/L of the incident wave
    (src/Outputs.jl:105-148)."""
    th1, ph1 = theta_phi(o)
    if j == 1:
        th, ph, sgn = th1, ph1, 1
    else:
        th, ph, sgn = _snell(o, th1), ph1 + 180, -1
    hh, vv = hv_hat(th, ph)
    t1, t2 = _t12(o)
    ct = _cosd(th)
    if isinstance(n, HorV):
        h, v = hh[:2], vv[:2]
        return np.array([[h @ t1, v @ t1], [h @ t2 / ct, v @ t2 / ct]], dtype=complex)
    Lh = ((hh + sgn * 1j * vv) / math.sqrt(2))[:2]
    Rh = ((hh - sgn * 1j * vv) / math.sqrt(2))[:2]
    # dot products taken without conjugation (src/Outputs.jl:145-146)
    return np.array([[t1 @ Rh, t1 @ Lh], [t2 @ Rh / ct, t2 @ Lh / ct]], dtype=complex)


def obsmat(i: int, n: Union[HorV, RorL], o: Result) -> np.ndarray:
    """Left-multiplying basis change of the scattered wave (src/Outputs.jl:159-200)."""
    th1, ph1 = theta_phi(o)
    if i == 1:
        th, ph, sgn = th1, ph1 + 180, 1
    else:
        th, ph, sgn = _snell(o, th1), ph1, -1
    hh, vv = hv_hat(th, ph)
    t1_2, bh = _t12(o)
    t1 = np.array([t1_2[0], t1_2[1], 0.0])
    t2 = np.array([bh[0], bh[1], sgn * math.tan(math.radians(th))])
    if isinstance(n, HorV):
        return np.array([[hh @ t1, hh @ t2], [vv @ t1, vv @ t2]], dtype=complex)
    Lh = (hh - sgn * 1j * vv) / math.sqrt(2)
    Rh = (hh + sgn * 1j * vv) / math.sqrt(2)
    # Julia's dot conjugates its first argument (src/Outputs.jl:196-197)
    return np.array([[np.vdot(Rh, t1), np.vdot(Rh, t2)],
                     [np.vdot(Lh, t1), np.vdot(Lh, t2)]], dtype=complex)


def _is_basis(x) -> bool:
    return isinstance(x, (HorV, RorL))


def getsijmn(i: int, j: int, m, n, o: Result) -> complex:
    """Entry (m, n) of GSM partition (i, j), in the TE/TM, H/V or R/L bases
    (src/Outputs.jl:73-93).  Integer or TEorTM indices address the modal GSM."""
    blk = _block(o, i, j)
    if not _is_basis(m) and not _is_basis(n):
        return complex(blk[int(m) - 1, int(n) - 1])
    b2 = blk[:2, :2]
    if not _is_basis(m):
        return complex((b2 @ sourcemat(j, n, o))[int(m) - 1, int(n) - 1])
    if not _is_basis(n):
        return complex((obsmat(i, m, o) @ b2)[int(m) - 1, int(n) - 1])
    return complex((obsmat(i, m, o) @ b2 @ sourcemat(j, n, o))[int(m) - 1, int(n) - 1])


class Outfun:
    """A labelled output functor (src/Outputs.jl:43-48)."""

    def __init__(self, f: Callable[[Result], float], label: str):
        self.f, self.label = f, label

    def __call__(self, o: Result):
        return self.f(o)

    def __repr__(self):
        return f"Outfun({self.label})"


def _db(x: float) -> float:
    return 10 * math.log10(x) if x > 0 else -math.inf


def _sij(kind: str, i: int, j: int, m, n) -> Outfun:
    f = {"": lambda z: z, "MAG": abs, "DB": lambda z: _db(abs(z) ** 2),
         "ANG": lambda z: math.degrees(cmath.phase(z)), "REAL": lambda z: z.real,
         "IMAG": lambda z: z.imag}[kind]
    return Outfun(lambda o: f(getsijmn(i, j, m, n, o)), f"S{i}{j}{kind}({m},{n})")


def ardb(i: int, j: int, n) -> Outfun:
    """Axial ratio in dB of the wave scattered into region i by incidence n in
    region j (src/Outputs.jl:271-277)."""
    def f(o):
        jP = 1j * getsijmn(i, j, 1, n, o) / getsijmn(i, j, 2, n, o)
        Q = (1 - jP) / (1 + jP)
        aq = abs(Q)
        if aq > 1:
            aq = 1 / aq
        return math.inf if aq == 1 else 20 * math.log10((1 + aq) / (1 - aq))
    return Outfun(f, f"AR{i}{j}({n})dB")


DIPD21 = Outfun(lambda o: math.degrees(cmath.phase(getsijmn(2, 1, 1, 1, o) /
                                                   getsijmn(2, 1, 2, 2, o))), "ΔIPD")
DIPD12 = Outfun(lambda o: math.degrees(cmath.phase(getsijmn(1, 2, 1, 1, o) /
                                                   getsijmn(1, 2, 2, 2, o))), "ΔIPD")
DIL21 = Outfun(lambda o: _db(abs(getsijmn(2, 1, 1, 1, o) / getsijmn(2, 1, 2, 2, o)) ** 2), "ΔIL")
DIL12 = Outfun(lambda o: _db(abs(getsijmn(1, 2, 1, 1, o) / getsijmn(1, 2, 2, 2, o)) ** 2), "ΔIL")
FGHZ = Outfun(lambda o: o.FGHz, "FGHZ")
FMHZ = Outfun(lambda o: o.FGHz * 1000, "FMHZ")
THETA = Outfun(lambda o: o.steering.get("theta", math.nan), "THETA")
PHI = Outfun(lambda o: o.steering.get("phi", math.nan), "PHI")
PSI1 = Outfun(lambda o: o.steering.get("psi1", math.nan), "PSI1")
PSI2 = Outfun(lambda o: o.steering.get("psi2", math.nan), "PSI2")

_SCALARS = {"FGHZ": FGHZ, "FMHZ": FMHZ, "THETA": THETA, "Θ": THETA, "PHI": PHI, "Φ": PHI,
            "PSI1": PSI1, "Ψ₁": PSI1, "PSI2": PSI2, "Ψ₂": PSI2, "DIPD21": DIPD21,
            "ΔIPD21": DIPD21, "DIPD12": DIPD12, "ΔIPD12": DIPD12, "DIL21": DIL21,
            "ΔIL21": DIL21, "DIL12": DIL12, "ΔIL12": DIL12}
_IDX = {"TE": TE, "TM": TM, "H": H, "V": V, "R": R, "L": L}


def _index(tok: str):
    tok = tok.strip().upper()
    return _IDX[tok] if tok in _IDX else int(tok)


def outputs(*names: str) -> Tuple[Outfun, ...]:
    """Output functors from names as accepted by @outputs, case-insensitive, e.g.
    outputs("FGHz", "s21db(te,te)", "ar21db(h)", "dipd21")."""
    out: List[Outfun] = []
    for nm in names:
        u = nm.strip().upper()
        if u in _SCALARS:
            out.append(_SCALARS[u])
            continue
        mt = re.fullmatch(r"S([12])([12])(MAG|DB|ANG|REAL|IMAG)?\((\w+),(\w+)\)", u.replace(" ", ""))
        if mt:
            out.append(_sij(mt.group(3) or "", int(mt.group(1)), int(mt.group(2)),
                            _index(mt.group(4)), _index(mt.group(5))))
            continue
        ma = re.fullmatch(r"AR([12])([12])DB\((\w+)\)", u.replace(" ", ""))
        if ma:
            out.append(ardb(int(ma.group(1)), int(ma.group(2)), _index(ma.group(3))))
            continue
        raise ValueError(f"Illegal output request: {nm}")
    return tuple(out)


def extract_result(results: Union[Result, Sequence[Result]], ops: Sequence[Outfun]) -> np.ndarray:
    """Matrix of outputs, one row per result (src/Outputs.jl:440-446)."""
    if isinstance(results, Result):
        return np.array([[o(results) for o in ops]])
    return np.array([[o(r) for o in ops] for r in results])


def _csv_value(v) -> str:
    """One CSV field as the reference's writedlm prints it: shortest round-trip
    reals, complex values as "re + imim" / "re - imim"."""
    if isinstance(v, complex) or isinstance(v, np.complexfloating):
        re_, im_ = float(v.real), float(v.imag)
        sign = "-" if math.copysign(1.0, im_) < 0 else "+"
        return f"{re_!r} {sign} {abs(im_)!r}im"
    if isinstance(v, (bool, np.bool_)):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return repr(float(v))


def write_outputs_csv(path: str, results: Sequence[Result], ops: Sequence[Outfun]) -> None:
    """The CSV file of one `outlist` entry of analyze (src/PSSFSS.jl:316-324): a
    header row of the outputs' labels, then one row of their values per result, in
    the analysis loop's order, comma-separated."""
    with open(path, "w") as io:
        io.write(",".join(o.label for o in ops) + "\n")
        for r in results:
            io.write(",".join(_csv_value(o(r)) for o in ops) + "\n")
