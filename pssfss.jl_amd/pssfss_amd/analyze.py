"""`analyze(strata, flist, steering)` — the per-point loop of src/PSSFSS.jl:114-334.

Loop shape, cut-off rule and Gblock cascade follow `_analyze`; the sheet solve
(fill -> LU -> RHS -> solve -> extract, `calculate_jtype_gsm` /
`calculate_mtype_gsm`) runs in the HIP library, batched over every point of the
sweep that shares a sheet (see `hip.SheetSolver`).

Steering is a dict whose first key is the outer loop, keys 'theta'/'phi' in
degrees or 'psi1'/'psi2' (unit-cell phasings in degrees, as the reference's
docstring states).  The reference's psi branch reads an undefined global
(src/PSSFSS.jl:612); here it works as documented.
"""
import math
from dataclasses import dataclass, field
from typing import Callable, List

import numpy as np

from . import gsm as G
from . import modes as Md
from .constants import C0, TWOPI, DBMIN
from .layers import Layer
from .rwg import setup_rwg
from .sheets import RWGSheet

_PHI_KEYS = ("phi", "ϕ", "Phi", "PHI", "Φ")
_THETA_KEYS = ("theta", "θ", "Theta", "THETA")
_PSI1_KEYS = ("psi1", "ψ₁", "Psi1", "PSI1", "ψ1")
_PSI2_KEYS = ("psi2", "ψ₂", "Psi2", "PSI2", "ψ2")


@dataclass
class Point:
    fghz: float
    k0: float
    beta00: np.ndarray
    steer: dict


def sind(x):
    return math.sin(math.radians(x))


def cosd(x):
    return math.cos(math.radians(x))


@dataclass
class Setup:
    layers: List[Layer]
    sheets: List[RWGSheet]
    junc: List[int]
    gbls: list
    gbldup: List[int]
    rwgdat: list
    uvec: List[float]
    usi: List[int]
    freqs: List[float]
    stkeys: tuple
    stvalues: list
    gsm_save: dict = field(default_factory=dict)

    def points(self):
        """Yield points in the reference's order: steer1 outer, steer2, freq innermost."""
        k1, k2 = self.stkeys
        psi_mode = k1 in _PSI1_KEYS + _PSI2_KEYS
        for so in self.stvalues[0]:
            for si in self.stvalues[1]:
                if psi_mode:
                    if k1 in _PSI1_KEYS:
                        p1, p2 = so, si
                    else:
                        p1, p2 = si, so
                    steer = {"psi1": p1, "psi2": p2}
                    sh = self.sheets[0]
                    b1, b2 = sh.beta1 * sh.upm, sh.beta2 * sh.upm
                    r1, r2 = math.radians(p1), math.radians(p2)
                    b00 = (r1 * b1 + r2 * b2) / TWOPI
                else:
                    if k1 in _PHI_KEYS:
                        th, ph = si, so
                    else:
                        th, ph = so, si
                    steer = {"theta": th, "phi": ph}
                    st = sind(th)
                    sp, cp = sind(ph), cosd(ph)
                for f in self.freqs:
                    k0 = ((TWOPI * f) * 1e9) / C0
                    if not psi_mode:
                        kk = k0 * math.sqrt((self.layers[0].epsr * self.layers[0].mur).real)
                        b00 = np.array([kk * st * cp, kk * st * sp])
                    yield Point(f, k0, np.array(b00, float), steer)

    def principal_ok(self) -> bool:
        """Skip rule of src/PSSFSS.jl:252-255 (principal modes cut off)."""
        a = np.angle(self.layers[0].gamma[0])
        b = np.angle(self.layers[-1].gamma[0])
        c = math.pi / 2
        tol = 1.4901161193847656e-08
        return abs(a - b) <= tol * max(abs(a), abs(b)) and abs(b - c) <= tol * max(abs(b), c)

    def cascade_point(self, k0: float, sheet_solver: Callable) -> G.GSM:
        """Gblock walk and cascade for one point (src/PSSFSS.jl:256-305).

        sheet_solver(region, sheet, u, rwgdat, s, isheet) returns the sheet GSM."""
        layers = self.layers
        n1 = len(layers[0].P)
        gsma = G.GSM.new(n1, n1)
        gsmc = G.GSM.new(n1, n1)
        G.cascade_layer(gsma, layers[0])
        saved = {}
        for ig, gbl in enumerate(self.gbls):
            i1 = gbl.rng[0]
            i2 = 1 + gbl.rng[-1]
            i_sheet = 0 if gbl.j == 0 else self.junc[gbl.j - 1]
            if i_sheet != 0:
                sheet = self.sheets[i_sheet - 1]
                region = layers[i1 - 1:i2]
                if self.gbldup[ig] > 0:
                    gsmb = saved[self.gbldup[ig] - 1].copy()
                else:
                    s = gbl.j - i1 + 1
                    if sheet.sheet_class in ("J", "M"):
                        gsmb = sheet_solver(region, sheet, self.uvec[i_sheet - 1],
                                            self.rwgdat[i_sheet - 1], s, i_sheet - 1)
                    elif sheet.sheet_class == "E":
                        gsmb = G.pecgsm(len(region[0].P), len(region[-1].P))
                    elif sheet.sheet_class == "H":
                        gsmb = G.pmcgsm(len(region[0].P), len(region[-1].P))
                    else:
                        raise ValueError(f"Illegal sheet class: {sheet.sheet_class}")
                    if self.gbldup[ig] < 0:
                        saved[ig] = gsmb.copy()
                if sheet.dx != 0 or sheet.dy != 0:
                    G.translate_gsm(gsmb, sheet.dx / sheet.upm, sheet.dy / sheet.upm,
                                    region[0], region[-1])
            else:
                gsmb = G.gsm_slab_interface(layers[i1 - 1], layers[i2 - 1], k0)
            gsmc = G.cascade(gsma, gsmb)
            G.cascade_layer(gsmc, layers[i2 - 1])
            gsma = gsmc
        return gsmc


def _get_gbldup(gbls, layers, sheets, junc):
    """src/PSSFSS.jl:675-704 (1-based Gblock references, -1 = save)."""
    gd = [0] * len(gbls)
    for g1, b1 in enumerate(gbls):
        if gd[g1] != 0 or b1.j == 0:
            continue
        j1, r1 = b1.j, list(b1.rng)
        n1 = len(layers[r1[0] - 1].P)
        n2 = len(layers[r1[-1]].P)
        for g2 in range(g1 + 1, len(gbls)):
            b2 = gbls[g2]
            r2, j2 = list(b2.rng), b2.j
            if gd[g2] != 0 or j2 == 0:
                continue
            if sheets[junc[j1 - 1] - 1] is not sheets[junc[j2 - 1] - 1]:
                continue
            if r1[-1] - j1 != r2[-1] - j2 or len(r1) != len(r2):
                continue
            if n1 != len(layers[r2[0] - 1].P) or n2 != len(layers[r2[-1]].P):
                continue
            same = all(i1 == r1[0] or layers[i1 - 1].electrical_equal(layers[i2 - 1])
                       for i1, i2 in zip(r1, r2))
            if not same:
                continue
            gd[g1] = -1
            gd[g2] = g1 + 1
    return gd


def _unique_indices(v):
    n = len(v)
    ui = list(range(n))
    for io in range(n - 1):
        if ui[io] != io:
            continue
        for it in range(io + 1, n):
            if ui[it] != it:
                continue
            if v[io] is v[it]:
                ui[it] = io
    return ui


def prepare(strata, flist, steering, mode_override=None, dbmin: float = DBMIN,
            classify=None) -> Setup:
    """One-time setup of `_analyze` (src/PSSFSS.jl:114-142, 213-226)."""
    import copy
    layers = [copy.deepcopy(s) for s in strata if isinstance(s, Layer)]
    sheets = [s for s in strata if isinstance(s, RWGSheet)]
    nl = len(layers)
    junc = [0] * (nl - 1)
    nlay = 0
    k = 0
    for s in strata:
        if isinstance(s, Layer):
            nlay += 1
        else:
            k += 1
            junc[nlay - 1] = k
    freqs = [float(f) for f in np.atleast_1d(np.asarray(flist, float))]
    stkeys = tuple(steering.keys())
    if len(stkeys) != 2:
        raise ValueError("steering must have exactly two keys")
    stvalues = [[float(x) for x in np.atleast_1d(np.asarray(v, float))] for v in steering.values()]
    _check_inputs(layers)
    k0min, k0max = [TWOPI * 1e9 / C0 * f for f in (min(freqs), max(freqs))]
    gbls = G.choose_gblocks(layers, sheets, junc, k0min)
    Md.choose_layer_modes(layers, sheets, junc, gbls, k0max, dbmin, mode_override=mode_override)
    gbldup = _get_gbldup(gbls, layers, sheets, junc)
    usi = _unique_indices(sheets)
    rwgdat = [setup_rwg(sh, classify=classify) if sh.style != "NULL" else None for sh in sheets]
    uvec = []
    for sh in sheets:
        if sh.style == "NULL":
            uvec.append(0.0)
        else:
            uf = 0.5 * sh.upm
            uvec.append(uf * max(float(np.linalg.norm(sh.beta1)), float(np.linalg.norm(sh.beta2))))
    return Setup(layers, sheets, junc, gbls, gbldup, rwgdat, uvec, usi, freqs, stkeys, stvalues)


def _check_inputs(layers):
    """src/PSSFSS.jl:619-633."""
    f, l = layers[0], layers[-1]
    if not (f.epsr.imag == 0 and l.epsr.imag == 0 and f.mur.imag == 0 and l.mur.imag == 0):
        raise ValueError("First and last layers must be lossless")
    if not (f.epsr == l.epsr and f.mur == l.mur):
        raise ValueError("First and last layers must have identical electrical parameters")


def analyze(strata, flist, steering, precision: str = "fp64", mode_override=None,
            ctx=None, return_stats: bool = False, dbmin: float = DBMIN, outlist=()):
    """Analyze the structure over all steering values and frequencies on the GPU.

    Returns a list of Result records (gsm, FGHz, steering, beta00, ...) in the
    reference's loop order.  precision: "fp64" (plain FP64 path) or "mixed" (FP32 LU
    and smooth sums, FP64 refinement).  Face-pair classes are formed on the device.
    outlist: (CSV file name, output functors) pairs as the reference's analyze
    takes them (src/PSSFSS.jl:67-110), e.g. [("s21.csv", outputs("FGHz",
    "s21db(te,te)"))]; each file gets a header of labels and one row per result."""
    from .hip import SheetSolver, get_context, face_pair_classes
    from .outputs import write_outputs_csv
    ctx = ctx or get_context()
    setup = prepare(strata, flist, steering, mode_override=mode_override, dbmin=dbmin,
                    classify=lambda sh: face_pair_classes(ctx, sh))
    # the device sheets (mesh and J/K cache) live for this call only
    with SheetSolver(ctx, setup, precision) as solver:
        results, stats = solver.run()
    for path, ops in outlist:
        write_outputs_csv(path, results, ops)
    return (results, stats) if return_stats else results
