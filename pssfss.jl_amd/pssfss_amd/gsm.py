"""Generalized scattering matrices and the transmission-line partial GSMs (src/GSMs.jl).

These are the per-point O(n_modes * n_layers) host computations that bracket the
sheet solve: the incident-field partial GSM with its TLGF and incident-voltage
vectors (gsm_electric_gblock / gsm_magnetic_gblock), the dielectric interface GSM,
Gblock partitioning, and the single-point Redheffer star product.  The batched
device cascade lives in the HIP library (`hip.cascade_batched`).
"""
from dataclasses import dataclass
from typing import List

import numpy as np

from .cmath_ref import mysqrt, sinhcosh
from .constants import ETA0, MIN_ELENGTH
from .layers import Layer, TE


@dataclass
class GSM:
    s11: np.ndarray
    s12: np.ndarray
    s21: np.ndarray
    s22: np.ndarray

    @staticmethod
    def new(n1: int, n2: int) -> "GSM":
        """src/GSMs.jl:23-29: zero reflection, unit diagonal transmission."""
        g = GSM(np.zeros((n1, n1), complex), np.zeros((n1, n2), complex),
                np.zeros((n2, n1), complex), np.zeros((n2, n2), complex))
        k = min(n1, n2)
        g.s12[np.arange(k), np.arange(k)] = 1
        g.s21[np.arange(k), np.arange(k)] = 1
        return g

    def __getitem__(self, ij):
        return {(1, 1): self.s11, (1, 2): self.s12, (2, 1): self.s21, (2, 2): self.s22}[ij]

    def copy(self) -> "GSM":
        return GSM(self.s11.copy(), self.s12.copy(), self.s21.copy(), self.s22.copy())


def pecgsm(n1, n2) -> GSM:
    g = GSM.new(n1, n2)
    g.s11[np.diag_indices(n1)] = -1
    g.s22[np.diag_indices(n2)] = -1
    g.s12[:] = 0
    g.s21[:] = 0
    return g


def pmcgsm(n1, n2) -> GSM:
    g = GSM.new(n1, n2)
    g.s11[np.diag_indices(n1)] = 1
    g.s22[np.diag_indices(n2)] = 1
    g.s12[:] = 0
    g.s21[:] = 0
    return g


def cascade(a: GSM, b: GSM) -> GSM:
    """Redheffer star product, eq. (3.35), src/GSMs.jl:125-141."""
    n2a = a.s22.shape[1]
    n1b = b.s11.shape[0]
    if n1b != n2a:
        raise ValueError("Non-conformable arrays")
    I = np.eye(n1b)
    g1 = np.linalg.solve(I - a.s22 @ b.s11, a.s21)
    s21 = b.s21 @ g1
    s11 = a.s11 + (a.s12 @ (b.s11 @ g1))
    g2 = np.linalg.solve(I - b.s11 @ a.s22, b.s12)
    s12 = a.s12 @ g2
    s22 = b.s22 + (b.s21 @ (a.s22 @ g2))
    return GSM(s11, s12, s21, s22)


def cascade_layer(a: GSM, layer: Layer, t: float = float("nan")) -> None:
    """Propagate the side-2 ports through `layer` (cascade!, src/GSMs.jl:155-176)."""
    n2 = a.s22.shape[0]
    if n2 != len(layer.gamma):
        raise ValueError("# modes not consistent")
    d = layer.width if t != t else t
    if d == 0:
        return
    for i in range(n2):
        p = np.exp(-layer.gamma[i] * d)
        a.s12[:, i] *= p
        a.s21[i, :] *= p
        a.s22[:, i] *= p
        a.s22[i, :] *= p


def find_mode_index(p, m, n, layer: Layer) -> int:
    """1-based index of mode (p, m, n) in `layer`, 0 if absent (src/GSMs.jl:884-891)."""
    for i in range(len(layer.P)):
        if layer.P[i] == p and layer.M[i] == m and layer.N[i] == n:
            return i + 1
    return 0


def _zdotcross(a, b):
    return a[0] * b[1] - a[1] * b[0]


def gsm_electric_gblock(layers: List[Layer], s: int, k0: float):
    """Incident partial GSM, TLGF and incident voltages for a J sheet (src/GSMs.jl:213-370).

    Returns (gsm, tlgfvi, vincs); the two matrices are (max(n1, n2), 2)."""
    N = len(layers)
    cosha = np.zeros(N + 1, complex)
    sinha = np.zeros(N + 1, complex)
    tanha = np.zeros(N + 1, complex)
    Z0 = np.zeros(N + 1, complex)
    Zl = np.zeros(N + 1, complex)
    Zr = np.zeros(N + 1, complex)
    k0sq = k0 * k0
    L1, LN = layers[0], layers[N - 1]
    n1, n2 = len(L1.P), len(LN.P)
    gsm = GSM.new(n1, n2)
    nm = max(n1, n2)
    vincs = np.zeros((nm, 2), complex)
    tlgf = np.zeros((nm, 2), complex)
    root = np.sqrt(abs(_zdotcross(LN.beta1, LN.beta2) / _zdotcross(L1.beta1, L1.beta2)))
    L = lambda i: layers[i - 1]   # 1-based layer access
    for q1 in range(n1):
        Z0[1] = 1 / L1.Y[q1]
        Zl[1] = Z0[1]
        b2 = float(L1.beta[q1] @ L1.beta[q1])
        for i in range(2, N + 1):
            k2 = k0sq * L(i).mur * L(i).epsr
            g = mysqrt(b2 - k2)
            if L1.P[q1] == TE:
                Z0[i] = (1j * (k0 * ETA0) * L(i).mur) / g
            else:
                Z0[i] = g / (1j * (k0 / ETA0) * L(i).epsr)
            if i < N:
                sinha[i], cosha[i] = sinhcosh(L(i).width * g)
                tanha[i] = sinha[i] / cosha[i]
                Zl[i] = Z0[i] * (Zl[i - 1] + Z0[i] * tanha[i]) / (Z0[i] + Zl[i - 1] * tanha[i])
        Zr[N - 1] = Z0[N]
        for i in range(N - 2, 0, -1):
            Zr[i] = Z0[i + 1] * (Zr[i + 1] + Z0[i + 1] * tanha[i + 1]) / (
                Z0[i + 1] + Zr[i + 1] * tanha[i + 1])
        cur = 2 / (Zr[1] + Z0[1])
        vol = cur * Zr[1]
        gsm.s11[q1, q1] = vol - 1
        for i in range(2, s + 1):
            vol = vol * cosha[i] - Z0[i] * cur * sinha[i]
            cur = vol / Zr[i]
        vincs[q1, 0] = vol
        q2 = find_mode_index(L1.P[q1], L1.M[q1], L1.N[q1], LN)
        if q2 != 0:
            for i in range(s + 1, N):
                vol = vol * cosha[i] - Z0[i] * cur * sinha[i]
                cur = vol / Zr[i]
            gsm.s21[q2 - 1, q1] = vol * L1.c[q1] / LN.c[q2 - 1] * root
        cur = -Zr[s] / (Zl[s] + Zr[s])
        vol = -Zl[s] * cur
        for i in range(s - 1, 0, -1):
            vol = vol * cosha[i + 1] + Z0[i + 1] * cur * sinha[i + 1]
            cur = -vol / Zl[i]
        tlgf[q1, 0] = vol
    for q2 in range(n2):
        Z0[N] = 1 / LN.Y[q2]
        Zr[N - 1] = Z0[N]
        b2 = float(LN.beta[q2] @ LN.beta[q2])
        for i in range(N - 1, 0, -1):
            k2 = k0sq * L(i).mur * L(i).epsr
            g = mysqrt(b2 - k2)
            if LN.P[q2] == TE:
                Z0[i] = (1j * (k0 * ETA0) * L(i).mur) / g
            else:
                Z0[i] = g / (1j * (k0 / ETA0) * L(i).epsr)
            if i > 1:
                sinha[i], cosha[i] = sinhcosh(L(i).width * g)
                tanha[i] = sinha[i] / cosha[i]
                Zr[i - 1] = Z0[i] * (Zr[i] + Z0[i] * tanha[i]) / (Z0[i] + Zr[i] * tanha[i])
        Zl[1] = Z0[1]
        for i in range(2, N):
            Zl[i] = Z0[i] * (Zl[i - 1] + Z0[i] * tanha[i]) / (Z0[i] + Zl[i - 1] * tanha[i])
        cur = -2 / (Zl[N - 1] + Z0[N])
        vol = -cur * Zl[N - 1]
        gsm.s22[q2, q2] = vol - 1
        for i in range(N - 2, s - 1, -1):
            vol = vol * cosha[i + 1] + Z0[i + 1] * cur * sinha[i + 1]
            cur = -vol / Zl[i]
        vincs[q2, 1] = vol
        q1 = find_mode_index(LN.P[q2], LN.M[q2], LN.N[q2], L1)
        if q1 != 0:
            for i in range(s - 1, 0, -1):
                vol = vol * cosha[i + 1] + Z0[i + 1] * cur * sinha[i + 1]
                cur = -vol / Zl[i]
            gsm.s12[q1 - 1, q2] = vol * LN.c[q2] / L1.c[q1 - 1] / root
        cur = Zl[s] / (Zl[s] + Zr[s])
        vol = Zr[s] * cur
        for i in range(s + 1, N):
            vol = vol * cosha[i] - Z0[i] * cur * sinha[i]
            cur = vol / Zr[i]
        tlgf[q2, 1] = vol
    return gsm, tlgf, vincs


def gsm_magnetic_gblock(layers: List[Layer], s: int, k0: float):
    """Partial GSM, TLGF and incident currents for an M sheet (src/GSMs.jl:401-526)."""
    N = len(layers)
    cosha = np.zeros(N + 1, complex)
    sinha = np.zeros(N + 1, complex)
    tanha = np.zeros(N + 1, complex)
    Z0 = np.zeros(N + 1, complex)
    Zl = np.zeros(N + 1, complex)
    Zr = np.zeros(N + 1, complex)
    k0sq = k0 * k0
    L1, LN = layers[0], layers[N - 1]
    n1, n2 = len(L1.P), len(LN.P)
    gsm = GSM.new(n1, n2)
    gsm.s12[:] = 0
    gsm.s21[:] = 0
    nm = max(n1, n2)
    iincs = np.zeros((nm, 2), complex)
    tlgf = np.zeros((nm, 2), complex)
    L = lambda i: layers[i - 1]
    for q1 in range(n1):
        Z0[1] = 1 / L1.Y[q1]
        Zl[1] = Z0[1]
        b2 = float(L1.beta[q1] @ L1.beta[q1])
        for i in range(2, s + 1):
            k2 = k0sq * L(i).mur * L(i).epsr
            g = mysqrt(b2 - k2)
            if L1.P[q1] == TE:
                Z0[i] = (1j * (k0 * ETA0) * L(i).mur) / g
            else:
                Z0[i] = g / (1j * (k0 / ETA0) * L(i).epsr)
            sinha[i], cosha[i] = sinhcosh(L(i).width * g)
            tanha[i] = sinha[i] / cosha[i]
            Zl[i] = Z0[i] * (Zl[i - 1] + Z0[i] * tanha[i]) / (Z0[i] + Zl[i - 1] * tanha[i])
        Zr[s] = 0
        for i in range(s - 1, 0, -1):
            Zr[i] = Z0[i + 1] * (Zr[i + 1] + Z0[i + 1] * tanha[i + 1]) / (
                Z0[i + 1] + Zr[i + 1] * tanha[i + 1])
        cur = 2 / (1 + Zr[1] / Z0[1])
        vol = cur * Zr[1]
        gsm.s11[q1, q1] = 1 - cur
        for i in range(2, s + 1):
            cur = cur * cosha[i] - vol * sinha[i] / Z0[i]
            vol = cur * Zr[i]
        iincs[q1, 0] = cur
        cur = 1 / Zl[s]
        vol = -1.0 + 0j
        for i in range(s - 1, 0, -1):
            cur = cur * cosha[i + 1] + vol * sinha[i + 1] / Z0[i + 1]
            vol = -cur * Zl[i]
        tlgf[q1, 0] = cur
    for q2 in range(n2):
        Z0[N] = 1 / LN.Y[q2]
        Zr[N - 1] = Z0[N]
        b2 = float(LN.beta[q2] @ LN.beta[q2])
        for i in range(N - 1, s, -1):
            k2 = k0sq * L(i).mur * L(i).epsr
            g = mysqrt(b2 - k2)
            if LN.P[q2] == TE:
                Z0[i] = (1j * (k0 * ETA0) * L(i).mur) / g
            else:
                Z0[i] = g / (1j * k0 / ETA0 * L(i).epsr)
            sinha[i], cosha[i] = sinhcosh(L(i).width * g)
            tanha[i] = sinha[i] / cosha[i]
            Zr[i - 1] = Z0[i] * (Zr[i] + Z0[i] * tanha[i]) / (Z0[i] + Zr[i] * tanha[i])
        Zl[s] = 0
        for i in range(s + 1, N):
            Zl[i] = Z0[i] * (Zl[i - 1] + Z0[i] * tanha[i]) / (Z0[i] + Zl[i - 1] * tanha[i])
        cur = -2 / (1 + Zl[N - 1] / Z0[N])
        vol = -cur * Zl[N - 1]
        gsm.s22[q2, q2] = 1 + cur
        for i in range(N - 2, s - 1, -1):
            cur = cur * cosha[i + 1] + vol * sinha[i + 1] / Z0[i + 1]
            vol = -cur * Zl[i]
        iincs[q2, 1] = cur
        cur = 1 / Zr[s]
        vol = 1.0 + 0j
        for i in range(s + 1, N):
            cur = cur * cosha[i] - vol * sinha[i] / Z0[i]
            vol = cur * Zr[i]
        tlgf[q2, 1] = cur
    return gsm, tlgf, iincs


def gsm_slab_interface(L1: Layer, L2: Layer, k0: float) -> GSM:
    """Dielectric junction GSM, eqs. (3.23)-(3.24), src/GSMs.jl:545-620."""
    n1, n2 = len(L1.P), len(L2.P)
    g = GSM.new(n1, n2)
    for i1 in range(n1):
        Y1 = L1.Y[i1]
        i2 = find_mode_index(L1.P[i1], L1.M[i1], L1.N[i1], L2)
        if i2 == 0:
            k2sq = float(L2.beta[0] @ L2.beta[0]) - L2.gamma[0] ** 2
            b2 = float(L1.beta[i1] @ L1.beta[i1])
            gm = mysqrt(b2 - k2sq)
            Y2 = gm / (1j * (k0 * ETA0) * L2.mur) if L1.P[i1] == TE else 1j * (k0 / ETA0) * L2.epsr / gm
        else:
            Y2 = L2.Y[i2 - 1]
        g.s11[i1, i1] = (Y1 - Y2) / (Y2 + Y1)
        if i2 != 0:
            g.s21[i2 - 1, i1] = 2 * mysqrt(Y2) * mysqrt(Y1) / (Y2 + Y1)
    for i2 in range(n2):
        Y2 = L2.Y[i2]
        i1 = find_mode_index(L2.P[i2], L2.M[i2], L2.N[i2], L1)
        if i1 == 0:
            k1sq = float(L1.beta[0] @ L1.beta[0]) - L1.gamma[0] ** 2
            b2 = float(L2.beta[i2] @ L2.beta[i2])
            gm = mysqrt(b2 - k1sq)
            Y1 = gm / (1j * (k0 * ETA0) * L1.mur) if L2.P[i2] == TE else 1j * (k0 / ETA0) * L1.epsr / gm
        else:
            Y1 = L1.Y[i1 - 1]
        R = (Y1 - Y2) / (Y2 + Y1)
        g.s22[i2, i2] = -R
        if i1 != 0:
            g.s12[i1 - 1, i2] = 2 * mysqrt(Y2) * mysqrt(Y1) / (Y2 + Y1)
    return g


def translate_gsm(g: GSM, dx: float, dy: float, layer1: Layer, layer2: Layer) -> None:
    """Phase shifts for a translated sheet (src/GSMs.jl:717-750)."""
    d = np.array([dx, dy])
    t1 = np.exp(1j * (layer1.beta @ d))
    t2 = np.exp(1j * (layer2.beta @ d))
    g.s11 *= t1[None, :] / t1[:, None]
    g.s12 *= t2[None, :] / t1[:, None]
    g.s21 *= t1[None, :] / t2[:, None]
    g.s22 *= t2[None, :] / t2[:, None]


@dataclass
class Gblock:
    rng: range       # 1-based interface numbers
    j: int           # interface holding the sheet, 0 if none


def find_unique_periods(junc, sheets):
    """src/Sheets.jl:139-169: class number of each sheet's unit cell, 0 where no sheet."""
    def cell(sh):
        v = np.concatenate([sh.s1, sh.s2]) / sh.upm
        return tuple(float(f"{x:.7e}") for x in v)     # round to 8 significant digits
    upa = [0] * len(junc)
    reps = []
    for i, js in enumerate(junc):
        if js == 0 or sheets[js - 1].style == "NULL":
            continue
        c = cell(sheets[js - 1])
        if c in reps:
            upa[i] = reps.index(c) + 1
        else:
            reps.append(c)
            upa[i] = len(reps)
    return upa


def choose_gblocks(layers: List[Layer], sheets, junc, k0min: float) -> List[Gblock]:
    """Gblock partition of the interfaces (src/GSMs.jl:771-868)."""
    nl = len(layers)
    nj = nl - 1
    ns = len(sheets)
    if ns == 0 or all(t.style == "NULL" for t in sheets):
        return [Gblock(range(i, i + 1), 0 if junc[i - 1] == 0 else i) for i in range(1, nj + 1)]
    el = [0.0] + [k0min * l.width * np.sqrt((l.epsr * l.mur).real) for l in layers]
    sint = [i + 1 for i in range(nj) if junc[i] != 0]
    owner = list(range(0, nj + 1))   # owner[i], 1-based
    sh = lambda j: sheets[junc[j - 1] - 1]
    if sheets[0].style != "NULL":
        for j in range(sint[0], 1, -1):
            if el[j] > MIN_ELENGTH:
                break
            owner[j - 1] = sint[0]
    if sheets[-1].style != "NULL":
        for j in range(1 + sint[ns - 1], nj + 1):
            if el[j] > MIN_ELENGTH:
                break
            owner[j] = sint[-1]
    for i in range(ns - 1):
        j1, j2 = sint[i], sint[i + 1]
        nwide = sum(1 for j in range(1 + j1, j2 + 1) if el[j] > MIN_ELENGTH)
        if nwide > 0:
            for j in range(j2, j1, -1):
                if el[j] > MIN_ELENGTH or sh(j2).style == "NULL":
                    break
                owner[j - 1] = j2
            for j in range(1 + j1, j2 + 1):
                if el[j] > MIN_ELENGTH or sh(j1).style == "NULL":
                    break
                owner[j] = j1
        else:
            n1, n2 = sh(j1).style == "NULL", sh(j2).style == "NULL"
            if n1 and not n2:
                for j in range(j1 + 1, j2 + 1):
                    owner[j] = j2
                continue
            if not n1 and n2:
                for j in range(j1, j2):
                    owner[j] = j1
                continue
            if n1 and n2:
                continue
            elmax = max(el[j1 + 1:j2 + 1])
            best = [j for j in range(1, nl + 1) if el[j] == elmax and j1 + 1 <= j <= j2]
            mid = (j1 + j2 + 1) // 2
            ib = int(np.argmin([abs(b - mid) for b in best]))
            lb = best[ib]
            for j in range(j1, lb):
                owner[j] = j1
            for j in range(lb, j2 + 1):
                owner[j] = j2
    gbl = []
    for j in range(1, nj + 1):
        if gbl and owner[j] == owner[gbl[-1].rng[0]]:
            gbl[-1].rng = range(gbl[-1].rng[0], j + 1)
        else:
            gbl.append(Gblock(range(j, j + 1), 0))
        if junc[j - 1] != 0:
            gbl[-1].j = j
    flat = [i for g in gbl for i in g.rng]
    if flat != list(range(1, nj + 1)):
        raise ValueError(f"Bad Gblock vector: {gbl}")
    return gbl
