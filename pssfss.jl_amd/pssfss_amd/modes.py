"""Floquet-mode selection and per-point modal constants (src/Modes.jl)."""
import math
from typing import List

import numpy as np

from .cmath_ref import mysqrt
from .constants import ETA0, TWOPI, DBMIN
from .layers import Layer, TE, TM
from .rings import ring

MNMAX_DEFAULT = 6   # src/Modes.jl:14


def equal_layer_periodicity(l1: Layer, l2: Layer) -> bool:
    """src/Modes.jl:335-338."""
    n = np.linalg.norm
    denom = n(l1.beta1) + n(l1.beta2) + n(l2.beta1) + n(l2.beta2)
    return (n(l1.beta1 - l2.beta1) / denom < 1e-5) and (n(l1.beta2 - l2.beta2) / denom < 1e-5)


def select_pmn(layer: Layer, k0max: float, dbmin: float = -math.inf) -> None:
    """Keep modes attenuated by less than dbmin through the layer (src/Modes.jl:399-431)."""
    b1, b2 = layer.beta1, layer.beta2
    kmaxsq = k0max ** 2 * (layer.epsr * layer.mur)
    keep = []
    for mode in range(len(layer.P)):
        m, n = layer.M[mode], layer.N[mode]
        if math.isfinite(dbmin):
            if m == 0 and n == 0:
                keep.append(True)
                continue
            beta = float(np.linalg.norm(m * b1 + n * b2))
            bmin = 0.0 if beta <= k0max else beta - k0max
            g = mysqrt(bmin ** 2 - kmaxsq)
            keep.append(g.real * layer.width < dbmin / 8.6859)
        else:
            bv = m * b1 + n * b2
            g = mysqrt(float(bv @ bv) - kmaxsq)
            keep.append(g.imag >= g.real)
    M = [layer.M[i] for i in range(len(keep)) if keep[i]]
    N = [layer.N[i] for i in range(len(keep)) if keep[i]]
    uniq = []
    for mn in zip(M, N):
        if mn not in uniq:
            uniq.append(mn)
    layer.M = [mn[0] for mn in uniq for _ in (0, 1)]
    layer.N = [mn[1] for mn in uniq for _ in (0, 1)]
    layer.P = [TE, TM] * len(uniq)


def fill_mmax_pmn(layer: Layer, MNmax: int, k0max: float, dbmin: float) -> None:
    """Candidate modes of rings 0..MNmax, then `select_pmn` (src/Modes.jl:367-382)."""
    if MNmax < 0:
        raise ValueError("MNmax must be nonnegative")
    layer.M, layer.N, layer.P = [], [], []
    for r in range(MNmax + 1):
        for (m, n) in ring(r):
            layer.M += [m, m]
            layer.N += [n, n]
            layer.P += [TE, TM]
    select_pmn(layer, k0max, dbmin)


def allocate_mode_arrays(layer: Layer) -> None:
    n = len(layer.P)
    layer.beta = np.zeros((n, 2))
    layer.gamma = np.zeros(n, complex)
    layer.Y = np.zeros(n, complex)
    layer.c = np.zeros(n, complex)
    layer.tvec = np.zeros((n, 2))


def setup_modes(layer: Layer, k0: float, kvec) -> None:
    """Per-point modal constants beta, gamma, Y, c, tvec (src/Modes.jl:298-322).

    Scalar arithmetic per mode, in the same operation order as the vector form
    (beta = (beta00 + m beta1) + n beta2), with the arrays written once."""
    nm = len(layer.M)
    if nm == 0:
        return
    b0x, b0y = float(kvec[0]), float(kvec[1])
    b1x, b1y = float(layer.beta1[0]), float(layer.beta1[1])
    b2x, b2y = float(layer.beta2[0]), float(layer.beta2[1])
    area = TWOPI ** 2 / abs(b1x * b2y - b1y * b2x)
    ksq = k0 ** 2 * layer.epsr * layer.mur
    if layer.beta.shape[0] != nm:
        allocate_mode_arrays(layer)
    zte = 1j * (k0 * ETA0) * layer.mur       # TE: Y = gamma / (j k0 eta0 mur)
    ytm = 1j * (k0 / ETA0) * layer.epsr      # TM: Y = j k0 epsr / (eta0 gamma)
    beta, gamma, Ys, cs, tvec = [], [], [], [], []
    for m, n, p in zip(layer.M, layer.N, layer.P):
        bx = (b0x + m * b1x) + n * b2x
        by = (b0y + m * b1y) + n * b2y
        b2s = bx * bx + by * by
        if b2s * area < 1e-14:
            hx, hy = 1.0, 0.0
        else:
            r = math.sqrt(b2s)
            hx, hy = bx / r, by / r
        g = mysqrt(b2s - ksq)
        if p == TE:
            Y = g / zte
            tvec.append((-hy, hx))
        else:
            Y = ytm / g
            tvec.append((hx, hy))
        beta.append((bx, by))
        gamma.append(g)
        Ys.append(Y)
        cs.append(mysqrt(1 / (area * Y)))
    layer.beta[:] = beta
    layer.gamma[:] = gamma
    layer.Y[:] = Ys
    layer.c[:] = cs
    layer.tvec[:] = tvec


def choose_layer_modes(layers: List[Layer], sheets, junc, gbls, k0max: float,
                       dbmin: float = DBMIN, mode_override=None) -> None:
    """Mode index selection per layer (src/Modes.jl:49-275).

    `mode_override`, when given, is an int MNmax used instead of the default 6
    for interior layers (the reference has no such knob; the 32-mode cascaded
    workload needs it)."""
    from .gsm import find_unique_periods
    nl = len(layers)
    ngbl = len(gbls)
    mset = [False] * nl
    for g in gbls:
        for i in list(g.rng)[1:]:
            mset[i - 1] = True
    pure_radome = not any(j != 0 and sheets[j - 1].style != "NULL" for j in junc)
    mnmax_default = MNMAX_DEFAULT if mode_override is None else int(mode_override)
    if pure_radome:
        for l in layers:
            l.P, l.M, l.N = [TE, TM], [0, 0], [0, 0]
            l.beta1 = np.array([TWOPI, 0.0])
            l.beta2 = np.array([0.0, TWOPI])
        notinablock = list(range(1, nl + 1))
    else:
        def nonnull(g):
            return g.j != 0 and sheets[junc[g.j - 1] - 1].style != "NULL"
        last_gbl = max(i for i, g in enumerate(gbls) if nonnull(g))
        first_gbl = min(i for i, g in enumerate(gbls) if nonnull(g))
        last_sheet = junc[gbls[last_gbl].j - 1]
        first_sheet = junc[gbls[first_gbl].j - 1]
        upa = find_unique_periods(junc, sheets)
        for igbl in range(ngbl):
            g = gbls[igbl]
            ijunc = g.j
            if ijunc == 0:
                continue
            isht = junc[ijunc - 1]
            sheet = sheets[isht - 1]
            if sheet.style == "NULL":
                continue
            upm = sheet.upm
            b1 = sheet.beta1 * upm
            b2 = sheet.beta2 * upm
            i1 = g.rng[0]
            i2 = 1 + g.rng[-1]
            for i in (i1, i2):
                if i == 1 or i == nl:
                    continue
                igo = max(igbl - 1, 0) if i == i1 else min(igbl + 1, ngbl - 1)
                ijo = gbls[igo].j
                isho = junc[ijo - 1] if ijo != 0 else 0
                if not mset[i - 1]:
                    layers[i - 1].beta1 = b1.copy()
                    layers[i - 1].beta2 = b2.copy()
                    MNmax = mnmax_default
                    if ijo != 0 and upa[ijunc - 1] != upa[ijo - 1] and sheets[isho - 1].style != "NULL":
                        MNmax = 0
                    fill_mmax_pmn(layers[i - 1], MNmax, k0max, dbmin)
                    mset[i - 1] = True
                else:
                    if (ijo != 0 and upa[ijunc - 1] != upa[ijo - 1] and len(layers[i - 1].P) != 2
                            and sheets[isho - 1].style != "NULL"):
                        raise ValueError(f"Unequal unit cells in sheets {isht} and {isho}.")
        guard = 0
        while not all(mset[1:nl - 1]):
            guard += 1
            if guard > 10 * nl:
                raise RuntimeError("mode assignment did not converge")
            for i in range(2, nl):
                if mset[i - 1]:
                    continue
                L, R = mset[i - 2], mset[i]
                if L and R:
                    if equal_layer_periodicity(layers[i - 2], layers[i]):
                        li = layers[i - 1]
                        li.beta1 = layers[i].beta1.copy()
                        li.beta2 = layers[i].beta2.copy()
                        li.P = list(layers[i - 2].P) + list(layers[i].P)
                        li.M = list(layers[i - 2].M) + list(layers[i].M)
                        li.N = list(layers[i - 2].N) + list(layers[i].N)
                        select_pmn(li, k0max, dbmin)
                    else:
                        li = layers[i - 1]
                        li.beta1 = layers[i].beta1.copy()
                        li.beta2 = layers[i].beta2.copy()
                        li.P, li.M, li.N = [TE, TM], [0, 0], [0, 0]
                    mset[i - 1] = True
                elif L and not R:
                    if i - 1 == 1:
                        break
                    _copy_pmn(layers[i - 1], layers[i - 2])
                    select_pmn(layers[i - 1], k0max, dbmin)
                    mset[i - 1] = True
                elif R and not L:
                    if i + 1 == nl:
                        break
                    _copy_pmn(layers[i - 1], layers[i])
                    select_pmn(layers[i - 1], k0max, dbmin)
                    mset[i - 1] = True
        for l, shi in ((layers[0], first_sheet), (layers[nl - 1], last_sheet)):
            sh = sheets[shi - 1]
            l.beta1 = sh.beta1 * sh.upm
            l.beta2 = sh.beta2 * sh.upm
            fill_mmax_pmn(l, 0, k0max, dbmin)
        inagblock = set()
        for g in gbls:
            inagblock |= set(list(g.rng)[1:])
        notinablock = sorted(set(range(1, nl + 1)) - inagblock)
        for a, b in zip(notinablock[:-1], notinablock[1:]):
            if b - a != 1:
                continue
            if equal_layer_periodicity(layers[a - 1], layers[a]):
                continue
            for l in (layers[a - 1], layers[a]):
                l.P, l.M, l.N = [TE, TM], [0, 0], [0, 0]
    for i in notinablock:
        allocate_mode_arrays(layers[i - 1])


def _copy_pmn(l1: Layer, l2: Layer) -> None:
    l1.beta1 = l2.beta1.copy()
    l1.beta2 = l2.beta2.copy()
    l1.P = list(l2.P)
    l1.M = list(l2.M)
    l1.N = list(l2.N)
