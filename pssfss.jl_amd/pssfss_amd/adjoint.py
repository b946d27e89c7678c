"""Geometry gradients of S-parameter objectives (no counterpart in the reference).

Mesh nodes move along directions vtan[k] (sheet units per unit parameter).  The
device library returns, per point, sum_e wbar_e dS_e/dt_k for the sheet's Gblock
GSM (adjoint solves with the same LU factors, the fill recomputed with forward-
mode derivatives and contracted on the fly).  The host supplies wbar = df/dS_sheet
from a linearised cascade (cascade_point_jvp) of the stack around the point.

`bilinear_patch` builds the corner-parameterised patch of the differentiation
workload: a rectstrip grid mapped bilinearly onto a quadrilateral with corners
(x0, y0), (x1, y1), (x2, y2), (x3, y3), eight parameters.
"""
from __future__ import annotations

import copy
import ctypes as C
from typing import List, Sequence, Tuple

import numpy as np

from . import gsm as G
from . import hip
from . import modes as Md
from .analyze import prepare
from .layers import Layer
from .sheets import rectstrip


# ------------------------------------------------------------------ geometry
def bilinear_patch(corners, Nx: int, Ny: int, Px: float, Py: float, units: str = "mm",
                   sheet_class: str = "J"):
    """Quadrilateral PEC patch meshed as the bilinear image of a rectstrip grid.

    corners: (4, 2) in the cell's units, counter-clockwise from the lower left.
    Returns (sheet, vtan) with vtan[k] = d(node xy)/d(parameter k), parameters
    ordered (x0, y0, x1, y1, x2, y2, x3, y3)."""
    c = np.asarray(corners, float).reshape(4, 2)
    sh = rectstrip(Lx=1.0, Ly=1.0, Px=Px, Py=Py, Nx=Nx, Ny=Ny, units=units,
                   sheet_class=sheet_class, fufp=False)
    x0, y0 = 0.5 * (Px - 1.0), 0.5 * (Py - 1.0)
    xi = np.clip(sh.rho[:, 0] - x0, 0.0, 1.0)
    eta = np.clip(sh.rho[:, 1] - y0, 0.0, 1.0)
    N = np.stack([(1 - xi) * (1 - eta), xi * (1 - eta), xi * eta, (1 - xi) * eta], axis=1)
    sh.rho = N @ c
    sh.style = "bilinear_patch"
    sh.xieta_check = False
    vtan = np.zeros((8, sh.rho.shape[0], 2))
    for k in range(4):
        vtan[2 * k, :, 0] = N[:, k]
        vtan[2 * k + 1, :, 1] = N[:, k]
    return sh, vtan


# ------------------------------------------------------- linearised cascade
def _cascade_jvp(a: G.GSM, da: G.GSM, b: G.GSM, db: G.GSM) -> Tuple[G.GSM, G.GSM]:
    """Star product (src/GSMs.jl:125-141) and its directional derivative."""
    n = b.s11.shape[0]
    I = np.eye(n)
    M1 = I - a.s22 @ b.s11
    dM1 = -(da.s22 @ b.s11 + a.s22 @ db.s11)
    g1 = np.linalg.solve(M1, a.s21)
    dg1 = np.linalg.solve(M1, da.s21 - dM1 @ g1)
    M2 = I - b.s11 @ a.s22
    dM2 = -(db.s11 @ a.s22 + b.s11 @ da.s22)
    g2 = np.linalg.solve(M2, b.s12)
    dg2 = np.linalg.solve(M2, db.s12 - dM2 @ g2)
    c = G.GSM(a.s11 + a.s12 @ (b.s11 @ g1), a.s12 @ g2, b.s21 @ g1,
              b.s22 + b.s21 @ (a.s22 @ g2))
    dc = G.GSM(da.s11 + da.s12 @ (b.s11 @ g1) + a.s12 @ (db.s11 @ g1 + b.s11 @ dg1),
               da.s12 @ g2 + a.s12 @ dg2,
               db.s21 @ g1 + b.s21 @ dg1,
               db.s22 + db.s21 @ (a.s22 @ g2) + b.s21 @ (da.s22 @ g2 + a.s22 @ dg2))
    return c, dc


def _zero_like(g: G.GSM) -> G.GSM:
    return G.GSM(*[np.zeros_like(m) for m in (g.s11, g.s12, g.s21, g.s22)])


def cascade_point_jvp(su, k0: float, sheet_gsm: G.GSM, dsheet: G.GSM):
    """Setup.cascade_point for a single-sheet stack, with the derivative of the
    total GSM for a perturbation dsheet of the sheet's Gblock GSM."""
    layers = su.layers
    n1 = len(layers[0].P)
    a = G.GSM.new(n1, n1)
    G.cascade_layer(a, layers[0])
    da = _zero_like(a)
    for ig, gbl in enumerate(su.gbls):
        i1, i2 = gbl.rng[0], 1 + gbl.rng[-1]
        i_sheet = 0 if gbl.j == 0 else su.junc[gbl.j - 1]
        if i_sheet != 0:
            sheet = su.sheets[i_sheet - 1]
            region = layers[i1 - 1:i2]
            b, db = sheet_gsm.copy(), dsheet.copy()
            if sheet.dx != 0 or sheet.dy != 0:
                for g in (b, db):
                    G.translate_gsm(g, sheet.dx / sheet.upm, sheet.dy / sheet.upm,
                                    region[0], region[-1])
        else:
            b = G.gsm_slab_interface(layers[i1 - 1], layers[i2 - 1], k0)
            db = _zero_like(b)
        a, da = _cascade_jvp(a, da, b, db)
        G.cascade_layer(a, layers[i2 - 1])
        G.cascade_layer(da, layers[i2 - 1])
    return a, da


def _blocks(g: G.GSM):
    return [g.s11, g.s12, g.s21, g.s22]


# ------------------------------------------------------------------ driver
def gradients(strata, freqs, steering, vtan, objectives: Sequence[Tuple[int, int, int, int]],
              ctx=None, precision: str = "fp64"):
    """Derivatives of f_o = |S_ij(m, n)|^2 of the total GSM (1-based i, j, m, n) with
    respect to the parameters whose node directions are vtan (nparam, nnode, 2).
    precision "fp64", or "mixed" (FP32 LU with FP64 refinement for the forward and
    the adjoint solves; the fill and its derivatives stay FP64).

    Returns (grad, values): grad (npoints, nobj, nparam), values (npoints, nobj) of f_o,
    for the points of the sweep in analyze order (cut-off points skipped)."""
    su = prepare(strata, freqs, steering)
    if len(su.sheets) != 1:
        raise ValueError("gradients: single-sheet structures")
    sheet = su.sheets[0]
    if sheet.sheet_class != "J":
        raise ValueError("gradients: electric-current ('J') sheets")
    ctx = ctx or hip.get_context()
    vt = np.ascontiguousarray(vtan, dtype=np.float64)
    nparam = vt.shape[0]
    if vt.shape[1:] != (sheet.rho.shape[0], 2):
        raise ValueError("vtan must be (nparam, nnode, 2)")
    ds = hip.DeviceSheet(ctx, sheet, su.rwgdat[0])
    try:
        return _gradients_on(ctx, ds, su, sheet, vt, nparam, objectives, precision)
    finally:
        ds.free()


def _gradients_on(ctx, ds, su, sheet, vt, nparam, objectives, precision):
    sheet_gbl = [ig for ig, g in enumerate(su.gbls) if g.j != 0]
    if len(sheet_gbl) != 1 or su.gbldup[sheet_gbl[0]] > 0:
        raise ValueError("gradients: one sheet Gblock")
    ig = sheet_gbl[0]
    gbl = su.gbls[ig]
    i1, i2 = gbl.rng[0], 1 + gbl.rng[-1]
    s = gbl.j - i1 + 1
    points, builders = [], []
    for pt in su.points():
        for l in su.layers:
            Md.setup_modes(l, pt.k0, pt.beta00)
        if not su.principal_ok():
            continue
        region = su.layers[i1 - 1:i2]
        builders.append(hip.PointBuilder(region, sheet, su.uvec[0], s, pt.k0, pt.beta00))
        points.append(pt)
    if not points:
        return np.zeros((0, len(objectives), nparam)), np.zeros((0, len(objectives)))
    sheet_gsms, _ = hip.sheet_gsm_batch(ctx, ds, builders, precision)
    nobj = len(objectives)
    wsets, values = [], []
    for pt, gs in zip(points, sheet_gsms):
        for l in su.layers:
            Md.setup_modes(l, pt.k0, pt.beta00)
        tot, _ = cascade_point_jvp(su, pt.k0, gs, _zero_like(gs))
        sv = [_blocks(tot)[2 * (i - 1) + (j - 1)][m - 1, n - 1] for (i, j, m, n) in objectives]
        values.append([abs(v) ** 2 for v in sv])
        w = np.zeros((nobj, sum(b.size for b in _blocks(gs))), complex)
        off = 0
        for bi, blk in enumerate(_blocks(gs)):
            for col in range(blk.shape[1]):          # column-major entry order
                for row in range(blk.shape[0]):
                    d = _zero_like(gs)
                    _blocks(d)[bi][row, col] = 1.0
                    _, dt = cascade_point_jvp(su, pt.k0, gs, d)
                    for o, (i, j, m, n) in enumerate(objectives):
                        dso = _blocks(dt)[2 * (i - 1) + (j - 1)][m - 1, n - 1]
                        w[o, off] = 2 * np.conj(sv[o]) * dso
                    off += 1
        wsets.append(w)
    npts = len(points)
    wbar = np.ascontiguousarray(np.stack(wsets)).view(np.float64)
    grad = np.zeros((npts, nobj, nparam), complex)
    pts = (hip.Point * npts)(*[b.point for b in builders])
    L = hip.lib()
    ctx.check(L.pssfss_sheet_adjoint(ctx.h, ds.h, npts, pts, nparam, hip._d(vt), nobj,
                                     hip._d(wbar), hip._d(grad.view(np.float64)),
                                     hip.PRECISION[precision]))
    return grad.real, np.array(values)
