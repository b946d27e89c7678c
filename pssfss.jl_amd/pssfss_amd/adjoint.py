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
                   sheet_class: str = "J", Rsheet: float = 0.0):
    """Quadrilateral patch (or, for sheet_class 'M', aperture) meshed as the
    bilinear image of a rectstrip grid; Rsheet is the surface resistance in ohms
    per square ('J' only).

    corners: (4, 2) in the cell's units, counter-clockwise from the lower left.
    Returns (sheet, vtan) with vtan[k] = d(node xy)/d(parameter k), parameters
    ordered (x0, y0, x1, y1, x2, y2, x3, y3)."""
    c = np.asarray(corners, float).reshape(4, 2)
    sh = rectstrip(Lx=1.0, Ly=1.0, Px=Px, Py=Py, Nx=Nx, Ny=Ny, units=units,
                   sheet_class=sheet_class, Rsheet=Rsheet, fufp=False)
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
def _check_adjoint_sheet(sheet, rwgdat, who: str):
    """The adjoint differentiates 'J' (resistive or not) and 'M' sheets whose face
    pairs are each their own class: a geometry direction moves the members of a
    translation class differently, so shared class terms have no shared derivative."""
    if sheet.sheet_class not in ("J", "M"):
        raise ValueError(f"{who}: 'J' or 'M' sheets")
    if not rwgdat.ufp_identity:
        raise ValueError(f"{who}: sheets meshed with fufp=False (no face-pair classes)")


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
    _check_adjoint_sheet(sheet, su.rwgdat[0], "gradients")
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


# ------------------------------------------------ reverse mode through the cascade
def _star_vjp(a: G.GSM, b: G.GSM, c_bar: G.GSM) -> Tuple[G.GSM, G.GSM]:
    """Cotangents of the star product's operands (src/GSMs.jl:125-141) for the
    bilinear pairing <X_bar, X> = sum X_bar * X (S is holomorphic in the blocks)."""
    n = b.s11.shape[0]
    I = np.eye(n)
    M1 = I - a.s22 @ b.s11
    g1 = np.linalg.solve(M1, a.s21)
    M2 = I - b.s11 @ a.s22
    g2 = np.linalg.solve(M2, b.s12)
    t1 = b.s11 @ g1
    t2 = a.s22 @ g2
    ab = G.GSM(c_bar.s11.copy(), c_bar.s11 @ t1.T + c_bar.s12 @ g2.T, None, None)
    bb = G.GSM(None, None, c_bar.s21 @ g1.T + c_bar.s22 @ t2.T, c_bar.s22.copy())
    t1b = a.s12.T @ c_bar.s11
    g1b = b.s11.T @ t1b + b.s21.T @ c_bar.s21
    t2b = b.s21.T @ c_bar.s22
    g2b = a.s12.T @ c_bar.s12 + a.s22.T @ t2b
    h1 = np.linalg.solve(M1.T, g1b)
    h2 = np.linalg.solve(M2.T, g2b)
    ab.s21 = h1
    ab.s22 = t2b @ g2.T + h1 @ (b.s11 @ g1).T + b.s11.T @ h2 @ g2.T
    bb.s12 = h2
    bb.s11 = t1b @ g1.T + a.s22.T @ h1 @ g1.T + h2 @ (a.s22 @ g2).T
    return ab, bb


def _scale(g: G.GSM, d1: np.ndarray, d2: np.ndarray, inv: bool = False) -> G.GSM:
    """diag(d1)^{+-1} [s11 s12; s21 s22] diag(d1, d2) pattern of the layer propagation
    (d1 = 1) and the translation phases."""
    r1 = 1 / d1 if inv else d1
    r2 = 1 / d2 if inv else d2
    return G.GSM((g.s11 * d1[None, :]) * r1[:, None], (g.s12 * d2[None, :]) * r1[:, None],
                 (g.s21 * d1[None, :]) * r2[:, None], (g.s22 * d2[None, :]) * r2[:, None])


def cascade_vjp(su, k0: float, blocks, w_total: G.GSM):
    """The Gblock walk of _analyze (src/PSSFSS.jl:256-305) at one point, forward and
    reverse: blocks[ig] is each Gblock's GSM before translation (layers set up at
    the point), w_total the cotangent of the total GSM.  Returns (total GSM,
    cotangents of blocks[ig])."""
    layers = su.layers
    n1 = len(layers[0].P)

    def prop(il):
        w = layers[il].width
        p = np.exp(-layers[il].gamma * w) if w != 0 else np.ones(len(layers[il].gamma))
        return p

    p0 = prop(0)
    a = _scale(G.GSM.new(n1, n1), np.ones(n1), p0)
    hist = []
    for ig, gbl in enumerate(su.gbls):
        i1, i2 = gbl.rng[0], 1 + gbl.rng[-1]
        b = blocks[ig]
        tr = None
        if gbl.j != 0:
            sheet = su.sheets[su.junc[gbl.j - 1] - 1]
            if sheet.dx != 0 or sheet.dy != 0:
                d = np.array([sheet.dx, sheet.dy]) / sheet.upm
                tr = (np.exp(1j * (layers[i1 - 1].beta @ d)), np.exp(1j * (layers[i2 - 1].beta @ d)))
                b = _scale(b, tr[0], tr[1], inv=True)
        c = G.cascade(a, b)
        p = prop(i2 - 1)
        hist.append((a, b, tr, p))
        a = G.GSM(c.s11, c.s12 * p[None, :], c.s21 * p[:, None], (c.s22 * p[None, :]) * p[:, None])
    total = a
    bars = [None] * len(su.gbls)
    ab = w_total
    for ig in range(len(su.gbls) - 1, -1, -1):
        a_in, b, tr, p = hist[ig]
        cb = G.GSM(ab.s11, ab.s12 * p[None, :], ab.s21 * p[:, None],
                   (ab.s22 * p[None, :]) * p[:, None])
        ab, bb = _star_vjp(a_in, b, cb)
        if tr is not None:
            bb = _scale(bb, tr[0], tr[1], inv=True)
        bars[ig] = bb
    return total, bars


def _wirtinger(f, g: G.GSM, rel: float = 1e-7) -> G.GSM:
    """w = df/dRe S - j df/dIm S of a real function of the total GSM by central
    differences on every entry (df = Re sum w dS)."""
    out = []
    scale = max(float(np.abs(m).max()) for m in _blocks(g) if m.size)
    h = rel * scale
    for bi, blk in enumerate(_blocks(g)):
        w = np.zeros_like(blk)
        for idx in np.ndindex(*blk.shape):
            for unit, sgn in ((1.0, 1.0), (1j, -1j)):
                gp, gm = g.copy(), g.copy()
                _blocks(gp)[bi][idx] += h * unit
                _blocks(gm)[bi][idx] -= h * unit
                w[idx] += sgn * (f(gp) - f(gm)) / (2 * h)
        out.append(w)
    return G.GSM(*out)


def gradients_stack(strata, freqs, steering, params, objective, ctx=None,
                    precision: str = "fp64", dbmin: float = None, mode_override=None):
    """Derivative of the mean over the sweep's points of objective(result) (a real
    output such as outputs.ardb(2, 1, H), evaluated on the Result of each point)
    with respect to geometry parameters of any of the stack's 'J' or 'M' sheets.

    params: list of (sheet index among the stack's sheets, vtan (nparam_s, nnode, 2)).
    Every sheet is solved on the device (fill, LU, solves); the cotangent of each
    sheet Gblock's GSM comes from a reverse pass through the cascade; one adjoint
    call per parameterised sheet contracts it with the recomputed fill derivatives.
    Returns (gradient over all parameters in params order, mean objective)."""
    from .outputs import make_result
    kw = {} if dbmin is None else {"dbmin": dbmin}
    su = prepare(strata, freqs, steering, mode_override=mode_override, **kw)
    ctx = ctx or hip.get_context()
    sheet_gbls = {}
    for ig, g in enumerate(su.gbls):
        if g.j != 0:
            sheet_gbls.setdefault(su.junc[g.j - 1] - 1, []).append(ig)
    for ish, _ in params:
        _check_adjoint_sheet(su.sheets[ish], su.rwgdat[ish], "gradients_stack")
    # forward: every point's sheet GSMs, batched per Gblock on the device
    pts = []
    for pt in su.points():
        for l in su.layers:
            Md.setup_modes(l, pt.k0, pt.beta00)
        if su.principal_ok():
            pts.append(pt)
    dsheets, builders, gsm = {}, {}, {}
    try:
        for ig, g in enumerate(su.gbls):
            if g.j == 0 or su.gbldup[ig] > 0:
                continue
            ish = su.junc[g.j - 1] - 1
            sheet = su.sheets[ish]
            if sheet.sheet_class not in ("J", "M"):
                continue
            if ish not in dsheets:
                dsheets[ish] = hip.DeviceSheet(ctx, sheet, su.rwgdat[ish])
            i1, i2 = g.rng[0], 1 + g.rng[-1]
            bl = []
            for pt in pts:
                for l in su.layers:
                    Md.setup_modes(l, pt.k0, pt.beta00)
                bl.append(hip.PointBuilder(su.layers[i1 - 1:i2], sheet, su.uvec[ish],
                                           g.j - i1 + 1, pt.k0, pt.beta00))
            builders[ig] = bl
            gs, _ = hip.sheet_gsm_batch(ctx, dsheets[ish], bl, precision)
            for ip, gg in enumerate(gs):
                gsm[(ip, ig)] = gg
        # reverse through each point's cascade
        npts = len(pts)
        values = []
        wbars = {}          # ig -> list over points of the Gblock GSM cotangent
        for ip, pt in enumerate(pts):
            for l in su.layers:
                Md.setup_modes(l, pt.k0, pt.beta00)
            blocks = []
            for ig, g in enumerate(su.gbls):
                i1, i2 = g.rng[0], 1 + g.rng[-1]
                if g.j == 0:
                    blocks.append(G.gsm_slab_interface(su.layers[i1 - 1], su.layers[i2 - 1], pt.k0))
                    continue
                src = su.gbldup[ig] - 1 if su.gbldup[ig] > 0 else ig
                sheet = su.sheets[su.junc[g.j - 1] - 1]
                if sheet.sheet_class == "E":
                    blocks.append(G.pecgsm(len(su.layers[i1 - 1].P), len(su.layers[i2 - 1].P)))
                elif sheet.sheet_class == "H":
                    blocks.append(G.pmcgsm(len(su.layers[i1 - 1].P), len(su.layers[i2 - 1].P)))
                else:
                    blocks.append(gsm[(ip, src)].copy())

            def f(gt, _pt=pt):
                return float(objective(make_result(gt, _pt.steer, _pt.beta00, _pt.fghz,
                                                   su.layers)))
            total, _ = cascade_vjp(su, pt.k0, blocks, _zero_like(G.GSM.new(
                len(su.layers[0].P), len(su.layers[-1].P))))
            values.append(f(total))
            w = _wirtinger(f, total)
            _, bars = cascade_vjp(su, pt.k0, blocks, w)
            for ig, g in enumerate(su.gbls):
                if g.j == 0 or bars[ig] is None:
                    continue
                src = su.gbldup[ig] - 1 if su.gbldup[ig] > 0 else ig
                if src not in builders:
                    continue
                acc = wbars.setdefault(src, [None] * npts)
                bb = bars[ig]
                acc[ip] = bb if acc[ip] is None else G.GSM(*[x + y for x, y in
                                                               zip(_blocks(acc[ip]), _blocks(bb))])
        # adjoint contraction per parameterised sheet
        grads = []
        L = hip.lib()
        for ish, vt in params:
            vt = np.ascontiguousarray(vt, dtype=np.float64)
            nparam = vt.shape[0]
            if vt.shape[1:] != (su.sheets[ish].rho.shape[0], 2):
                raise ValueError("vtan must be (nparam, nnode, 2)")
            gsum = np.zeros(nparam)
            for ig in sheet_gbls.get(ish, []):
                if ig not in builders:
                    continue
                wb = np.ascontiguousarray(np.stack([
                    np.concatenate([m.ravel(order="F") for m in _blocks(wbars[ig][ip])])[None, :]
                    for ip in range(npts)])).view(np.float64)
                gr = np.zeros((npts, 1, nparam), complex)
                cpts = (hip.Point * npts)(*[b.point for b in builders[ig]])
                ctx.check(L.pssfss_sheet_adjoint(ctx.h, dsheets[ish].h, npts, cpts, nparam,
                                                 hip._d(vt), 1, hip._d(wb),
                                                 hip._d(gr.view(np.float64)),
                                                 hip.PRECISION[precision]))
                gsum += gr.real.sum(axis=(0, 1))
            grads.append(gsum / npts)
        return np.concatenate(grads), float(np.mean(values))
    finally:
        for d in dsheets.values():
            d.free()


def meander_amplitude_tangent(sheet_kwargs: dict, rel: float = 1e-6):
    """d(node xy)/dh of sheets.meander at fixed mesh topology (the structured mesher
    is affine in h between the integer rounding points of its segment counts):
    returns (sheet, vtan (1, nnode, 2))."""
    from .sheets import meander
    h = float(sheet_kwargs["h"])
    d = rel * h
    sh = meander(**sheet_kwargs)
    kp = dict(sheet_kwargs, h=h + d)
    km = dict(sheet_kwargs, h=h - d)
    sp, sm = meander(**kp), meander(**km)
    if not (np.array_equal(sp.fv, sh.fv) and np.array_equal(sm.fv, sh.fv)):
        raise ValueError("meander topology changes at this amplitude")
    return sh, ((sp.rho - sm.rho) / (2 * d))[None, :, :]
