"""Adjoint geometry gradients (no reference counterpart) against central finite
differences.  Objective |S21|^2 co-polar TE and TM of a corner-parameterised patch
on the two-layer stack (eps_r 3.0, 1.0 mm; eps_r 1.1 foam, 3.0 mm).

The reference algorithm is piecewise smooth: its 6-point table interpolant jumps
across table cells, the spatial sums stop at a 1e-8 ring test and the singular
integrals switch formula at |p0/l| = 1e-4, so S moves by ~1e-9 when a quadrature
point crosses a cell.  The adjoint differentiates the smooth piece the base point
lies in (branches and cells are taken on values).  The reference gradient is
therefore the oracle's central difference with those decisions frozen at the base
point (oracle.FrozenDecisions): recorded there, replayed at the displaced
geometries.  It is smooth to ~1e-8 across steps of 1e-6 to 1e-4 of the period, so
central differences at 1e-6 of the period are a 1e-5-grade reference.

Bars: 1e-5 relative (to the largest derivative of each objective and point) for the
frozen-decision references, on the 5 x 5 mesh (all 8 parameters, normal and
oblique incidence) and the full 39 x 39 mesh (two parameters).  The unfrozen
Richardson difference of the GPU path itself is kept as a coarser check (3e-5:
cell crossings along its larger steps limit it)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CORNERS = np.array([[2.1, 2.4], [9.7, 2.0], [10.1, 9.6], [2.3, 9.9]])
P = 12.0
OBJ = [(2, 1, 1, 1), (2, 1, 2, 2)]


def _stack(pk, corners, n=5, **kw):
    from pssfss_amd.adjoint import bilinear_patch
    sh, vt = bilinear_patch(corners, n, n, P, P, **kw)
    L = pk.Layer.make
    strata = [L(), sh, L(width=1.0, units="mm", epsr=3.0), L(width=3.0, units="mm", epsr=1.1),
              L()]
    return strata, vt


def _objectives(results):
    out = []
    for r in results:
        g = r["gsm"]
        blk = {(1, 1): g.s11, (1, 2): g.s12, (2, 1): g.s21, (2, 2): g.s22}
        out.append([abs(blk[(i, j)][m - 1, n - 1]) ** 2 for (i, j, m, n) in OBJ])
    return np.array(out)


def _cd(pk, analyze, k, freqs, steer, h):
    c1 = CORNERS.copy().reshape(-1)
    c2 = CORNERS.copy().reshape(-1)
    c1[k] += h
    c2[k] -= h
    s1, _ = _stack(pk, c1.reshape(4, 2))
    s2, _ = _stack(pk, c2.reshape(4, 2))
    return (_objectives(analyze(s1, freqs, steer)) - _objectives(analyze(s2, freqs, steer))) / (2 * h)


def _fd(pk, analyze, k, freqs, steer, h):
    return (4 * _cd(pk, analyze, k, freqs, steer, h) - _cd(pk, analyze, k, freqs, steer, 2 * h)) / 3


def test_adjoint_vs_unfrozen_fd_of_gpu_path(ctx, pk):
    from pssfss_amd.adjoint import gradients
    freqs = [8.0, 12.0]
    steer = {"theta": [0.0], "phi": [0.0]}
    strata, vt = _stack(pk, CORNERS)
    print("\n[case] adjoint gradients, 5x5 bilinear patch, 8 parameters, 2 freqs", flush=True)
    grad, vals = gradients(strata, freqs, steer, vt, OBJ, ctx=ctx)
    assert grad.shape == (2, 2, 8)
    fwd = _objectives(pk.analyze.analyze(strata, freqs, steer, precision="fp64", ctx=ctx))
    assert np.abs(vals - fwd).max() < 1e-12
    h = 1e-3 * P

    def gpu(s, f, st):
        return pk.analyze.analyze(s, f, st, precision="fp64", ctx=ctx)
    fd = np.stack([_fd(pk, gpu, k, freqs, steer, h) for k in range(8)], axis=-1)
    scale = np.abs(fd).max(axis=-1, keepdims=True)
    rel = np.abs(grad - fd) / scale
    print(f"  max |grad - FD| / max|FD| = {rel.max():.2e}, median {np.median(rel):.2e}; "
          f"max|FD| = {scale.max():.3e}")
    assert rel.max() < 3e-5
    assert np.median(rel) < 5e-6


def _frozen_fd(pk, orc, corners, n, params, freqs, steer, h, **kw):
    """Central differences of the oracle with its decisions frozen at `corners`."""
    strata, _ = _stack(pk, corners, n, **kw)
    nufp = pk.analyze.prepare(strata, freqs, steer).rwgdat[0].nufp
    fz = orc.FrozenDecisions(nufp, len(freqs) * len(steer["theta"]) * len(steer["phi"]))
    base = _objectives(orc.analyze(pk, strata, freqs, steer, decisions=(fz, "record")))
    out = []
    for k in params:
        c1 = corners.copy().reshape(-1)
        c2 = corners.copy().reshape(-1)
        c1[k] += h
        c2[k] -= h
        f1 = _objectives(orc.analyze(pk, _stack(pk, c1.reshape(4, 2), n, **kw)[0], freqs,
                                     steer, decisions=(fz, "replay")))
        f2 = _objectives(orc.analyze(pk, _stack(pk, c2.reshape(4, 2), n, **kw)[0], freqs,
                                     steer, decisions=(fz, "replay")))
        out.append((f1 - f2) / (2 * h))
    return np.stack(out, axis=-1), base


@pytest.mark.parametrize("steer", [{"theta": [0.0], "phi": [0.0]},
                                   {"theta": [20.0], "phi": [30.0]}])
def test_adjoint_vs_frozen_oracle_fd(ctx, pk, orc, steer):
    from pssfss_amd.adjoint import gradients
    freqs = [8.0, 12.0]
    print(f"\n[case] adjoint vs frozen-decision oracle FD, 5x5 patch, 8 parameters, {steer}",
          flush=True)
    strata, vt = _stack(pk, CORNERS)
    grad, vals = gradients(strata, freqs, steer, vt, OBJ, ctx=ctx)
    fd, base = _frozen_fd(pk, orc, CORNERS, 5, range(8), freqs, steer, 1e-6 * P)
    assert np.abs(vals - base).max() < 1e-12
    rel = np.abs(grad - fd) / np.abs(fd).max(axis=-1, keepdims=True)
    print(f"  max |grad - FD| / max|FD| = {rel.max():.2e}, median {np.median(rel):.2e}")
    assert rel.max() < 1e-5


@pytest.mark.parametrize("kind", ["aperture", "resistive"])
def test_adjoint_magnetic_and_resistive_vs_frozen_oracle_fd(ctx, pk, orc, kind):
    """The same corner parameters for a quadrilateral aperture ('M' sheet: the
    magnetic class terms, sources z x t and extraction sigma tl c) and for a patch
    of 20 ohm/square surface resistance (the self-pair resistance term moves with
    each face's area and edge lengths), at oblique incidence; bar 1e-5."""
    from pssfss_amd.adjoint import gradients
    kw = {"sheet_class": "M"} if kind == "aperture" else {"Rsheet": 20.0}
    freqs = [8.0, 12.0]
    steer = {"theta": [20.0], "phi": [30.0]}
    print(f"\n[case] adjoint vs frozen-decision oracle FD, 5x5 {kind}, 8 parameters",
          flush=True)
    strata, vt = _stack(pk, CORNERS, **kw)
    grad, vals = gradients(strata, freqs, steer, vt, OBJ, ctx=ctx)
    fd, base = _frozen_fd(pk, orc, CORNERS, 5, range(8), freqs, steer, 1e-6 * P, **kw)
    assert np.abs(vals - base).max() < 1e-12
    rel = np.abs(grad - fd) / np.abs(fd).max(axis=-1, keepdims=True)
    print(f"  max |grad - FD| / max|FD| = {rel.max():.2e}, median {np.median(rel):.2e}; "
          f"max|FD| = {np.abs(fd).max():.3e}")
    assert rel.max() < 1e-5


def test_adjoint_full_mesh_vs_frozen_oracle_fd(ctx, pk, orc):
    """The full 39 x 39 mesh (4485 unknowns) at 10 GHz, parameters x0 and y2."""
    from pssfss_amd.adjoint import gradients
    freqs = [10.0]
    steer = {"theta": [0.0], "phi": [0.0]}
    corners = np.array([[2.0, 2.0], [10.0, 2.0], [10.0, 10.0], [2.0, 10.0]])
    print("\n[case] adjoint vs frozen-decision oracle FD, 39x39 patch, parameters 0 and 5",
          flush=True)
    strata, vt = _stack(pk, corners, 39)
    grad, _ = gradients(strata, freqs, steer, vt, OBJ, ctx=ctx)
    fd, _ = _frozen_fd(pk, orc, corners, 39, (0, 5), freqs, steer, 1e-6 * P)
    rel = np.abs(grad[:, :, [0, 5]] - fd) / np.abs(fd).max(axis=-1, keepdims=True)
    print(f"  grad {grad[0, :, [0, 5]].ravel()}, FD {fd[0].ravel()}; max rel {rel.max():.2e}")
    assert rel.max() < 1e-5


def test_adjoint_mixed_vs_fp64(ctx, pk):
    """The mixed form (FP32 LU, FP64 refinement of the forward and the transposed
    systems) against the FP64 adjoint on the same points: gradients agree to 1e-6
    relative to the largest gradient of each objective and point."""
    from pssfss_amd.adjoint import gradients
    freqs = [8.0, 12.0]
    steer = {"theta": [0.0], "phi": [0.0]}
    strata, vt = _stack(pk, CORNERS)
    print("\n[case] mixed vs fp64 adjoint, 5x5 bilinear patch, 8 parameters, 2 freqs", flush=True)
    g64, v64 = gradients(strata, freqs, steer, vt, OBJ, ctx=ctx)
    g32, v32 = gradients(strata, freqs, steer, vt, OBJ, ctx=ctx, precision="mixed")
    rel = np.abs(g32 - g64) / np.abs(g64).max(axis=-1, keepdims=True)
    print(f"  max |grad_mixed - grad_fp64| / max|grad| = {rel.max():.2e}; "
          f"max |f_mixed - f_fp64| = {np.abs(v32 - v64).max():.2e}")
    assert rel.max() < 1e-6
    assert np.abs(v32 - v64).max() < 1e-6


def test_stack_gradient_meander_axial_ratio(ctx, pk, orc):
    """The cascaded 4-sheet meanderline polarizer (rotated 45 deg, so that it turns a
    linear wave circular): gradient of the mean axial ratio AR21(H) over 10 GHz at
    theta 0 and 40 deg with respect to the 4 meander amplitudes, through the reverse
    pass of the cascade and one adjoint call per sheet, against frozen-decision
    oracle central differences at 1e-6 of the amplitude (bar 1e-5)."""
    from pssfss_amd import adjoint as A
    from pssfss_amd import outputs as O
    print("\n[case] stack gradient: mean AR21(H) of the 4-sheet meander w.r.t. amplitudes",
          flush=True)
    hs = [2.2, 2.5, 2.7, 2.4]
    L = pk.Layer.make

    def build(hv):
        sheets, vts = [], []
        for h in hv:
            sh, vt = A.meander_amplitude_tangent(dict(a=6.0, b=10.0, h=h, w1=0.3, w2=0.3,
                                                      ntri=120, units="mm", fufp=False,
                                                      rot=45.0))
            sheets.append(sh)
            vts.append(vt)
        foam = lambda: L(width=5.0, units="mm", epsr=1.07)
        skin = lambda: L(width=0.25, units="mm", epsr=2.8)
        return [L(), skin(), sheets[0], foam(), sheets[1], foam(), sheets[2], foam(),
                sheets[3], skin(), L()], vts
    freqs, steer = [10.0], {"phi": [0.0], "theta": [0.0, 40.0]}
    obj = O.ardb(2, 1, O.H)
    strata, vts = build(hs)
    grad, val = A.gradients_stack(strata, freqs, steer, [(i, vts[i]) for i in range(4)], obj,
                                  ctx=ctx, dbmin=30.0)
    su = pk.analyze.prepare(strata, freqs, steer, dbmin=30.0)
    fz = orc.FrozenDecisions([r.nufp for r in su.rwgdat], 2)
    base = np.mean([obj(r) for r in orc.analyze(pk, strata, freqs, steer, dbmin=30.0,
                                                 decisions=(fz, "record"))])
    assert abs(val - base) < 1e-9 * abs(base)
    fd = []
    for i in range(4):
        d = 1e-6 * hs[i]
        vals = []
        for sgn in (1, -1):
            hv = list(hs)
            hv[i] += sgn * d
            s2, _ = build(hv)
            res = orc.analyze(pk, s2, freqs, steer, dbmin=30.0, decisions=(fz, "replay"))
            vals.append(np.mean([obj(r) for r in res]))
        fd.append((vals[0] - vals[1]) / (2 * d))
    fd = np.array(fd)
    rel = np.abs(grad - fd).max() / np.abs(fd).max()
    print(f"  mean AR {val:.4f} dB; grad {grad}, FD {fd}; max rel {rel:.2e}")
    assert rel < 1e-5
