"""End-to-end GPU parity at the workloads' full mesh sizes (through the C ABI),
against the oracle (FP64 CPU restatement with its own transmission-line side):
S to 1e-9 for the FP64 path and 1e-6 for the mixed path.

* the 5 mm patch in free space: rectstrip 14 x 14 (392 faces, 560 unknowns), all
  21 frequencies 2..30 GHz at normal incidence;
* the annular ring on its slab: the full 2016-face mesh at one oblique point;
* the two-layer patch: the full 39 x 39 bilinear mesh (4485 unknowns) at one point;
* the ring with magnetic currents ('M' sheet, the aperture dual) at 30 degrees.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"\n[case] {msg}", flush=True)


def _max_dS(res, ref):
    m = 0.0
    for a, b in zip(res, ref):
        for blk in ("s11", "s12", "s21", "s22"):
            m = max(m, np.abs(getattr(a["gsm"], blk) - getattr(b["gsm"], blk)).max())
    return m


def _check(pk, orc, ctx, strata, freqs, steer, npts):
    ref = orc.analyze(pk, strata, freqs, steer)
    res = pk.analyze.analyze(strata, freqs, steer, precision="fp64", ctx=ctx)
    assert len(res) == len(ref) == npts
    d = _max_dS(res, ref)
    resm = pk.analyze.analyze(strata, freqs, steer, precision="mixed", ctx=ctx)
    dm = _max_dS(resm, ref)
    print(f"  max |dS| fp64 {d:.3e}, mixed {dm:.3e}", flush=True)
    assert d <= 1e-9
    assert dm <= 1e-6


def test_patch_free_space_full(ctx, pk, orc):
    _say("patch in free space, rectstrip 14 x 14, 21 frequencies 2..30 GHz")
    strata, freqs, steer = pk.workloads.patch_free_space()
    _check(pk, orc, ctx, strata, freqs, steer, 21)


def test_ring_on_slab_full_mesh_point(ctx, pk, orc):
    _say("ring on slab, full mesh (2016 faces), 17 GHz, theta 30, phi 45")
    strata, _, _ = pk.workloads.ring_on_slab()
    _check(pk, orc, ctx, strata, [17.0], {"phi": [45.0], "theta": [30.0]}, 1)


def test_two_layer_patch_full_mesh_point(ctx, pk, orc):
    _say("two-layer patch, 39 x 39 bilinear mesh (4485 unknowns), 10 GHz")
    from pssfss_amd import adjoint as A
    corners = [[2.0, 2.0], [10.0, 2.0], [10.0, 10.0], [2.0, 10.0]]
    sh, _ = A.bilinear_patch(corners, 39, 39, 12.0, 12.0)
    L = pk.Layer.make
    strata = [L(), sh, L(width=1.0, units="mm", epsr=3.0), L(width=3.0, units="mm", epsr=1.1),
              L()]
    _check(pk, orc, ctx, strata, [10.0], {"theta": [0.0], "phi": [0.0]}, 1)


def test_magnetic_ring_oblique(ctx, pk, orc):
    _say("ring aperture ('M' sheet, nr=3, nphi=48) on slab, theta 30, phi 0/45, 2 freqs")
    sh = pk.sheets.annular_ring(a=3.0, b=3.6, P=9.0, nr=3, nphi=48, units="mm",
                                sheet_class="M")
    L = pk.Layer.make
    strata = [L(), sh, L(width=0.787, units="mm", epsr=2.2, tandel=0.0009), L()]
    _check(pk, orc, ctx, strata, [9.0, 16.0], {"phi": [0.0, 45.0], "theta": [30.0]}, 4)
