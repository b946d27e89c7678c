"""d Z / d k0 of the fill (no reference counterpart) at fixed phasing and u,
against central differences in k0 of the GPU FP64 fill and of the oracle's fill
(the CPU restatement of fillz with its own tables and spatial sums).

At fixed phasing Z is a smooth function of k0 as long as the table size M does
not change, since no quadrature point moves: the difference quotients below
check that M is the same at every step.  The check uses the Richardson-
extrapolated central difference (4 D(h) - D(2h)) / 3 (O(h^4) truncation) with
h = 1e-3 k0; both agree with the kernels' derivative to ~4e-12 relative.

Z is dominated by its scalar-potential term, which scales as 1/k0, so the test
also requires that dZ departs from -Z/k0 (the prefactor alone) by at least 1e4
times the bar: what the tables, the extraction constants and the vector
potential contribute is then resolved to better than 1e-4 of itself (measured:
~1e-7 on the strip at 2 GHz, where that part is 3e-5 of dZ, and ~1e-10 on the
ring at 17 GHz, where it is 4e-2).  Both sheet classes."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"\n[case] {msg}", flush=True)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def _ring_case(pk, cls="J"):
    strata, _, _ = pk.workloads.ring_on_slab(nr=2, nphi=24, nfreq=1)
    layers = [s for s in strata if isinstance(s, pk.Layer)]
    sh = [s for s in strata if not isinstance(s, pk.Layer)][0]
    sh.sheet_class = cls
    k0 = 2 * math.pi * 17e9 / 299792458.0
    kinc = k0 * math.sin(math.radians(30)) * np.array([math.cos(0.7), math.sin(0.7)])
    u = 0.5 * max(np.linalg.norm(sh.beta1), np.linalg.norm(sh.beta2)) * sh.upm
    return layers, sh, u, k0, kinc, 1


def _strip_case(pk, cls="J"):
    FGHz = 2.0
    k0 = 2 * math.pi / (11.80285 / FGHz * 0.0254)
    L = pk.Layer.make
    layers = [L(), L(epsr=3.4, tandel=0.003, width=1, units="mil"),
              L(epsr=3.4, tandel=0.003, width=1, units="mil")]
    sh = pk.rectstrip(Nx=12, Ny=12, Lx=20, Ly=50, Px=100, Py=100, units="mil", sheet_class=cls)
    u = 0.5 * max(np.linalg.norm(sh.beta1), np.linalg.norm(sh.beta2)) * sh.upm
    return layers, sh, u, k0, np.zeros(2), 2


def _richardson(f, k0, h):
    d1 = (f(k0 + h) - f(k0 - h)) / (2 * h)
    d2 = (f(k0 + 2 * h) - f(k0 - 2 * h)) / (4 * h)
    return (4 * d1 - d2) / 3


@pytest.mark.parametrize("cls", ["J", "M"])
@pytest.mark.parametrize("case", ["ring", "strip"])
def test_fill_dk0_vs_central_differences(ctx, pk, case, cls):
    _say(f"fill d/dk0 {case} class {cls}: GPU derivative against differences of the GPU fill")
    from pssfss_amd.hip import DeviceSheet, PointBuilder
    layers, sh, u, k0, kinc, s = (_ring_case if case == "ring" else _strip_case)(pk, cls)
    rw = pk.setup_rwg(sh)
    ds = DeviceSheet(ctx, sh, rw)
    pb = PointBuilder(layers, sh, u, s, k0, kinc, with_modes=False)
    Z0, info0 = ds.fill(pb)
    Z, dZ, info = ds.fill_dk0(pb)
    assert np.array_equal(Z, Z0) and info.mmax == info0.mmax
    mm = set()

    def fill_at(k):
        pb.point.k0 = k
        z, inf = ds.fill(pb)
        mm.add(inf.mmax)
        return z

    h = 1e-3 * k0
    fd = _richardson(fill_at, k0, h)
    ds.free()
    assert mm == {info.mmax}, f"table size changed along the steps: {mm}"
    err = _rel(dZ, fd)
    dev = _rel(dZ, -Z / k0)
    print(f"  nbf={rw.nbf} M={info.mmax}  |dZ - FD| / |dZ| = {err:.3e}  "
          f"|dZ + Z/k0| / |dZ| = {dev:.3e}")
    assert err <= 1e-9
    assert dev >= 1e4 * 1e-9


@pytest.mark.parametrize("cls", ["J", "M"])
def test_fill_dk0_vs_oracle_differences(ctx, pk, orc, cls):
    _say(f"fill d/dk0 ring class {cls}, oblique: GPU derivative against differences of the "
         "oracle fill")
    from pssfss_amd.hip import DeviceSheet, PointBuilder
    layers, sh, u, k0, kinc, s = _ring_case(pk, cls)
    rw = pk.setup_rwg(sh)
    ds = DeviceSheet(ctx, sh, rw)
    pb = PointBuilder(layers, sh, u, s, k0, kinc, with_modes=False)
    _, dZ, info = ds.fill_dk0(pb)
    ds.free()
    psi1, psi2 = pb.psi1, pb.psi2
    upm = sh.upm
    b1m, b2m = sh.beta1 * upm, sh.beta2 * upm
    b00 = (psi1 * b1m + psi2 * b2m) / (2 * math.pi)
    JK, rr, ri = orc.filljk(sh, rw, u / upm, psi1, psi2)
    mm = set()

    def fill_at(k):
        tabs = orc.modal_tables(cls == "M", k, u, psi1, psi2, layers, s, b1m, b2m, b00, 1e-7)
        mm.add(tabs[0])
        return orc.fill(sh, rw, JK, rr, ri, k, u, layers, s, tabs[:3], b1m, b2m, psi1, psi2)

    fd = _richardson(fill_at, k0, 1e-3 * k0)
    assert mm == {info.mmax}, f"table size changed along the steps: {mm}"
    err = _rel(dZ, fd)
    print(f"  |dZ - FD(oracle)| / |dZ| = {err:.3e}")
    assert err <= 1e-9
