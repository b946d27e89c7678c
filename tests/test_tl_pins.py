"""Pins of the per-point transmission-line side of the path (CPU only).

* Mode and Gblock selection against the reference's own goldens
  (test/Modes_test.jl:17-60, test/GSMs_test.jl:140-271).
* The oracle's transmission-line restatement (nodal networks, oracle/tl_oracle.py)
  against the package's recursions (src/GSMs.jl:213-620 restated in
  pssfss_amd/gsm.py): two independent computations of the same circuits.
* Closed forms: a bare dielectric slab at oblique incidence (TE and TM) against
  the two-interface Airy formula; the incident partial GSM of a sheet Gblock
  against the same region's junction-and-layer cascade (a vanishing sheet); a
  full-cell resistive sheet; power conservation of a lossless dielectric stack
  carrying a PEC sheet at 30 degrees.
"""
import math

import numpy as np
import pytest


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _split(pk, strata):
    layers = [s for s in strata if isinstance(s, pk.Layer)]
    sheets = [s for s in strata if not isinstance(s, pk.Layer)]
    junc = [0] * (len(layers) - 1)
    nlay, k = 0, 0
    for s in strata:
        if isinstance(s, pk.Layer):
            nlay += 1
        else:
            k += 1
            junc[nlay - 1] = k
    return layers, sheets, junc


def _polyring(pk, P=0.5117):
    return pk.polyring(sides=4, orient=45, a=[0.1902], b=[0.2169], units="inch",
                       s1=[P, 0], s2=[0, P], ntri=100)


def _k0(fghz):
    return 2 * math.pi * fghz * 1e9 / 299792458.0


# ------------------------------------------------------------------ selection goldens
@pytest.mark.parametrize("case,expect", [
    (1, [2, 0, 0, 18, 0, 0, 2]),
    (2, [2, 0, 0, 2, 0, 0, 0, 2]),
    (3, [2, 0, 0, 42, 0, 0, 0, 2]),
])
def test_choose_layer_modes_golden(pk, case, expect):
    """test/Modes_test.jl:17-60 (the sheets supply only their lattices)."""
    L = pk.Layer.make
    duroid = lambda: L(width=20, units="mil", epsr=3.6, tandel=0.0013)
    if case == 1:
        strata = [L(), duroid(), _polyring(pk), duroid(), L(width=0.2, units="inch"), duroid(),
                  _polyring(pk), duroid(), L()]
    elif case == 2:
        strata = [L(), duroid(), _polyring(pk), duroid(), L(width=0.1, units="inch"),
                  L(width=0.1, units="inch"), duroid(), _polyring(pk, 0.9005), duroid(), L()]
    else:
        strata = [L(), duroid(), _polyring(pk), duroid(), L(width=0.1, units="inch"),
                  L(width=0.1, units="inch"), duroid(), _polyring(pk), duroid(), L()]
    layers, sheets, junc = _split(pk, strata)
    gbl = pk.gsm.choose_gblocks(layers, sheets, junc, _k0(2.0))
    pk.modes.choose_layer_modes(layers, sheets, junc, gbl, _k0(8.0), 25.0)
    assert [len(l.P) for l in layers] == expect


def test_choose_gblocks_golden(pk):
    """test/GSMs_test.jl:140-271: four Gblock partitions."""
    L = pk.Layer.make
    Gb = lambda r, j: (list(r), j)
    got = lambda g: [(list(b.rng), b.j) for b in g]
    rs = lambda: pk.rectstrip(units="cm", Px=1, Py=1, Lx=0.5, Ly=0.5, Nx=10, Ny=10)
    strata = [L(), rs(), L(width=1, units="mil", epsr=2.2), L(width=1, units="mil"),
              L(width=3, units="mm"), rs(), L(width=1, units="mil", epsr=2.2), L()]
    layers, sheets, junc = _split(pk, strata)
    assert got(pk.gsm.choose_gblocks(layers, sheets, junc, _k0(10.0))) == \
        [Gb(range(1, 4), 1), Gb(range(4, 6), 4)]
    duroid = lambda: L(width=20, units="mil", epsr=3.6, tandel=0.0013)
    strata = [L(), duroid(), _polyring(pk), duroid(), L(width=0.05, units="inch"), duroid(),
              _polyring(pk), duroid(), L()]
    layers, sheets, junc = _split(pk, strata)
    assert got(pk.gsm.choose_gblocks(layers, sheets, junc, _k0(2.0))) == \
        [Gb(range(1, 4), 2), Gb(range(4, 7), 5)]
    strata = [L(), duroid(), _polyring(pk), duroid(), duroid(), _polyring(pk), duroid(), L()]
    layers, sheets, junc = _split(pk, strata)
    assert got(pk.gsm.choose_gblocks(layers, sheets, junc, _k0(2.0))) == \
        [Gb(range(1, 3), 2), Gb(range(3, 6), 4)]
    strata = [L(), duroid(), _polyring(pk), duroid(), L(width=5, units="mil"), duroid(),
              _polyring(pk), duroid(), L(width=5, units="mil"), duroid(), _polyring(pk),
              duroid(), L()]
    layers, sheets, junc = _split(pk, strata)
    assert got(pk.gsm.choose_gblocks(layers, sheets, junc, _k0(2.0))) == \
        [Gb(range(1, 3), 2), Gb(range(3, 6), 5), Gb(range(6, 10), 8)]


# ------------------------------------------------- two independent TL computations
def _stacks(pk):
    L = pk.Layer.make
    sh = lambda c="J": pk.rectstrip(Lx=5, Ly=5, Px=10, Py=10, Nx=4, Ny=4, units="mm",
                                    sheet_class=c)
    lossy = lambda: L(width=0.5, units="mm", epsr=2.2, tandel=0.01)
    return [
        ([L(), sh(), L(width=1.0, units="mm", epsr=3.0), L(width=3.0, units="mm", epsr=1.1), L()],
         10.0, 0.0, 0.0),
        ([L(), lossy(), sh(), L(width=0.787, units="mm", epsr=2.2, tandel=0.0009), L()],
         17.0, 30.0, 45.0),
        ([L(), lossy(), sh("M"), L(width=0.787, units="mm", epsr=2.2, tandel=0.0009), L()],
         17.0, 30.0, 45.0),
        ([L(), sh("M"), L(width=2, units="mm", epsr=4.0), L()], 12.0, 40.0, 10.0),
        ([L(), L(width=2, units="mm", epsr=4.0), sh("M"), L()], 12.0, 40.0, 10.0),
        ([L(), L(width=2, units="mm", epsr=4.0, mur=1.5, mtandel=0.02), sh(), L()],
         22.0, 20.0, 0.0),
    ]


def _point(pk, tl, strata, fghz, th, ph):
    su = pk.analyze.prepare(strata, [fghz], {"theta": [th], "phi": [ph]})
    pt = list(su.points())[0]
    for l in su.layers:
        pk.modes.setup_modes(l, pt.k0, pt.beta00)
    ml = [tl.ModalLayer(l) for l in su.layers]
    for l in ml:
        l.setup(pt.k0, pt.beta00)
    return su, ml, pt


@pytest.mark.parametrize("icase", range(6))
def test_tl_oracle_matches_package(pk, icase):
    """Modal constants, Gblock partial GSMs / TLGFs / incident vectors and junction
    GSMs: the oracle's nodal solutions equal the package's recursions to 1e-12."""
    import tl_oracle as T
    strata, f, th, ph = _stacks(pk)[icase]
    su, ml, pt = _point(pk, T, strata, f, th, ph)
    for l, o in zip(su.layers, ml):
        if len(l.P):
            for a, b in ((l.beta, o.beta), (l.gamma, o.gamma), (l.Y, o.Y), (l.c, o.c),
                         (l.tvec, o.tvec)):
                assert _rel(b, a) < 1e-13
    for g in su.gbls:
        i1, i2 = g.rng[0], 1 + g.rng[-1]
        if g.j == 0:
            a = pk.gsm.gsm_slab_interface(su.layers[i1 - 1], su.layers[i2 - 1], pt.k0)
            b = T.slab_interface(ml[i1 - 1], ml[i2 - 1], pt.k0)
            for x, y in zip((a.s11, a.s12, a.s21, a.s22), b):
                assert _rel(y, x) < 1e-12
            continue
        s = g.j - i1 + 1
        sheet = su.sheets[su.junc[g.j - 1] - 1]
        region, mr = su.layers[i1 - 1:i2], ml[i1 - 1:i2]
        if sheet.sheet_class == "M":
            a, tl, vi = pk.gsm.gsm_magnetic_gblock(region, s, pt.k0)
            b, tl2, vi2 = T.gblock_magnetic(mr, s, pt.k0)
        else:
            a, tl, vi = pk.gsm.gsm_electric_gblock(region, s, pt.k0)
            b, tl2, vi2 = T.gblock_electric(mr, s, pt.k0)
        for x, y in zip((a.s11, a.s12, a.s21, a.s22, tl, vi), list(b) + [tl2, vi2]):
            assert _rel(y, x) < 1e-12


def test_star_product_matches_package(pk):
    import tl_oracle as T
    rng = np.random.default_rng(1)

    def rg(n1, n2):
        return [0.3 * (rng.standard_normal(s) + 1j * rng.standard_normal(s))
                for s in ((n1, n1), (n1, n2), (n2, n1), (n2, n2))]
    for n1, n2, n3 in ((1, 1, 1), (3, 5, 4), (10, 20, 15)):
        a, b = rg(n1, n2), rg(n2, n3)
        c = pk.cascade(pk.GSM(*a), pk.GSM(*b))
        for x, y in zip((c.s11, c.s12, c.s21, c.s22), T.star(a, b)):
            assert _rel(y, x) < 1e-12


# ------------------------------------------------------------------ closed forms
@pytest.mark.parametrize("th", [0.0, 35.0, 70.0])
@pytest.mark.parametrize("lossy", [False, True])
def test_dielectric_slab_fresnel(pk, th, lossy):
    """A bare slab (no sheet) at oblique incidence, TE and TM: the package's walk and
    the oracle's walk against r = (r12 + r23 P^2) / (1 + r12 r23 P^2),
    t = t12 t23 P / (1 + r12 r23 P^2), P = exp(-gamma d), with the modal (transverse
    field) impedances Z_TE = j w mu / gamma, Z_TM = gamma / (j w eps)."""
    import tl_oracle as T
    L = pk.Layer.make
    er, tand, d_mm = 4.3, (0.02 if lossy else 0.0), 1.7
    strata = [L(), L(width=d_mm, units="mm", epsr=er, tandel=tand), L()]
    f = 13.0
    su = pk.analyze.prepare(strata, [f], {"theta": [th], "phi": [20.0]})
    pt = list(su.points())[0]
    for l in su.layers:
        pk.modes.setup_modes(l, pt.k0, pt.beta00)
    gpk = su.cascade_point(pt.k0, None)
    ml = [T.ModalLayer(l) for l in su.layers]
    for l in ml:
        l.setup(pt.k0, pt.beta00)
    gor = T.walk(su, ml, pt.k0, None)
    k0 = pt.k0
    eta0 = T.ETA0
    kt2 = (k0 * math.sin(math.radians(th))) ** 2
    eps2 = er * complex(1, -tand)
    g1 = T.branch_sqrt(kt2 - k0 ** 2)
    g2 = T.branch_sqrt(kt2 - k0 ** 2 * eps2)
    P = np.exp(-g2 * d_mm * 1e-3)
    for mode in range(2):     # mode 0 TE, 1 TM (the layers' 00 pair)
        if su.layers[0].P[mode] == T.TE:
            Z1, Z2 = 1j * k0 * eta0 / g1, 1j * k0 * eta0 / g2
        else:
            Z1, Z2 = g1 / (1j * k0 / eta0), g2 / (1j * k0 * eps2 / eta0)
        r12, r23 = (Z2 - Z1) / (Z2 + Z1), (Z1 - Z2) / (Z1 + Z2)
        t12, t23 = 2 * Z2 / (Z1 + Z2), 2 * Z1 / (Z1 + Z2)
        r = (r12 + r23 * P * P) / (1 + r12 * r23 * P * P)
        t = t12 * t23 * P / (1 + r12 * r23 * P * P)
        for g in (gpk, pk.GSM(*gor)):
            assert abs(g.s11[mode, mode] - r) < 1e-12
            assert abs(g.s22[mode, mode] - r) < 1e-12
            assert abs(g.s21[mode, mode] - t) < 1e-12
            assert abs(g.s12[mode, mode] - t) < 1e-12
            assert abs(g.s21[1 - mode, mode]) < 1e-14


def _with_modes(T, layer, like, k0, b00):
    """An oracle copy of `layer` carrying the mode list and lattice of `like` (a
    Gblock's interior layers have no modes of their own)."""
    m = T.ModalLayer(layer)
    m.P, m.M, m.N = list(like.P), like.M.copy(), like.N.copy()
    m.beta1, m.beta2, m.key = like.beta1.copy(), like.beta2.copy(), dict(like.key)
    m.setup(k0, b00)
    return m


def _slab_cascade(T, part, k0):
    """Junctions and interior layers of `part` (oracle layers), first to last."""
    ref = T.new_gsm(len(part[0].P), len(part[0].P))
    for i in range(len(part) - 1):
        ref = T.star(ref, T.slab_interface(part[i], part[i + 1], k0))
        if i + 1 < len(part) - 1:
            ref = T.propagate(ref, part[i + 1])
    return ref


@pytest.mark.parametrize("icase", [0, 1, 5])
def test_vanishing_sheet_is_the_slab_stack(pk, icase):
    """The incident partial GSM of a J-sheet Gblock (the region with the sheet
    removed) equals the region's own junction-and-layer cascade, for the
    package's recursion and the oracle's nodal solution."""
    import tl_oracle as T
    strata, f, th, ph = _stacks(pk)[icase]
    su, ml, pt = _point(pk, T, strata, f, th, ph)
    for g in su.gbls:
        if g.j == 0:
            continue
        i1, i2 = g.rng[0], 1 + g.rng[-1]
        mr = ml[i1 - 1:i2]
        aux = [mr[0]] + [_with_modes(T, l, mr[-1], pt.k0, pt.beta00)
                         for l in su.layers[i1:i2 - 1]] + [mr[-1]]
        ref = _slab_cascade(T, aux, pt.k0)
        s = g.j - i1 + 1
        a, _, _ = pk.gsm.gsm_electric_gblock(su.layers[i1 - 1:i2], s, pt.k0)
        b, _, _ = T.gblock_electric(mr, s, pt.k0)
        for x, y, z in zip((a.s11, a.s12, a.s21, a.s22), b, ref):
            assert _rel(x, z) < 1e-12
            assert _rel(y, z) < 1e-12


@pytest.mark.parametrize("side", ["left", "right"])
def test_magnetic_gblock_is_a_shorted_stack(pk, side):
    """The incident partial GSM of an M-sheet Gblock is the region with a short
    circuit (PEC) at the sheet's interface: each side's reflection equals the
    cascade of that side's junctions and layers ending in a PEC."""
    import tl_oracle as T
    L = pk.Layer.make
    sh = pk.rectstrip(Lx=5, Ly=5, Px=10, Py=10, Nx=4, Ny=4, units="mm", sheet_class="M")
    strata = [L(), L(width=0.6, units="mm", epsr=2.5, tandel=0.01), sh,
              L(width=0.9, units="mm", epsr=3.3), L()]
    su, ml, pt = _point(pk, T, strata, 15.0, 25.0, 30.0)
    g = su.gbls[0]
    i1, i2 = g.rng[0], 1 + g.rng[-1]
    mr = ml[i1 - 1:i2]
    s = g.j - i1 + 1
    a, _, _ = pk.gsm.gsm_magnetic_gblock(su.layers[i1 - 1:i2], s, pt.k0)
    b, _, _ = T.gblock_magnetic(mr, s, pt.k0)
    raw = su.layers[i1 - 1:i2]
    if side == "left":      # first region, the layers up to interface s, then a PEC
        part = [mr[0]] + [_with_modes(T, l, mr[0], pt.k0, pt.beta00) for l in raw[1:s]]
        got = (a.s11, b[0])
    else:                   # last region back to interface s, then a PEC
        part = [mr[-1]] + [_with_modes(T, l, mr[-1], pt.k0, pt.beta00)
                           for l in raw[s:-1][::-1]]
        got = (a.s22, b[3])
    ref = T.new_gsm(len(part[0].P), len(part[0].P))
    for i in range(len(part) - 1):
        ref = T.star(ref, T.slab_interface(part[i], part[i + 1], pt.k0))
        ref = T.propagate(ref, part[i + 1])
    ref = T.star(ref, T.pec(len(part[0].P), len(part[0].P)))
    for x in got:
        assert _rel(x, ref[0]) < 1e-12


def test_resistive_sheet_closed_form_oracle(pk, orc):
    """The oracle's surface-resistance loading (src/FillZY.jl:221-231): a uniform
    R ohm/sq sheet filling the cell reflects -eta0/(eta0 + 2R) and transmits
    2R/(eta0 + 2R) at normal incidence."""
    import tl_oracle as T
    R = 120.0
    sh = pk.rectstrip(Lx=10, Ly=10, Px=10, Py=10, Nx=8, Ny=8, units="mm", Rsheet=R)
    res = orc.analyze(pk, [pk.Layer.make(), sh, pk.Layer.make()], [3.0, 9.0],
                      {"theta": [0.0], "phi": [0.0]})
    G = -T.ETA0 / (T.ETA0 + 2 * R)
    Tt = 2 * R / (T.ETA0 + 2 * R)
    for r in res:
        g = r["gsm"]
        assert np.abs(g.s11 - G * np.eye(2)).max() < 1e-4
        assert np.abs(g.s21 - Tt * np.eye(2)).max() < 1e-4


def test_power_conservation_dielectric_oblique(pk, orc):
    """A lossless stack (eps_r 2.5 slab, free space outside) carrying a PEC patch, at
    theta = 30 deg below the first grating lobe: each incident mode's scattered
    power sums to one (|S11|^2 + |S21|^2 over the propagating TE/TM pair), within
    the near-symmetry of centroid testing.  This pins the modal normalisation
    c = sqrt(1/(A Y)), the area factors and the Gblock TL Green's functions of the
    sheet solve in a dielectric at oblique incidence."""
    L = pk.Layer.make
    sh = pk.rectstrip(Lx=4, Ly=4, Px=8, Py=8, Nx=6, Ny=6, units="mm")
    strata = [L(), sh, L(width=1.2, units="mm", epsr=2.5), L()]
    res = orc.analyze(pk, strata, [9.0, 14.0], {"theta": [30.0], "phi": [0.0, 30.0]})
    assert len(res) == 4
    for r in res:
        g = r["gsm"]
        for c in range(2):
            p = np.sum(np.abs(g.s11[:, c]) ** 2) + np.sum(np.abs(g.s21[:, c]) ** 2)
            assert abs(p - 1) < 2e-3, p


@pytest.mark.parametrize("eps_end", [1.0, 2.2])
def test_principal_cutoff_rule(pk, eps_end):
    """The point-skip rule of src/PSSFSS.jl:252-255 (principal modes of both end media
    propagating), in the package's sweep and in the oracle's own, against its closed
    form: with unit-cell phasings (psi1, psi2) the principal mode's transverse
    wavenumber is |beta00| = |(psi1 beta1 + psi2 beta2) / 2 pi|, and it propagates in
    the end media (identical, src/PSSFSS.jl:619-633) iff |beta00| < k0 sqrt(eps_r mu_r)."""
    import tl_oracle as T
    from pssfss_amd import modes as Md
    P = 10.0
    sh = pk.sheets.rectstrip(Lx=5, Ly=5, Px=P, Py=P, Nx=3, Ny=3, units="mm")
    strata = [pk.Layer.make(epsr=eps_end), sh, pk.Layer.make(width=1.0, units="mm", epsr=3.0),
              pk.Layer.make(epsr=eps_end)]
    freqs = [2.0, 5.0, 9.0, 14.0, 21.0, 29.0]
    steer = {"psi1": [0.0, 45.0, 120.0, 170.0], "psi2": [0.0, 80.0]}
    su = pk.analyze.prepare(strata, freqs, steer)
    kept_pkg, kept_orc, expect = [], [], []
    for pt in su.points():
        for l in su.layers:
            Md.setup_modes(l, pt.k0, pt.beta00)
        kept_pkg.append(su.principal_ok())
        b = np.hypot(math.radians(pt.steer["psi1"]), math.radians(pt.steer["psi2"])) / (P * 1e-3)
        kmin = pt.k0 * math.sqrt(eps_end)
        assert abs(b - kmin) > 1e-3 * kmin          # no point on the boundary
        expect.append(b < kmin)
    ml = [T.ModalLayer(l) for l in su.layers]
    for fghz, k0, b00, st in T.points(su):
        for l in ml:
            l.setup(k0, b00)
        kept_orc.append(T.principal_propagates(ml[0], ml[-1]))
    assert kept_pkg == expect
    assert kept_orc == expect
    assert 0 < sum(expect) < len(expect)            # both outcomes occur
