"""Pins the oracle (FP64 CPU restatement) before it is trusted as the checker.

Every check comes from outside the oracle's own formulas: the reference's golden
numbers (test/*.jl), closed forms (Collin's grating, the 1-mode cascade), brute
force (long lattice sums, direct modal series, adaptive cubature of the singular
integrals) and invariants (independence of the smoothing parameter u).
"""
import math

import numpy as np
import pytest
from scipy import integrate

INCH = 0.0254


def _fillzy_case(pk, cls):
    FGHz = 2.0
    k0 = 2 * math.pi / (11.80285 / FGHz * INCH)
    L = pk.Layer.make
    layers = [L(), L(epsr=3.4, tandel=0.003, width=1, units="mil"),
              L(epsr=3.4, tandel=0.003, width=1, units="mil")]
    sh = pk.rectstrip(Nx=20, Ny=20, Lx=20, Ly=50, Px=100, Py=100, units="mil", sheet_class=cls)
    u = 0.5 * max(np.linalg.norm(sh.beta1), np.linalg.norm(sh.beta2)) * sh.upm
    return layers, sh, u, k0


@pytest.mark.parametrize("cls", ["J", "M"])
def test_fill_golden(pk, orc, cls):
    """test/FillZY_test.jl:19-80 golden entries of fillz / filly."""
    layers, sh, u, k0 = _fillzy_case(pk, cls)
    rw = pk.setup_rwg(sh)
    upm = sh.upm
    JK, rr, ri = orc.filljk(sh, rw, u / upm, 0.0, 0.0)
    b1m, b2m = sh.beta1 * upm, sh.beta2 * upm
    tabs = orc.modal_tables(cls == "M", k0, u, 0.0, 0.0, layers, 2, b1m, b2m, np.zeros(2), 1e-7)
    Z = orc.fill(sh, rw, JK, rr, ri, k0, u, layers, 2, tabs[:3], b1m, b2m, 0.0, 0.0)
    if cls == "J":
        assert abs(Z[19, 29] - (0.8990238712012797 - 354.1682770720877j)) < 0.005
        assert abs(Z[99, 99] - (110.95156005861463 - 37529.820828933574j)) < 0.01
        # far tighter than the reference's tolerance: the restatement follows it
        assert abs(Z[19, 29] - (0.8990238712012797 - 354.1682770720877j)) < 1e-9
    else:
        assert abs(Z[19, 29] - (1.2168396602563192e-6 - 0.02476973884018213j)) < 1e-4
        assert abs(Z[99, 99] - (2.2127855040521283e-6 - 3.5131557304457406j)) < 1e-4
        assert abs(Z[99, 99] - (2.2127855040521283e-6 - 3.5131557304457406j)) < 1e-12


# test/PGF_test.jl:19-85: 5-layer magneto-dielectric stack, oblique lattice.
def _pgf_case(pk):
    from pssfss_amd.sheets import s1s2_to_beta
    k0 = 2 * math.pi / (11.80285 / 2.0 * INCH)
    L = pk.Layer.make
    layers = [L(), L(width=20, units="mil", epsr=3.6, tandel=0.0013, mur=5.0),
              L(width=10, units="mil", epsr=8.0, tandel=0.0013, mur=2.0),
              L(width=50, units="mil", epsr=2.6, tandel=0.0003), L()]
    psi1, psi2 = 0.4, -0.6
    s1 = np.array([1.14 * INCH, 0.0])
    s2 = np.array([0.5707, 0.9885]) * INCH
    b1, b2 = s1s2_to_beta(s1, s2)
    b00 = (psi1 * b1 + psi2 * b2) / (2 * math.pi)
    rho = np.array([2, 6]) / 1000 * INCH
    return k0, layers, psi1, psi2, s1, s2, b1, b2, b00, rho


PGF_EXACT = {0.1: ((-232.15311689927114 - 10.286495939161526j, -11.778556545589899 - 160.87504453495308j),
                   (-5905.287097073966 - 116.24899221632135j, -65.290894751642 - 115.6027231717079j)),
             0.9: ((-86.47040978143367 - 9.887172408455307j, 133.51945406766538 - 160.47675058617088j),
                   (-2516.511258383936 - 111.36619289656556j, 138.66489521333057 - 115.0436702287188j))}


@pytest.mark.parametrize("ufact", [0.1, 0.9])
@pytest.mark.parametrize("magnetic", [0, 1])
def test_modal_series_golden(pk, orc, ufact, magnetic):
    k0, layers, psi1, psi2, s1, s2, b1, b2, b00, rho = _pgf_case(pk)
    u = ufact * max(np.linalg.norm(b1), np.linalg.norm(b2))
    M, t1, t2, conv = orc.modal_tables(magnetic, k0, u, psi1, psi2, layers, 2, b1, b2, b00, 5e-12)
    assert conv
    v1 = orc.sigma(t1, M, b1, b2, psi1, psi2, rho)
    v2 = orc.sigma(t2, M, b1, b2, psi1, psi2, rho)
    e1, e2 = PGF_EXACT[ufact][magnetic]
    assert abs(v1 - e1) / abs(e1) < 2e-4
    assert abs(v2 - e2) / abs(e2) < 2e-4


def test_green_functions_u_independent(pk, orc):
    """test/PGF_test.jl:88-126: spectral + spatial parts agree for u = 0.1 and 0.9."""
    k0, layers, psi1, psi2, s1, s2, b1, b2, b00, rho = _pgf_case(pk)
    s = 2
    ms, ms1 = layers[s - 1].mur, layers[s].mur
    es, es1 = layers[s - 1].epsr, layers[s].epsr
    mut = 2 * ms * ms1 / (ms + ms1)
    epsb = (es + es1) / 2
    G = []
    for uf in (0.1, 0.9):
        u = uf * max(np.linalg.norm(b1), np.linalg.norm(b2))
        js, ks, _ = orc.jksums(u * rho, psi1, psi2, u * s1, u * s2, True)
        Me, t1, t2, _ = orc.modal_tables(0, k0, u, psi1, psi2, layers, s, b1, b2, b00, 5e-12)
        sm1 = orc.sigma(t1, Me, b1, b2, psi1, psi2, rho)
        sm2 = orc.sigma(t2, Me, b1, b2, psi1, psi2, rho)
        Mm, p1, p2, _ = orc.modal_tables(1, k0, u, psi1, psi2, layers, s, b1, b2, b00, 5e-12)
        sp1 = orc.sigma(p1, Mm, b1, b2, psi1, psi2, rho)
        sp2 = orc.sigma(p2, Mm, b1, b2, psi1, psi2, rho)
        c3, d3 = orc.c3d3(k0, u, ms, es, ms1, es1)
        c3s, _ = orc.c3d3(k0, u, es, ms, es, ms)
        c3p, _ = orc.c3d3(k0, u, es1, ms1, es1, ms1)
        _, d3s = orc.c3d3(k0, u, es, ms, es, ms)
        _, d3p = orc.c3d3(k0, u, es1, ms1, es1, ms1)
        GA = mut * (sm1 + u / (4 * math.pi) * (js + c3 / u ** 2 * ks))
        GPhi = 1 / epsb * (sm2 + u / (4 * math.pi) * (js + d3 / u ** 2 * ks))
        GF = -(sp1 + u * epsb / math.pi * js + (c3s * es + c3p * es1) / (2 * math.pi * u) * ks)
        GPsi = sp2 + u / (math.pi * mut) * js + (d3s / ms + d3p / ms1) / (2 * math.pi * u) * ks
        G.append((GA, GPhi, GF, GPsi))
    for a, b in zip(G[0], G[1]):
        assert 2 * abs(b - a) / abs(b + a) < 2e-4


def test_jksums_brute_force(orc):
    """Converged spatial sums equal a 60-ring brute-force lattice sum to ~1e-8."""
    urho = np.array([0.37, -0.81])
    us1 = np.array([math.pi, 0.0])
    us2 = np.array([0.9, 2.7])
    psi1, psi2 = 0.7, -1.3
    js, ks, rings = orc.jksums(urho, psi1, psi2, us1, us2, True)
    arg = np.linalg.norm(urho)
    J = (math.exp(-arg) - 1) / arg
    K = math.exp(-arg) + 0j
    R = 60
    m, n = np.meshgrid(np.arange(-R, R + 1), np.arange(-R, R + 1), indexing="ij")
    mask = (m != 0) | (n != 0)
    d = urho[None, None, :] - (m[..., None] * us1 + n[..., None] * us2)
    a = np.linalg.norm(d, axis=-1)[mask]
    ph = (m * psi1 + n * psi2)[mask]
    term = np.exp(-a) * np.exp(-1j * ph)
    J = J + np.sum(term / a)
    K = K + np.sum(term)
    assert 5 < rings < 65
    assert abs(js - J) / abs(J) < 1e-7
    assert abs(ks - K) / abs(K) < 1e-7


def test_modal_series_direct(pk, orc):
    """Table interpolation vs a directly summed modal series (PGF.jl:665-760)."""
    from pssfss_amd.cmath_ref import mysqrt
    k0, layers, psi1, psi2, s1, s2, b1, b2, b00, rho = _pgf_case(pk)
    u = 0.9 * max(np.linalg.norm(b1), np.linalg.norm(b2))
    M, t1, t2, _ = orc.modal_tables(0, k0, u, psi1, psi2, layers, 2, b1, b2, b00, 5e-12)
    v1 = orc.sigma(t1, M, b1, b2, psi1, psi2, rho)
    # direct spectral sum of the Kummer-subtracted summand, with the Floquet phase
    s = 2
    k0sq = k0 * k0
    ms, ms1 = layers[s - 1].mur, layers[s].mur
    es, es1 = layers[s - 1].epsr, layers[s].epsr
    mut = 2 * ms * ms1 / (ms + ms1)
    c3, _ = orc.c3d3(k0, u, ms, es, ms1, es1)
    area = 4 * math.pi ** 2 / abs(b1[0] * b2[1] - b1[1] * b2[0])
    R = 700
    m, n = np.meshgrid(np.arange(-R, R + 1), np.arange(-R, R + 1), indexing="ij")
    bv = b00[None, None, :] + m[..., None] * b1 + n[..., None] * b2
    bsq = np.maximum(np.sum(bv * bv, axis=-1), 1e-10 * k0sq)
    kap = np.sqrt(bsq + u * u)

    def msq(z):
        r = np.sqrt(z.astype(complex))
        flip = (r.real == 0) & (r.imag < 0)
        r[flip] = -r[flip]
        return r
    g = msq(bsq - k0sq * layers[0].epsr * layers[0].mur)
    zl = layers[0].mur / g
    for i in range(1, s):
        g = msq(bsq - k0sq * layers[i].epsr * layers[i].mur)
        z0 = layers[i].mur / g
        th = np.tanh(layers[i].width * g)
        zl = z0 * (zl + z0 * th) / (z0 + zl * th)
    g = msq(bsq - k0sq * layers[-1].epsr * layers[-1].mur)
    zr = layers[-1].mur / g
    for i in range(len(layers) - 2, s - 1, -1):
        g = msq(bsq - k0sq * layers[i].epsr * layers[i].mur)
        z0 = layers[i].mur / g
        th = np.tanh(layers[i].width * g)
        zr = z0 * (zr + z0 * th) / (z0 + zr * th)
    vte = 2 * (zl * zr / (zl + zr)) / mut
    summ = vte - (1 + c3 / kap ** 2) / kap
    ph = np.exp(-1j * (bv @ rho))
    direct = np.sum(summ * ph) / (2 * area)
    assert abs(v1 - direct) / abs(direct) < 2e-4


def test_singular_integrals_cubature(pk, orc):
    """Analytic rinv and rho_r (src/Zint.jl:126-165) vs adaptive cubature
    (test/Zint_test.jl:53-102), including a self pair split at the centroid."""
    sh = pk.annular_ring(a=0.4957, b=0.5557, P=1.1414, nr=2, nphi=16, units="inch")
    rw = pk.setup_rwg(sh)
    ul = 0.4 * max(np.linalg.norm(sh.beta1), np.linalg.norm(sh.beta2))
    JK, rr, ri = orc.filljk(sh, rw, ul, 0.0, 0.0)
    nface = sh.nface
    for fp in [0, 1, 5, nface + 3, 7 * nface + 2, nface * nface - 1]:
        ifm, ifs = fp % nface, fp // nface
        rs = sh.rho[sh.fv[:, ifs]]
        rm = sh.rho[sh.fv[:, ifm]].mean(axis=0)

        def integ(tri):
            a = tri[0]
            e1 = tri[1] - tri[0]
            e2 = tri[2] - tri[0]
            area2 = abs(e1[0] * e2[1] - e1[1] * e2[0])
            out = []
            for comp in range(3):
                def f(y, x):
                    r = a + x * e1 + y * e2 - rm
                    d = math.hypot(*r)
                    return 1 / d if comp == 0 else r[comp - 1] / d
                v, _ = integrate.dblquad(f, 0, 1, 0, lambda x: 1 - x, epsabs=0, epsrel=1e-11)
                out.append(v * area2)
            return np.array(out), area2
        if ifm == ifs:
            tot = np.zeros(3)
            for tri in ([rm, rs[0], rs[1]], [rm, rs[1], rs[2]], [rm, rs[2], rs[0]]):
                v, _ = integ(np.array(tri))
                tot += v
            e1, e2 = rs[1] - rs[0], rs[2] - rs[0]
            area2 = abs(e1[0] * e2[1] - e1[1] * e2[0])
        else:
            tot, area2 = integ(rs)
        rinv_ref = tot[0] / area2
        rr_ref = tot[1:] / area2
        assert abs(ul * ri[fp] - rinv_ref) / abs(rinv_ref) < 1e-7
        assert np.linalg.norm(rr[fp] - rr_ref) / np.linalg.norm(rr_ref) < 1e-7


def test_rwgbfft_golden(pk, orc):
    """test/RWGData_test.jl:125-145."""
    sh = pk.rectstrip(Lx=1, Ly=1, Px=1, Py=1, Nx=2, Ny=1, units="mm", fufp=False)
    rw = pk.setup_rwg(sh)
    ft = orc.rwgbfft(sh, rw, np.zeros(2), 0.0, 0.0)
    ref0 = np.array([[1.666666666666667e-4, -3.333333333333334e-4],
                     [3.333333333333334e-4, 3.333333333333334e-4],
                     [1.6666666666666663e-4, -3.333333333333334e-4]])
    assert np.allclose(ft[:3], ref0, rtol=1.5e-8, atol=0)
    sh = pk.rectstrip(Lx=1, Ly=1, Px=1, Py=1, Nx=2, Ny=1, units="inch", fufp=False)
    rw = pk.setup_rwg(sh)
    ft = orc.rwgbfft(sh, rw, np.array([20.0, -30.3]), 0.8, -0.7)
    ref = np.array([[0.003986416685586114 - 0.0010511300702857503j,
                     -0.008079847138694855 + 0.0021304773084817956j],
                    [0.008263557134402535 - 0.0010871639073400372j,
                     0.008284527320353664 - 0.0010899227712197804j],
                    [0.004122638196063312 - 1.5707327530113332e-5j,
                     -0.008355946971820058 + 3.1836312009155855e-5j]])
    for i in range(3):
        assert np.linalg.norm(ft[i] - ref[i]) <= 1.5e-8 * np.linalg.norm(ref[i])


def test_fft_against_numpy(orc):
    rng = np.random.default_rng(0)
    for M in (2, 6, 30, 80, 96, 150):
        a = rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M))
        buf = np.asfortranarray(a).ravel(order="F").copy().view(np.float64)
        orc.lib().orc_fft2(buf, M)
        got = buf.view(np.complex128).reshape((M, M), order="F")
        assert np.abs(got - np.fft.fft2(a)).max() / np.abs(a).sum() * M < 1e-12


def _grating(kP, nterms=30):
    sum1 = 1.3862943611198906
    sum3 = 7.2123414189575710
    x = kP / (4 * math.pi)
    th = x * sum1 + x ** 3 / 6 * sum3
    for n in range(1, nterms + 1):
        a = x / (n - 0.5)
        b = x / n
        th += math.asin(a) - (a + a ** 3 / 6) - (math.asin(b) - (b + b ** 3 / 6))
    G = math.sin(th) * complex(math.cos(-math.pi / 2 - th), math.sin(-math.pi / 2 - th))
    return G, 1 + G


@pytest.mark.parametrize("cls", ["J", "M"])
def test_grating_closed_form(pk, orc, cls):
    """test/full_test.jl:27-93: Collin's strip grating, J and M formulations."""
    FGHz = 0.8
    r, t = _grating(2 * math.pi * FGHz)
    period = 299792458.0 * 100 / 2.54 * 1e-9
    Py, Ly = period, period / 2
    Px = Lx = Ly / 10
    sh = pk.rectstrip(Px=Px, Py=Py, Lx=Lx, Ly=Ly, Nx=round(40 * Lx / Ly), Ny=40, units="inch",
                      sheet_class=cls)
    res = orc.analyze(pk, [pk.Layer.make(), sh, pk.Layer.make()], [FGHz],
                      {"theta": [0.0], "phi": [0.0]})
    g = res[0]["gsm"]
    ex = {"s11": np.array([[r, 0], [0, -t]]), "s22": np.array([[r, 0], [0, -t]]),
          "s21": np.array([[t, 0], [0, -r]]), "s12": np.array([[t, 0], [0, -r]])}
    for k, v in ex.items():
        assert np.abs(getattr(g, k) - v).max() < 0.01


def test_power_conservation_lossless(pk, orc):
    """Lossless PEC patch below the first grating lobe: |S11|^2 + |S21|^2 = 1
    within the near-symmetry of centroid testing."""
    strata, freqs, steer = pk.workloads.patch_free_space(n=8, nfreq=3)
    res = orc.analyze(pk, strata, [5.0, 12.0, 20.0], steer)
    for r in res:
        g = r["gsm"]
        for c in range(2):
            p = np.sum(np.abs(g.s11[:, c]) ** 2) + np.sum(np.abs(g.s21[:, c]) ** 2)
            assert abs(p - 1) < 1e-3


def test_frozen_decisions_make_fd_smooth(pk, orc):
    """With its discrete choices recorded at a base geometry and replayed, the
    oracle is smooth in the node positions: central differences at 1e-6 and 1e-4 of
    the period agree to 1e-7 relative (unfrozen differences at 1e-6 carry the
    ~1e-9 jumps of cell crossings, ~1e-3 relative)."""
    from pssfss_amd.adjoint import bilinear_patch
    P = 12.0
    corners = np.array([[2.1, 2.4], [9.7, 2.0], [10.1, 9.6], [2.3, 9.9]])
    L = pk.Layer.make

    def strata(c):
        sh, _ = bilinear_patch(c, 5, 5, P, P)
        return [L(), sh, L(width=1.0, units="mm", epsr=3.0), L(width=3.0, units="mm", epsr=1.1),
                L()]

    def f(res):
        g = res[0]["gsm"]
        return np.array([abs(g.s21[0, 0]) ** 2, abs(g.s21[1, 1]) ** 2])
    freqs, steer = [10.0], {"theta": [0.0], "phi": [0.0]}
    nufp = pk.analyze.prepare(strata(corners), freqs, steer).rwgdat[0].nufp
    fz = orc.FrozenDecisions(nufp, 1)
    orc.analyze(pk, strata(corners), freqs, steer, decisions=(fz, "record"))
    for k in (0, 5):
        d = []
        for h in (1e-6 * P, 1e-4 * P):
            c1, c2 = corners.copy().reshape(-1), corners.copy().reshape(-1)
            c1[k] += h
            c2[k] -= h
            d.append((f(orc.analyze(pk, strata(c1.reshape(4, 2)), freqs, steer,
                                    decisions=(fz, "replay")))
                      - f(orc.analyze(pk, strata(c2.reshape(4, 2)), freqs, steer,
                                      decisions=(fz, "replay")))) / (2 * h))
        assert np.abs(d[0] - d[1]).max() < 1e-7 * np.abs(d[0]).max()
