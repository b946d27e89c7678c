"""Host-side reverse mode of the adjoint (CPU): the star-product and whole-walk
cotangents against directional finite differences, and the meander amplitude
tangent against the mesher."""
import numpy as np


def _rg(rng, n1, n2, s=0.3):
    f = lambda r, c: s * (rng.standard_normal((r, c)) + 1j * rng.standard_normal((r, c)))
    return f(n1, n1), f(n1, n2), f(n2, n1), f(n2, n2)


def _pair(a, b):
    return sum(np.sum(x * y) for x, y in zip(a, b))


def test_star_vjp_matches_directional_derivative(pk):
    from pssfss_amd import adjoint as A
    G = pk.gsm
    rng = np.random.default_rng(7)
    a, b = G.GSM(*_rg(rng, 3, 5)), G.GSM(*_rg(rng, 5, 4))
    da, db = G.GSM(*_rg(rng, 3, 5)), G.GSM(*_rg(rng, 5, 4))
    cb = G.GSM(*_rg(rng, 3, 4))
    ab, bb = A._star_vjp(a, b, cb)
    lin = _pair(A._blocks(ab), A._blocks(da)) + _pair(A._blocks(bb), A._blocks(db))
    eps = 1e-6

    def L(t):
        aa = G.GSM(*[x + t * y for x, y in zip(A._blocks(a), A._blocks(da))])
        bbb = G.GSM(*[x + t * y for x, y in zip(A._blocks(b), A._blocks(db))])
        return _pair(A._blocks(cb), A._blocks(G.cascade(aa, bbb)))
    fd = (L(eps) - L(-eps)) / (2 * eps)
    assert abs(fd - lin) < 1e-8 * abs(lin)


def test_cascade_vjp_walk(pk):
    """Forward value of the reverse-mode walk equals Setup.cascade_point, and its
    block cotangents give the directional derivative of <W, S_total> (meander
    polarizer stack: four sheet Gblocks, 30-mode foam layers, theta 40 deg)."""
    from pssfss_amd import adjoint as A
    G = pk.gsm
    strata, freqs, steer = pk.workloads.meander_polarizer(ntri=60, nfreq=1, thetas=(40.0,))
    su = pk.analyze.prepare(strata, [11.0], steer)
    pt = list(su.points())[0]
    for l in su.layers:
        pk.modes.setup_modes(l, pt.k0, pt.beta00)
    rng = np.random.default_rng(3)
    blocks = []
    for g in su.gbls:
        i1, i2 = g.rng[0], 1 + g.rng[-1]
        n1, n2 = len(su.layers[i1 - 1].P), len(su.layers[i2 - 1].P)
        if g.j == 0:
            blocks.append(G.gsm_slab_interface(su.layers[i1 - 1], su.layers[i2 - 1], pt.k0))
        else:
            blocks.append(G.GSM(*_rg(rng, n1, n2, 0.2)))
    it = iter([b for b, g in zip(blocks, su.gbls) if g.j != 0])
    ref = su.cascade_point(pt.k0, lambda *a: next(it))
    n1, n2 = len(su.layers[0].P), len(su.layers[-1].P)
    W = G.GSM(*_rg(rng, n1, n2))
    total, bars = A.cascade_vjp(su, pt.k0, blocks, W)
    for x, y in zip(A._blocks(total), A._blocks(ref)):
        assert np.abs(x - y).max() < 1e-12
    ks = [k for k, g in enumerate(su.gbls) if g.j != 0]
    dirs = {k: G.GSM(*[rng.standard_normal(m.shape) + 1j * rng.standard_normal(m.shape)
                       for m in A._blocks(blocks[k])]) for k in ks}
    lin = sum(_pair(A._blocks(bars[k]), A._blocks(dirs[k])) for k in ks)
    eps = 1e-6

    def L(t):
        bl = [G.GSM(*[x + t * y for x, y in zip(A._blocks(b), A._blocks(dirs[k]))])
              if k in dirs else b for k, b in enumerate(blocks)]
        tot, _ = A.cascade_vjp(su, pt.k0, bl, W)
        return _pair(A._blocks(W), A._blocks(tot))
    fd = (L(eps) - L(-eps)) / (2 * eps)
    assert abs(fd - lin) < 1e-7 * abs(lin)


def test_meander_amplitude_tangent(pk):
    from pssfss_amd import adjoint as A
    kw = dict(a=6.0, b=10.0, h=2.4, w1=0.3, w2=0.3, ntri=400, units="mm", fufp=False)
    sh, vt = A.meander_amplitude_tangent(kw)
    d = 1e-3
    sp = pk.sheets.meander(**dict(kw, h=2.4 + d))
    assert np.array_equal(sp.fv, sh.fv)
    assert np.abs((sp.rho - sh.rho) / d - vt[0]).max() < 1e-9
