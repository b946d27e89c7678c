"""Host-side logic (mesh, RWG maps, rings, GSM algebra) against the reference's
exact fixtures, and the C-ABI library's exports (no GPU work)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from conftest import gpu_available


def test_rings(pk):
    """test/Rings_test.jl."""
    from pssfss_amd.rings import ring
    assert ring(0) == [(0, 0)]
    assert ring(1) == [(1, -1), (1, 0), (1, 1), (-1, -1), (-1, 0), (-1, 1), (0, -1), (0, 1)]
    assert ring(2) == [(2, -2), (2, -1), (2, 0), (2, 1), (2, 2), (-2, -2), (-2, -1), (-2, 0),
                       (-2, 1), (-2, 2), (-1, -2), (0, -2), (1, -2), (-1, 2), (0, 2), (1, 2)]
    assert len(ring(10)) == 80
    assert all(max(abs(m), abs(n)) == 100 for m, n in ring(100))


@pytest.mark.parametrize("fufp", [True, False])
def test_rwgdata_maps(pk, fufp):
    """test/RWGData_test.jl:7-37 (exact maps of a 4-face periodic strip)."""
    sh = pk.rectstrip(Lx=1, Ly=1, Px=1, Py=1, Nx=2, Ny=1, units="mm", fufp=fufp)
    d = pk.setup_rwg(sh)
    assert (d.bfe + 1).tolist() == [[8, 6, 9, 1, 2, 5], [8, 6, 9, 3, 4, 7]]
    assert (d.bff + 1).tolist() == [[1, 2, 3, 2, 4, 1], [2, 3, 4, 1, 3, 4]]
    assert (d.ebf + 1).tolist() == [4, 5, 4, 5, 6, 2, 6, 1, 3]
    assert d.eci.tolist() == [3, 3, 4, 4, 1, 0, 2, 0, 0]
    if fufp:
        assert (d.ufpm + 1).tolist() == [[1, 5, 9, 11], [2, 6, 10, 12], [3, 7, 1, 5],
                                         [4, 8, 2, 6]]
        classes = [(d.ufp_members[d.ufp_offsets[i]:d.ufp_offsets[i + 1]] + 1).tolist()
                   for i in range(d.nufp)]
        assert classes == [[1, 11], [2, 12], [3], [4], [5, 15], [6, 16], [7], [8], [9], [10],
                           [13], [14]]
        assert d.nufp == 12
    else:
        assert d.ufp_identity and d.nufp == 16


def test_recttri_counts(pk):
    """src/Sheets.jl:398-480 counts; fe[i] is opposite fv[i]."""
    sh = pk.recttri([0, 0], [3, 2], 3, 2)
    assert sh.rho.shape[0] == 12 and sh.nedge == 3 * 6 + 5 and sh.nface == 12
    for f in range(sh.nface):
        for i in range(3):
            e = sh.fe[i, f]
            assert sh.fv[i, f] not in (sh.e1[e], sh.e2[e])


def test_annulus_mesh(pk):
    sh = pk.annular_ring(a=3.0, b=3.6, P=9.0, nr=6, nphi=168, units="mm")
    d = pk.setup_rwg(sh)
    assert sh.nface == 2016 and d.nbf == (3 * 6 - 1) * 168
    assert d.ufp_identity and d.nufp == sh.nface ** 2


def test_structured_ufp_count(pk):
    """Translation classes of the 14x14 patch mesh: far fewer than nface^2."""
    strata, _, _ = pk.workloads.patch_free_space()
    d = pk.setup_rwg(strata[1])
    assert strata[1].nface == 392 and d.nbf == 560
    assert d.nufp < 392 ** 2 // 20
    # every face pair belongs to exactly one class
    assert d.ufp_offsets[-1] == 392 ** 2
    assert np.array_equal(np.sort(d.ufp_members), np.arange(392 ** 2))


def test_cascade_identity_and_one_mode(pk):
    """test/GSMs_test.jl:17-59: identity cascade and the 1-mode closed form."""
    a = pk.GSM.new(10, 20)
    b = pk.GSM.new(20, 15)
    c = pk.cascade(a, b)
    assert np.linalg.norm(c.s11) == 0 and np.linalg.norm(c.s22) == 0
    assert np.array_equal(c.s12, np.eye(10, 15)) and np.array_equal(c.s21, np.eye(15, 10))
    al, ga = 0.35, 0.23
    be, de = math.sqrt(1 - al ** 2), math.sqrt(1 - ga ** 2)
    a = pk.GSM.new(1, 1)
    b = pk.GSM.new(1, 1)
    b.s11[:] = b.s22[:] = ga
    b.s12[:] = b.s21[:] = de
    c = pk.cascade(a, b)
    D = 1 / (1 - a.s22[0, 0] * b.s11[0, 0])
    assert abs(c.s11[0, 0] - (a.s11[0, 0] + a.s12[0, 0] * a.s21[0, 0] * b.s11[0, 0] / D)) <= 1e-16
    assert abs(c.s21[0, 0] - a.s21[0, 0] * b.s21[0, 0] / D) <= 1e-16
    assert abs(c.s22[0, 0] - (b.s22[0, 0] + b.s12[0, 0] * b.s21[0, 0] * a.s22[0, 0] / D)) <= 1e-16


def test_cascade_layer(pk):
    """test/GSMs_test.jl:62-94: propagation through a slab."""
    a = pk.GSM.new(1, 1)
    a.s11[:] = a.s22[:] = 0.35
    a.s12[:] = a.s21[:] = math.sqrt(1 - 0.35 ** 2)
    s22 = a.s22.copy()
    lay = pk.Layer.make(width=25, units="mm")
    lay.gamma = np.array([1j])
    p = np.exp(-1j * 0.025)
    s12 = a.s12 * p
    pk.cascade_layer(a, lay)
    assert abs(a.s12[0, 0] - s12[0, 0]) <= 1e-16
    assert abs(a.s22[0, 0] - s22[0, 0] * p * p) <= 1e-15


def test_gblocks_and_modes(pk):
    """Thin slab joins the sheet's Gblock; end regions carry the TE/TM pair."""
    strata, freqs, steer = pk.workloads.ring_on_slab(nr=2, nphi=24, nfreq=3)
    su = pk.analyze.prepare(strata, freqs, steer)
    assert [(list(g.rng), g.j) for g in su.gbls] == [([1, 2], 1)]
    assert len(su.layers[0].P) == 2 and len(su.layers[-1].P) == 2
    pts = list(su.points())
    assert len(pts) == 6 and pts[0].steer == {"theta": 30.0, "phi": 0.0}


def test_library_exports(pk):
    """libpssfss_hip.so loads and exports every entry point declared in the header."""
    from pssfss_amd import hip
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(root, "include", "pssfss_hip.h")).read()
    declared = set(re.findall(r"\b(pssfss_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(hip.EXPORTS)
    L = ctypes.CDLL(hip.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name


def test_struct_layout_matches_header(pk):
    from pssfss_amd import hip
    assert ctypes.sizeof(hip.Point) == 184
    assert ctypes.sizeof(hip.SheetDesc) == 192
    assert ctypes.sizeof(hip.PointInfo) == 32
    assert ctypes.sizeof(hip.Stats) == 192


def test_recttri_and_combine_goldens(pk):
    """test/RWGSheet_test.jl:5-80: exact nodes, edges and face maps (1-based there)."""
    from pssfss_amd.sheets import recttri, combine
    sh1 = recttri([0., 0.], [1., 1.], 1, 1)
    assert sh1.rho.tolist() == [[0, 0], [1, 0], [0, 1], [1, 1]]
    assert (sh1.e1 + 1).tolist() == [1, 3, 1, 2, 1] and (sh1.e2 + 1).tolist() == [2, 4, 3, 4, 4]
    assert (sh1.fe + 1).T.tolist() == [[2, 3, 5], [4, 5, 1]]
    assert (sh1.fv + 1).T.tolist() == [[1, 4, 3], [1, 2, 4]]
    cases = [
        ([1., 0.], [2., 1.], "x", 1.0, [[0, 0], [1, 0], [0, 1], [1, 1], [2, 0], [2, 1]],
         [1, 3, 1, 2, 1, 2, 4, 5, 2], [2, 4, 3, 4, 4, 5, 6, 6, 6],
         [[1, 4, 3], [1, 2, 4], [2, 6, 4], [2, 5, 6]], [[2, 3, 5], [4, 5, 1], [7, 4, 9], [8, 9, 6]]),
        ([0., 1.], [1., 2.], "y", 1.0, [[0, 0], [1, 0], [0, 1], [1, 1], [0, 2], [1, 2]],
         [1, 3, 1, 2, 1, 5, 3, 4, 3], [2, 4, 3, 4, 4, 6, 5, 6, 6],
         [[1, 4, 3], [1, 2, 4], [3, 6, 5], [3, 4, 6]], [[2, 3, 5], [4, 5, 1], [6, 7, 9], [8, 9, 2]]),
        ([1., 1.], [2., 2.], "y", 1.0, [[0, 0], [1, 0], [0, 1], [1, 1], [2, 1], [1, 2], [2, 2]],
         [1, 3, 1, 2, 1, 4, 6, 4, 5, 4], [2, 4, 3, 4, 4, 5, 7, 6, 7, 7],
         [[1, 4, 3], [1, 2, 4], [4, 7, 6], [4, 5, 7]],
         [[2, 3, 5], [4, 5, 1], [7, 8, 10], [9, 10, 6]]),
        ([2., 0.], [3., 1.], " ", float("nan"),
         [[0, 0], [1, 0], [0, 1], [1, 1], [2, 0], [3, 0], [2, 1], [3, 1]],
         [1, 3, 1, 2, 1, 5, 7, 5, 6, 5], [2, 4, 3, 4, 4, 6, 8, 7, 8, 8],
         [[1, 4, 3], [1, 2, 4], [5, 8, 7], [5, 6, 8]],
         [[2, 3, 5], [4, 5, 1], [7, 8, 10], [9, 10, 6]]),
    ]
    for bl, tr, dc, dv, rho, e1, e2, fv, fe in cases:
        sh3 = combine(sh1, recttri(bl, tr, 1, 1), dc, dv)
        assert sh3.rho.tolist() == rho
        assert (sh3.e1 + 1).tolist() == e1 and (sh3.e2 + 1).tolist() == e2
        assert (sh3.fv + 1).T.tolist() == fv and (sh3.fe + 1).T.tolist() == fe


@pytest.mark.parametrize("ntri,h", [(200, 2.0), (800, 2.6), (8000, 3.0)])
def test_meander_mesh(pk, ntri, h):
    """Conforming single-piece mesh of exactly the meander's area
    a w2 + 2 w1 (h - w2) (ten rectangles of src/Elements.jl:617-729), reaching
    both x = 0 and x = a so its basis functions wrap across the cell."""
    a, b, w1, w2 = 6.0, 10.0, 0.3, 0.3
    sh = pk.sheets.meander(a=a, b=b, h=h, w1=w1, w2=w2, ntri=ntri, units="mm")
    v = sh.rho[sh.fv]                       # (3, nface, 2)
    area = 0.5 * np.abs((v[1, :, 0] - v[0, :, 0]) * (v[2, :, 1] - v[0, :, 1]) -
                        (v[2, :, 0] - v[0, :, 0]) * (v[1, :, 1] - v[0, :, 1]))
    assert abs(area.sum() - (a * w2 + 2 * w1 * (h - w2))) < 1e-9
    assert np.all(area > 0)
    uses = np.bincount(sh.fe.reshape(-1), minlength=sh.nedge)
    assert set(uses.tolist()) <= {1, 2}
    assert sh.rho.shape[0] - sh.nedge + sh.nface == 1          # one simply connected piece
    assert 0.5 * ntri < sh.nface < 1.5 * ntri
    assert abs(sh.rho[:, 0].min()) < 1e-12 and abs(sh.rho[:, 0].max() - a) < 1e-12
    if ntri <= 800:
        d = pk.setup_rwg(sh)
        assert np.count_nonzero(d.eci > 0) > 0 and d.nbf > 0


@pytest.mark.parametrize("kind", ["loaded", "loaded_solid", "jerusalem", "jerusalem_loaded"])
def test_cross_meshes_area_and_conformity(pk, kind):
    """Structured cross meshes: the triangle areas add up to the closed-form metal
    area, every triangle is counter-clockwise, and every edge is shared by at most
    two faces (conforming mesh)."""
    S = pk.sheets
    if kind == "loaded":
        sh = S.loadedcross(s1=[10, 0], s2=[0, 10], L1=8, L2=2, w=0.5, ntri=600, units="mm")
        area = (2 * 8 * 2 - 4) - (2 * 7 * 1 - 1)
    elif kind == "loaded_solid":
        sh = S.loadedcross(s1=[10, 0], s2=[0, 10], L1=8, L2=2, w=1.5, ntri=300, units="mm")
        area = 2 * 8 * 2 - 4
    elif kind == "jerusalem":
        sh = S.jerusalemcross(P=12, L1=10, L2=0.6, A=5, B=0.6, w=0.6, ntri=800, units="mm")
        area = (2 * 10 * 0.6 - 0.36) + 4 * 5 * 0.6 - 4 * 0.36
    else:
        sh = S.jerusalemcross(P=12, L1=10, L2=2, A=5, B=2, w=0.5, ntri=1500, units="mm")
        area = ((2 * 10 * 2 - 4) + 4 * 10 - 16) - ((2 * 9 - 1) + 4 * 4 - 4)
    p = sh.rho[sh.fv]
    a = 0.5 * ((p[1, :, 0] - p[0, :, 0]) * (p[2, :, 1] - p[0, :, 1])
               - (p[2, :, 0] - p[0, :, 0]) * (p[1, :, 1] - p[0, :, 1]))
    assert (a > 0).all()
    assert abs(a.sum() - area) < 1e-9 * area
    cnt = np.bincount(sh.fe.reshape(-1), minlength=sh.nedge)
    assert cnt.max() <= 2 and cnt.min() >= 1
    rw = pk.rwg.setup_rwg(sh)
    assert rw.nbf == int((cnt == 2).sum())


@pytest.mark.parametrize("a,b,sides,ntri", [([0.0], [3.0], 4, 200), ([1.0], [2.0], 6, 500),
                                             ([0.0, 2.0], [1.0, 3.0], 8, 800),
                                             ([1.0, 2.5], [2.0, 3.5], 3, 600)])
def test_polyring_meshes(pk, a, b, sides, ntri):
    """Structured polyring meshes: the triangle areas add up to the polygon rings'
    closed-form area (sides/2 sin(2 pi/sides) (b^2 - a^2) each), every triangle is
    counter-clockwise, every edge is shared by at most two faces, and the Euler
    characteristic V - E + F counts one per solid polygon and zero per annulus."""
    sh = pk.polyring(s1=[10, 0], s2=[0, 10], a=a, b=b, sides=sides, ntri=ntri, units="mm")
    p = sh.rho[sh.fv]
    ar = 0.5 * ((p[1, :, 0] - p[0, :, 0]) * (p[2, :, 1] - p[0, :, 1])
                - (p[2, :, 0] - p[0, :, 0]) * (p[1, :, 1] - p[0, :, 1]))
    exact = sum(0.5 * sides * math.sin(2 * math.pi / sides) * (bb * bb - aa * aa)
                for aa, bb in zip(a, b))
    assert (ar > 0).all()
    assert abs(ar.sum() - exact) < 1e-12 * exact
    assert 0.5 * ntri <= sh.nface <= 2 * ntri
    cnt = np.bincount(sh.fe.reshape(-1), minlength=sh.nedge)
    assert cnt.max() <= 2 and cnt.min() >= 1
    assert sh.rho.shape[0] - sh.nedge + sh.nface == sum(1 for aa in a if aa == 0)
    # every node on its polygon: |x - c| projected on the vertex directions
    c = np.array([5.0, 5.0])
    r = np.linalg.norm(sh.rho - c, axis=1)
    assert r.max() <= max(b) + 1e-12
    rw = pk.rwg.setup_rwg(sh)
    assert rw.nbf == int((cnt == 2).sum())


@pytest.mark.parametrize("kw,msg", [(dict(a=[1.0], b=[1.0]), "a\\[1\\] >= b\\[1\\]"),
                                    (dict(a=[0.0, 1.0], b=[1.5, 2.0]), "b\\[1\\] >= a\\[2\\]"),
                                    (dict(a=[0.0], b=[1.0, 2.0]), "length"),
                                    (dict(a=[1.0], b=[2.0], mesher="triangle"), "mesher"),
                                    (dict(a=[0.0], b=[1.0], sides=2), "sides")])
def test_polyring_argument_checks(pk, kw, msg):
    args = dict(s1=[10, 0], s2=[0, 10], sides=6, ntri=100, units="mm")
    args.update(kw)
    with pytest.raises(ValueError, match=msg):
        pk.polyring(**args)


def test_cross_batch_generator(pk):
    """Seeded generator: reproducible, parameters inside the stated ranges, meshes
    between about 800 and 4000 faces."""
    a = pk.workloads.cross_batch()
    b = pk.workloads.cross_batch()
    assert len(a) == 16
    for (sa, fa, st, pa), (sb, _, _, pb) in zip(a, b):
        assert pa == pb and np.array_equal(sa[1].rho, sb[1].rho)
        P = pa["P"]
        assert 8 <= P <= 15
        assert 0.6 * P <= pa["L1"] <= 0.9 * P
        assert 700 <= sa[1].nface <= 4500
        assert len(fa) == 64 and fa[0] == 1.0 and fa[-1] == 30.0
        assert st["theta"] == [0.0, 15.0, 30.0, 45.0]

@pytest.mark.skipif(gpu_available(), reason="checks the status without a device")
def test_context_without_device(pk):
    """Without a GPU, context creation fails with PSSFSS_EHIP and no handle, and a
    null context is reported as such (no crash, no fallback)."""
    import ctypes as C
    from pssfss_amd import hip
    h = C.c_void_p()
    rc = hip.lib().pssfss_ctx_create(0, C.byref(h))
    assert rc == -2 and h.value is None
    assert hip.lib().pssfss_last_error(None) == b"null context"


def test_sheet_data_round_trip(pk, tmp_path):
    """write_sheet_data / read_sheet_data keep every field, and the read sheet
    gives the same RWG data."""
    sh = pk.sheets.meander(a=6.0, b=10.0, h=2.5, w1=0.3, w2=0.3, ntri=600, units="mm")
    path = str(tmp_path / "sheet.npz")
    pk.sheets.write_sheet_data(path, sh)
    sh2 = pk.sheets.read_sheet_data(path)
    for k in pk.sheets._SHEET_ARRAYS:
        assert np.array_equal(getattr(sh, k), getattr(sh2, k)), k
    for k in pk.sheets._SHEET_SCALARS:
        assert getattr(sh, k) == getattr(sh2, k), k
    r1, r2 = pk.setup_rwg(sh), pk.setup_rwg(sh2)
    assert r1.nbf == r2.nbf and r1.nufp == r2.nufp
    assert np.array_equal(r1.bff, r2.bff) and np.array_equal(r1.ebf, r2.ebf)
    np.savez(str(tmp_path / "other.npz"), x=np.zeros(3))
    with pytest.raises(ValueError, match="does not contain sheet data"):
        pk.sheets.read_sheet_data(str(tmp_path / "other.npz"))


@pytest.mark.parametrize("kw, msg", [(dict(fufp=True), "fufp=False"),
                                     (dict(fufp=False), None)])
def test_adjoint_sheet_checks(pk, kw, msg):
    """The adjoint takes 'J' and 'M' sheets whose face pairs are each their own
    class; a sheet meshed with translation classes is refused before any device
    work."""
    from pssfss_amd.adjoint import _check_adjoint_sheet
    from pssfss_amd.rwg import setup_rwg
    for cls in ("J", "M"):
        sh = pk.sheets.rectstrip(Lx=6.0, Ly=6.0, Px=12.0, Py=12.0, Nx=4, Ny=4, units="mm",
                                 sheet_class=cls, **kw)
        rd = setup_rwg(sh)
        if msg is None:
            _check_adjoint_sheet(sh, rd, "gradients")
        else:
            with pytest.raises(ValueError, match=msg):
                _check_adjoint_sheet(sh, rd, "gradients")
