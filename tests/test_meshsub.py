"""Unstructured mesher (pssfss_amd.meshsub), which replaces Triangle for the
reference's meshsub (src/Meshsub.jl:42-111) and the cell-filling outer ring of
polyring (src/Elements.jl:393-480).  Mesh coordinates cannot be pinned to
Triangle's; these checks pin what any valid conforming triangulation of the
region must satisfy: exact area, edges shared by at most two faces with the
single-face edges on the outline, Euler characteristic, counter-clockwise faces,
no face in a hole, the reference's +-10 % triangle-count rule, a bounded minimum
angle, and reproducibility for a seed."""
import math

import numpy as np
import pytest


def _areas(sh):
    p = sh.rho[sh.fv]
    return 0.5 * ((p[1, :, 0] - p[0, :, 0]) * (p[2, :, 1] - p[0, :, 1])
                  - (p[2, :, 0] - p[0, :, 0]) * (p[1, :, 1] - p[0, :, 1]))


def _min_angle_deg(sh):
    p = sh.rho[sh.fv]
    out = []
    for i in range(3):
        u = p[(i + 1) % 3] - p[i]
        v = p[(i + 2) % 3] - p[i]
        c = np.sum(u * v, 1) / np.linalg.norm(u, axis=1) / np.linalg.norm(v, axis=1)
        out.append(np.degrees(np.arccos(np.clip(c, -1, 1))))
    return float(np.min(out))


def _on_outline(pts, loops, tol=1e-9):
    """Distance of each point to the nearest outline segment is below tol."""
    d = np.full(len(pts), np.inf)
    for loop in loops:
        L = np.asarray(loop, float)
        for a, b in zip(L, np.roll(L, -1, axis=0)):
            ab = b - a
            t = np.clip(((pts - a) @ ab) / (ab @ ab), 0, 1)
            d = np.minimum(d, np.linalg.norm(pts - (a + t[:, None] * ab), axis=1))
    return d < tol


def _check(sh, loops, area, ntri, holes, euler):
    ar = _areas(sh)
    assert (ar > 0).all(), "clockwise or degenerate face"
    assert abs(ar.sum() - area) < 1e-12 * area
    assert 0.9 * ntri <= sh.nface <= 1.1 * ntri
    cnt = np.bincount(sh.fe.reshape(-1), minlength=sh.nedge)
    assert cnt.max() <= 2 and cnt.min() >= 1
    # edges of one face lie on the outline (no cracks inside the region)
    bd = np.flatnonzero(cnt == 1)
    mid = 0.5 * (sh.rho[sh.e1[bd]] + sh.rho[sh.e2[bd]])
    assert _on_outline(mid, loops).all()
    assert sh.rho.shape[0] - sh.nedge + sh.nface == euler
    cen = sh.rho[sh.fv].mean(axis=0)
    for inside_hole in holes:
        assert not inside_hole(cen).any()
    assert _min_angle_deg(sh) > 20.0


def test_meshsub_square(pk):
    M = pk.meshsub
    loop = [(0.0, 0.0), (2.0, 0.0), (2.0, 1.0), (0.0, 1.0)]
    P, S = M.outline_loops([loop])
    sh = M.meshsub(points=P, seglist=S, area=2.0 / 300, ntri=300)
    _check(sh, [loop], 2.0, 300, [], 1)
    # the same seed gives the same mesh; another seed a valid one
    sh2 = M.meshsub(points=P, seglist=S, area=2.0 / 300, ntri=300)
    assert np.array_equal(sh.fv, sh2.fv) and np.array_equal(sh.rho, sh2.rho)
    sh3 = M.meshsub(points=P, seglist=S, area=2.0 / 300, ntri=300, seed=7)
    _check(sh3, [loop], 2.0, 300, [], 1)


def test_meshsub_region_with_holes(pk):
    """A square with two square holes: Euler characteristic 1 - 2 and no face
    centroid inside a hole; a hole point inside the region is refused."""
    M = pk.meshsub
    outer = [(0, 0), (10, 0), (10, 10), (0, 10)]
    h1 = [(2, 2), (4, 2), (4, 4), (2, 4)]
    h2 = [(6, 5), (8, 5), (8, 8), (6, 8)]
    P, S = M.outline_loops([outer, h1, h2])
    area = 100 - 4 - 6
    sh = M.meshsub(points=P, seglist=S, area=area / 500, ntri=500,
                   holes=np.array([[3.0, 7.0], [3.0, 6.5]]))
    in_h1 = lambda c: (c[:, 0] > 2) & (c[:, 0] < 4) & (c[:, 1] > 2) & (c[:, 1] < 4)
    in_h2 = lambda c: (c[:, 0] > 6) & (c[:, 0] < 8) & (c[:, 1] > 5) & (c[:, 1] < 8)
    _check(sh, [outer, h1, h2], area, 500, [in_h1, in_h2], -1)
    with pytest.raises(ValueError, match="hole"):
        M.meshsub(points=P, seglist=S, area=area / 500, ntri=500, holes=np.array([[5.0], [1.0]]))


@pytest.mark.parametrize("sides,a,ns1", [(4, 3.0, 12), (6, 2.5, 10), (8, 4.0, 16)])
def test_polyring_fill_cell(pk, sides, a, ns1):
    """Cell-filling outer ring (b[end] = -ns1): metal area = cell - polygon; the cell
    sides keep ns1 (and ns2) nodes that coincide under the lattice translations, so
    setup_rwg pairs every boundary edge with one across the cell."""
    s1, s2 = [10.0, 0.0], [0.0, 12.0]
    sh = pk.polyring(s1=s1, s2=s2, a=[a], b=[-ns1], sides=sides, ntri=600, units="mm",
                     orient=22.5)
    assert sh.xieta_check
    poly = 0.5 * sides * math.sin(2 * math.pi / sides) * a * a
    area = 120.0 - poly
    assert abs(_areas(sh).sum() - area) < 1e-12 * area
    assert 0.9 * 600 <= sh.nface <= 1.1 * 600
    ns2 = int(round(ns1 * 12.0 / 10.0))
    x, y = sh.rho[:, 0], sh.rho[:, 1]
    tol = 1e-9
    assert np.sort(x[np.abs(y) < tol]).tolist() == pytest.approx(
        np.sort(x[np.abs(y - 12) < tol]).tolist())
    assert np.sort(y[np.abs(x) < tol]).tolist() == pytest.approx(
        np.sort(y[np.abs(x - 10) < tol]).tolist())
    assert int((np.abs(y) < tol).sum()) == ns1 + 1 and int((np.abs(x) < tol).sum()) == ns2 + 1
    cnt = np.bincount(sh.fe.reshape(-1), minlength=sh.nedge)
    rw = pk.rwg.setup_rwg(sh)
    # interior edges, plus one basis function per pair of opposite cell-boundary edges
    assert rw.nbf == int((cnt == 2).sum()) + ns1 + ns2
    # the cell sides keep their given nodes (never split), so the triangles on them
    # can be flatter than elsewhere
    assert _min_angle_deg(sh) > 15.0


@pytest.mark.parametrize("w", [0.8, 2.0])
def test_loadedcross_delaunay(pk, w):
    """The unstructured loaded cross covers exactly the structured one's region."""
    kw = dict(s1=[10, 0], s2=[0, 10], L1=8, L2=3, w=w, ntri=700, units="mm")
    a = pk.loadedcross(mesher="delaunay", **kw)
    b = pk.loadedcross(**kw)
    assert abs(_areas(a).sum() - _areas(b).sum()) < 1e-12 * _areas(b).sum()
    assert 0.9 * 700 <= a.nface <= 1.1 * 700
    cnt = np.bincount(a.fe.reshape(-1), minlength=a.nedge)
    assert cnt.max() <= 2
    assert a.rho.shape[0] - a.nedge + a.nface == (0 if 2 * w < 3 else 1)
    assert pk.rwg.setup_rwg(a).nbf == int((cnt == 2).sum())


@pytest.mark.parametrize("w", [0.2, 1.0])
def test_jerusalemcross_delaunay(pk, w):
    kw = dict(P=10, L1=8, L2=1, A=5, B=1, w=w, ntri=900, units="mm")
    a = pk.jerusalemcross(mesher="delaunay", **kw)
    b = pk.jerusalemcross(**kw)
    assert abs(_areas(a).sum() - _areas(b).sum()) < 1e-12 * _areas(b).sum()
    assert 0.9 * 900 <= a.nface <= 1.1 * 900
    assert _min_angle_deg(a) > 20.0
