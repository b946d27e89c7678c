"""Unstructured triangulation of polygonal regions with a target triangle count.

The reference triangulates element outlines with Triangle (`meshsub`,
src/Meshsub.jl:42-111: constrained conforming Delaunay, maximum triangle area,
30-degree minimum angle, the area iterated up to 14 times until the triangle count
is within 10 % of `ntri`).  Triangle is not available here; this is our own mesher
with the same interface and the same `ntri` iteration, so the meshes differ from
Triangle's (parity of mesh coordinates is unpinned) while the result is a valid
conforming triangulation of the same region:

1. every boundary segment is split into pieces no longer than the target edge h
   (segments marked fixed keep their end points: the unit-cell sides of a
   cell-filling element, whose nodes must match across the period), then split
   again while another boundary point lies inside its diametral circle;
2. interior points are the nodes of an equilateral lattice of spacing h, with a
   small seeded jitter, that lie inside the region (even-odd rule over all
   segments: nested outlines alternate metal and hole), outside every segment's
   diametral circle and at least h/2 from the boundary points;
3. scipy's Delaunay triangulation of all points, keeping the triangles whose
   centroid is inside the region; a segment's diametral circle holds no other
   point, so it is an edge of the triangulation (Gabriel property) -- any segment
   that is nevertheless missing is split and the step repeated;
4. two passes of Laplacian smoothing of the interior points (a point moves only
   to a position that keeps the conditions of step 2), each followed by a fresh
   triangulation.

Faces are counter-clockwise; edge i of a face is opposite its vertex i.
"""
import math

import numpy as np
from scipy.spatial import Delaunay, cKDTree

from .sheets import RWGSheet, _build_edges

SQ3 = math.sqrt(3.0)


def _inside(pts: np.ndarray, A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """Even-odd rule: a horizontal ray from each point crosses the segments
    (A[k], B[k]) an odd number of times for points inside the region."""
    x, y = pts[:, 0:1], pts[:, 1:2]
    ay, by = A[None, :, 1], B[None, :, 1]
    ax, bx = A[None, :, 0], B[None, :, 0]
    straddle = (ay > y) != (by > y)
    with np.errstate(divide="ignore", invalid="ignore"):
        xc = ax + (y - ay) * (bx - ax) / (by - ay)
    cross = straddle & (x < xc)
    return (np.count_nonzero(cross, axis=1) % 2) == 1


def _encroaching(pts: np.ndarray, A: np.ndarray, B: np.ndarray, chunk: int = 2048) -> np.ndarray:
    """For each point, whether it lies strictly inside the diametral circle of some
    segment ((p - a).(p - b) < 0), the points equal to an end point excluded."""
    out = np.zeros(len(pts), bool)
    L2 = np.sum((B - A) ** 2, axis=1)
    for i0 in range(0, len(pts), chunk):
        p = pts[i0:i0 + chunk, None, :]
        d = np.sum((p - A[None]) * (p - B[None]), axis=2)
        out[i0:i0 + chunk] = np.any(d < -1e-12 * L2[None], axis=1)
    return out


def _split_segments(P: list, segs: list, fixed: list, h: float):
    """Pieces no longer than h for the free segments."""
    out, ofix = [], []
    for (i, j), fx in zip(segs, fixed):
        if fx:
            out.append((i, j))
            ofix.append(True)
            continue
        a, b = np.asarray(P[i]), np.asarray(P[j])
        k = max(1, int(math.ceil(np.linalg.norm(b - a) / h - 1e-9)))
        prev = i
        for q in range(1, k):
            P.append(tuple(a + (b - a) * (q / k)))
            out.append((prev, len(P) - 1))
            ofix.append(False)
            prev = len(P) - 1
        out.append((prev, j))
        ofix.append(False)
    return out, ofix


def _split_encroached(P: list, segs: list, fixed: list, max_rounds: int = 30):
    """Split free segments whose diametral circle holds another boundary point."""
    for _ in range(max_rounds):
        pts = np.asarray(P)
        A = pts[[s[0] for s in segs]]
        B = pts[[s[1] for s in segs]]
        tree = cKDTree(pts)
        bad = []
        for k, ((i, j), fx) in enumerate(zip(segs, fixed)):
            if fx:
                continue
            c = 0.5 * (A[k] + B[k])
            r = 0.5 * np.linalg.norm(B[k] - A[k])
            for q in tree.query_ball_point(c, r * (1 - 1e-9)):
                if q != i and q != j:
                    bad.append(k)
                    break
        if not bad:
            return segs, fixed
        badset = set(bad)
        nsegs, nfix = [], []
        for k, ((i, j), fx) in enumerate(zip(segs, fixed)):
            if k in badset:
                P.append(tuple(0.5 * (np.asarray(P[i]) + np.asarray(P[j]))))
                m = len(P) - 1
                nsegs += [(i, m), (m, j)]
                nfix += [False, False]
            else:
                nsegs.append((i, j))
                nfix.append(fx)
        segs, fixed = nsegs, nfix
    raise RuntimeError("meshsub: boundary refinement did not settle")


def _interior_points(bpts: np.ndarray, A: np.ndarray, B: np.ndarray, h: float,
                     rng: np.random.Generator) -> np.ndarray:
    lo, hi = bpts.min(axis=0), bpts.max(axis=0)
    dy = 0.5 * SQ3 * h
    ny = int(math.ceil((hi[1] - lo[1]) / dy)) + 1
    nx = int(math.ceil((hi[0] - lo[0]) / h)) + 2
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    x = lo[0] + (ii + 0.5 * (jj % 2)) * h
    y = lo[1] + jj * dy + 0.5 * dy
    cand = np.column_stack([x.ravel(), y.ravel()])
    cand = cand + rng.uniform(-1e-3, 1e-3, cand.shape) * h
    cand = cand[_inside(cand, A, B)]
    if len(cand) == 0:
        return cand
    d, _ = cKDTree(bpts).query(cand)
    cand = cand[d >= 0.5 * h]
    return cand[~_encroaching(cand, A, B)]


def _triangulate(pts: np.ndarray, A: np.ndarray, B: np.ndarray) -> np.ndarray:
    tri = Delaunay(pts, qhull_options="Qbb Qc Qz Q12").simplices
    cen = pts[tri].mean(axis=1)
    return tri[_inside(cen, A, B)]


def _missing_segments(tri: np.ndarray, segs: list) -> list:
    have = set()
    for a, b, c in tri:
        for u, v in ((a, b), (b, c), (c, a)):
            have.add((min(u, v), max(u, v)))
    return [k for k, (i, j) in enumerate(segs) if (min(i, j), max(i, j)) not in have]


def _mesh_once(points: np.ndarray, segs0: list, fixed0: list, area: float, seed: int):
    h = math.sqrt(4.0 * area / SQ3)
    P = [tuple(p) for p in points]
    segs, fixed = _split_segments(P, list(segs0), list(fixed0), h)
    rng = np.random.default_rng(seed)
    for _ in range(20):
        segs, fixed = _split_encroached(P, segs, fixed)
        bpts = np.asarray(P)
        A = bpts[[s[0] for s in segs]]
        B = bpts[[s[1] for s in segs]]
        inner = _interior_points(bpts, A, B, h, rng)
        pts = np.vstack([bpts, inner]) if len(inner) else bpts
        tri = _triangulate(pts, A, B)
        miss = _missing_segments(tri, segs)
        if not miss:
            break
        if any(fixed[k] for k in miss):
            raise RuntimeError("meshsub: a fixed boundary segment is not conforming; use "
                               "more boundary points")
        missing = set(miss)
        nsegs, nfix = [], []
        for k, ((i, j), fx) in enumerate(zip(segs, fixed)):
            if k in missing:
                P.append(tuple(0.5 * (np.asarray(P[i]) + np.asarray(P[j]))))
                m = len(P) - 1
                nsegs += [(i, m), (m, j)]
                nfix += [False, False]
            else:
                nsegs.append((i, j))
                nfix.append(fx)
        segs, fixed = nsegs, nfix
    else:
        raise RuntimeError("meshsub: could not recover every boundary segment")
    nb = len(bpts)
    # Laplacian smoothing of the interior points, each pass re-triangulated
    for _ in range(2):
        if len(pts) == nb:
            break
        nbr_sum = np.zeros_like(pts)
        cnt = np.zeros(len(pts))
        for u, v in ((0, 1), (1, 2), (2, 0)):
            np.add.at(nbr_sum, tri[:, u], pts[tri[:, v]])
            np.add.at(nbr_sum, tri[:, v], pts[tri[:, u]])
            np.add.at(cnt, tri[:, u], 1)
            np.add.at(cnt, tri[:, v], 1)
        new = pts.copy()
        idx = np.arange(nb, len(pts))
        ok = cnt[idx] > 0
        idx = idx[ok]
        prop = nbr_sum[idx] / cnt[idx, None]
        d, _ = cKDTree(bpts).query(prop)
        good = _inside(prop, A, B) & (d >= 0.25 * h) & ~_encroaching(prop, A, B)
        new[idx[good]] = prop[good]
        tri2 = _triangulate(new, A, B)
        if _missing_segments(tri2, segs):
            break
        pts, tri = new, tri2
    # drop interior points no triangle uses (none expected) and renumber
    used = np.zeros(len(pts), bool)
    used[tri.ravel()] = True
    used[:nb] = True
    remap = -np.ones(len(pts), np.int64)
    remap[used] = np.arange(np.count_nonzero(used))
    pts = pts[used]
    tri = remap[tri]
    # counter-clockwise faces
    p0, p1, p2 = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
    cr = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p1[:, 1] - p0[:, 1]) * (p2[:, 0] - p0[:, 0])
    flip = cr < 0
    tri[flip, 1], tri[flip, 2] = tri[flip, 2].copy(), tri[flip, 1].copy()
    return pts, tri


def meshsub(*, points, seglist, area: float, ntri: int, segmarkers=None, holes=None,
            fixed=None, seed: int = 0, max_iter: int = 14) -> RWGSheet:
    """Triangulate the region bounded by the segments (src/Meshsub.jl:42-111).

    points: (2, npoint) coordinates; seglist: (2, nseg) 0-based point
    indices of the boundary segments; area: initial target triangle area; ntri:
    desired triangle count (the area is rescaled until the count is within 10 %,
    at most max_iter triangulations, as the reference iterates); holes: points
    inside regions not to be meshed (checked against the even-odd rule, which
    decides the region); fixed: per segment, True to keep it unsplit; segmarkers:
    accepted for the reference's signature.  Returns an RWGSheet with rho, e1, e2,
    fv, fe and zero fr."""
    pts = np.asarray(points, float)
    if pts.ndim != 2 or pts.shape[0] != 2:
        raise ValueError("points first dimension must be 2")
    pts = pts.T.copy()
    sl = np.asarray(seglist, np.int64)
    if sl.ndim != 2 or sl.shape[0] != 2:
        raise ValueError("seglist first dimension must be 2")
    if segmarkers is not None and len(segmarkers) != sl.shape[1]:
        raise ValueError("wrong length(segmarkers)")
    if not ntri > 0 or not area > 0:
        raise ValueError("ntri and area must be positive")
    segs = [(int(a), int(b)) for a, b in sl.T]
    fx = [False] * len(segs) if fixed is None else [bool(v) for v in fixed]
    A = pts[sl[0]]
    B = pts[sl[1]]
    if holes is not None:
        hp = np.asarray(holes, float)
        if hp.size:
            hp = hp.reshape(2, -1).T
            if np.any(_inside(hp, A, B)):
                raise ValueError("a hole point lies inside the region")
    a = float(area)
    best = None
    for _ in range(max_iter):
        rho, tri = _mesh_once(pts, segs, fx, a, seed)
        n = len(tri)
        if best is None or abs(n - ntri) < abs(len(best[1]) - ntri):
            best = (rho, tri)
        if 0.9 * ntri <= n <= 1.1 * ntri:
            break
        a *= n / ntri
    rho, tri = best
    fv = tri.T.copy()
    e1, e2, fe = _build_edges(fv)
    sh = RWGSheet()
    sh.rho = rho
    sh.e1, sh.e2, sh.fv, sh.fe = e1, e2, fv, fe
    sh.fr = np.zeros(fv.shape[1])
    return sh


def polygon_loop(center, radius: float, sides: int, orient_deg: float, start: int):
    """Vertices of a regular polygon (first at angle orient) and its closed segment
    loop, numbered from `start`."""
    t0 = math.radians(orient_deg)
    al = 2 * math.pi / sides
    v = [tuple(np.asarray(center) + radius * np.array([math.cos(t0 + k * al),
                                                       math.sin(t0 + k * al)]))
         for k in range(sides)]
    seg = [(start + k, start + (k + 1) % sides) for k in range(sides)]
    return v, seg


def outline_loops(loops) -> tuple:
    """Points and segments of closed outlines given as vertex lists."""
    pts, segs = [], []
    for loop in loops:
        s = len(pts)
        pts.extend(tuple(map(float, p)) for p in loop)
        n = len(loop)
        segs.extend((s + k, s + (k + 1) % n) for k in range(n))
    return np.asarray(pts).T, np.asarray(segs, np.int64).T


def cross_outline(L1: float, L2: float):
    """Counter-clockwise outline of a plus-shaped cross centred at the origin:
    tip-to-tip length L1, arm width L2."""
    a, b = 0.5 * L1, 0.5 * L2
    return [(b, -a), (b, -b), (a, -b), (a, b), (b, b), (b, a), (-b, a), (-b, b), (-a, b),
            (-a, -b), (-b, -b), (-b, -a)]


def jcross_outline(L1: float, L2: float, A: float, B: float):
    """Counter-clockwise outline of a Jerusalem cross centred at the origin: arms of
    width L2 spanning L1, end caps of length A (across the arm) and width B."""
    a, b, c = 0.5 * L1, 0.5 * L2, 0.5 * A
    q = [(a - B, -b), (a - B, -c), (a, -c), (a, c), (a - B, c), (a - B, b)]   # +x cap

    def rot(p, k):
        x, y = p
        for _ in range(k):
            x, y = -y, x
        return (x, y)

    pts = []
    for k in range(4):
        pts.append(rot((b, -b), k))
        pts.extend(rot(p, k) for p in q)
    return pts
