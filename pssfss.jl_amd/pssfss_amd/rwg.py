"""RWG basis-function maps and unique-face-pair classes (`setup_rwg`, src/RWG.jl:90-324).

Basis functions are numbered exactly as the reference numbers them: first the
interior edges in lexicographic order of their (plus face, minus face) pair
(src/RWG.jl:152-163), then the eta=0 / eta=1 periodic pairs in order of the eta=0
edge index, then the xi=0 / xi=1 pairs (src/RWG.jl:166-254).

Unique face pairs.  The reference groups face pairs whose 6-vector
(c_m - r1_n, c_m - r2_n, c_m - r3_n) agree within 1e-6 |c_1 - c_2| using a KD-tree
over all nface^2 pairs (src/RWG.jl:287-319).  Here the same classes come from
hashing: the vectors are rounded to a grid far finer than the mesh spacing, and
every class is checked against the reference radius afterwards.  Classes are
ordered by their smallest member (column-major face-pair index ifm + nface*ifs),
members ascending, as the reference's scan produces them.
"""
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .sheets import RWGSheet


@dataclass
class RWGData:
    bfe: np.ndarray      # (2, nbf) 0-based edge indices of plus/minus half basis functions
    bff: np.ndarray      # (2, nbf) 0-based plus/minus face indices
    ebf: np.ndarray      # (nedge,) basis function of each edge, -1 if none
    eci: np.ndarray      # (nedge,) cell-boundary code 0..4
    nufp: int
    ufp_identity: bool   # True: every face pair is its own class (nufp = nface^2)
    ufp_offsets: Optional[np.ndarray] = None   # (nufp+1,) CSR offsets into ufp_members
    ufp_members: Optional[np.ndarray] = None   # face-pair indices ifm + nface*ifs, 0-based
    ufpm: Optional[np.ndarray] = None          # (nface, nface) class of each face pair

    @property
    def nbf(self) -> int:
        return self.bfe.shape[1]

    def ufp_rep(self) -> np.ndarray:
        """Representative face-pair index of every class (its smallest member)."""
        if self.ufp_identity:
            return np.arange(self.nufp, dtype=np.int64)
        return self.ufp_members[self.ufp_offsets[:-1]]


def _edge_xieta(sheet: RWGSheet):
    twopi = 2.0 * np.pi
    r1 = sheet.rho[sheet.e1]
    r2 = sheet.rho[sheet.e2]
    b1, b2 = sheet.beta1, sheet.beta2
    xi_i = (b1[0] * r1[:, 0] + b1[1] * r1[:, 1]) / twopi
    xi_t = (b1[0] * r2[:, 0] + b1[1] * r2[:, 1]) / twopi
    eta_i = (b2[0] * r1[:, 0] + b2[1] * r1[:, 1]) / twopi
    eta_t = (b2[0] * r2[:, 0] + b2[1] * r2[:, 1]) / twopi
    return xi_i, xi_t, eta_i, eta_t


def setup_rwg(sheet: RWGSheet, fufp: Optional[bool] = None, classify=None) -> RWGData:
    """Basis-function maps and face-pair classes.  classify(sheet) -> (offsets,
    members, ufpm) replaces the host classification (e.g. the device one of
    hip.face_pair_classes)."""
    tol = 1.0e-5
    nface = sheet.nface
    nedge = sheet.nedge
    fe = sheet.fe
    eci = np.zeros(nedge, np.int64)
    ie_xi0, ie_xi1, ie_eta0, ie_eta1 = [], [], [], []

    # Interior edges: shared by exactly two faces.
    edge_faces = [[] for _ in range(nedge)]
    for f in range(nface):
        for i in range(3):
            edge_faces[fe[i, f]].append(f)
    pairs = []
    for e in range(nedge):
        fl = edge_faces[e]
        if len(fl) == 2:
            fp, fm = min(fl), max(fl)
            pairs.append((fp, fm, e))
        elif len(fl) > 2:
            raise ValueError(f"edge {e + 1} is shared by {len(fl)} faces")
    pairs.sort()

    if sheet.xieta_check:
        xi_i, xi_t, eta_i, eta_t = _edge_xieta(sheet)
        for e in range(nedge):
            if abs(xi_i[e]) < tol and abs(xi_t[e]) < tol:
                eci[e] = 1
                ie_xi0.append(e)
            elif abs(xi_i[e] - 1) < tol and abs(xi_t[e] - 1) < tol:
                eci[e] = 2
                ie_xi1.append(e)
            elif abs(eta_i[e]) < tol and abs(eta_t[e]) < tol:
                eci[e] = 3
                ie_eta0.append(e)
            elif abs(eta_i[e] - 1) < tol and abs(eta_t[e] - 1) < tol:
                eci[e] = 4
                ie_eta1.append(e)
    if len(ie_xi0) != len(ie_xi1):
        raise ValueError("Inconsistent # edges at ξ=0 and ξ=1")
    if len(ie_eta0) != len(ie_eta1):
        raise ValueError("Inconsistent # edges at η=0 and η=1")

    nbf = len(pairs) + len(ie_xi0) + len(ie_eta0)
    bfe = np.zeros((2, nbf), np.int64)
    bff = np.zeros((2, nbf), np.int64)
    ebf = np.full(nedge, -1, np.int64)
    i = 0
    for fp, fm, e in pairs:
        bfe[:, i] = (e, e)
        bff[:, i] = (fp, fm)
        ebf[e] = i
        i += 1

    def first_face(e):
        for f in range(nface):
            if e in fe[:, f]:
                return f
        raise ValueError(f"Unable to locate face containing edge {e + 1}")

    if sheet.xieta_check:
        xi_i, xi_t, eta_i, eta_t = _edge_xieta(sheet)
    for (list0, list1, lo, hi, lab) in (
            (ie_eta0, ie_eta1, None, None, "η"), (ie_xi0, ie_xi1, None, None, "ξ")):
        for ge0 in list0:
            face0 = first_face(ge0)
            if lab == "η":
                v0 = (xi_i[ge0], xi_t[ge0])
            else:
                v0 = (eta_i[ge0], eta_t[ge0])
            mn0, mx0 = min(v0), max(v0)
            ge1 = -1
            for ge in list1:
                v1 = (xi_i[ge], xi_t[ge]) if lab == "η" else (eta_i[ge], eta_t[ge])
                if abs(min(v1) - mn0) < tol and abs(max(v1) - mx0) < tol:
                    ge1 = ge
                    break
            if ge1 < 0:
                raise ValueError(f"Unable to locate edge at {lab}=1 corresponding to edge {ge0 + 1}")
            face1 = first_face(ge1)
            bfe[:, i] = (ge0, ge1)
            bff[:, i] = (face0, face1)
            ebf[ge0] = i
            ebf[ge1] = i
            i += 1
    if i != nbf:
        raise ValueError("Inconsistent number of basis functions")

    use_fufp = sheet.fufp if fufp is None else fufp
    if not use_fufp:
        return RWGData(bfe, bff, ebf, eci, nface * nface, True)
    offsets, members, ufpm = (classify or unique_face_pairs)(sheet)
    return RWGData(bfe, bff, ebf, eci, len(offsets) - 1, False, offsets, members, ufpm)


def unique_face_pairs(sheet: RWGSheet):
    """Translation classes of face pairs (column-major index ifm + nface*ifs)."""
    nface = sheet.nface
    rho = sheet.rho
    r1 = rho[sheet.fv[0]]
    r2 = rho[sheet.fv[1]]
    r3 = rho[sheet.fv[2]]
    centroid = (r1 + r2 + r3) / 3.0
    radius = 1e-6 * np.linalg.norm(centroid[0] - centroid[1])
    # Quantization step: far above rounding noise, far below any mesh spacing.
    q = 64.0 * radius
    # data[mn] for mn = m + nface*n (m observation face, n source face)
    cm = centroid[None, :, :]                    # indexed [n, m]
    d1 = cm - r1[:, None, :]
    d2 = cm - r2[:, None, :]
    d3 = cm - r3[:, None, :]
    data = np.concatenate([d1, d2, d3], axis=2).reshape(nface * nface, 6)
    keys = np.rint(data / q).astype(np.int64)
    # Classes of equal 6-component keys: a 64-bit mix of the key sorts fast; every
    # class is then checked member by member against its first key, and the exact
    # row-wise grouping is used if a collision ever merged different keys.
    mult = np.array([0x9E3779B97F4A7C15, 0xC2B2AE3D27D4EB4F, 0x165667B19E3779F9,
                     0xD6E8FEB86659FD93, 0xFF51AFD7ED558CCD, 0xC4CEB9FE1A85EC53], np.uint64)
    with np.errstate(over="ignore"):
        h = (keys.astype(np.uint64) * mult).sum(axis=1, dtype=np.uint64)
        h ^= h >> np.uint64(31)
    _, first_idx, inverse = np.unique(h, return_index=True, return_inverse=True)
    inverse = inverse.reshape(-1)
    if not np.array_equal(keys, keys[first_idx[inverse]]):
        _, first_idx, inverse = np.unique(keys, axis=0, return_index=True,
                                          return_inverse=True)
        inverse = inverse.reshape(-1)
    # Order classes by smallest member.
    order = np.argsort(first_idx, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    cls = rank[inverse]
    members = np.argsort(cls, kind="stable").astype(np.int64)
    counts = np.bincount(cls, minlength=len(order))
    offsets = np.zeros(len(order) + 1, np.int64)
    offsets[1:] = np.cumsum(counts)
    # Verify every member lies within the reference radius of its representative.
    rep = members[offsets[:-1]]
    dist = np.linalg.norm(data[members] - data[np.repeat(rep, counts)], axis=1)
    if dist.size and dist.max() >= radius:
        raise ValueError("unique-face-pair classes are not within the grouping radius; "
                         "mesh has near-coincident face pairs at the quantization scale")
    ufpm = cls.reshape(nface, nface).T.copy()   # ufpm[m, n]
    return offsets, members, ufpm
