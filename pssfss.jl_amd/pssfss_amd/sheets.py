"""`RWGSheet` and the structured triangulations the sweep workloads use.

Field meanings follow src/Sheets.jl:19-58.  Indices are stored 0-based here
(the reference is 1-based); the C ABI takes 1-based int64 arrays as the Julia
arrays are, and converts once at upload.

Local numbering convention (relied on by the fill): edge `fe[i, f]` is the edge
opposite vertex `fv[i, f]`, as produced by `recttri` (src/Sheets.jl:458-477).
"""
from dataclasses import dataclass, field
import math

import numpy as np

from .constants import units_per_meter


def s1s2_to_beta(s1, s2):
    """Reciprocal lattice vectors from direct lattice vectors (src/Elements.jl:27-36)."""
    s1 = np.asarray(s1, float)
    s2 = np.asarray(s2, float)
    cross = s1[0] * s2[1] - s1[1] * s2[0]
    fact = 2.0 * math.pi / abs(cross)
    # s2 x zhat = (s2y, -s2x); zhat x s1 = (-s1y, s1x)
    b1 = np.array([fact * s2[1], fact * -s2[0]])
    b2 = np.array([fact * -s1[1], fact * s1[0]])
    return b1 + 0.0, b2 + 0.0


@dataclass(eq=False)
class RWGSheet:
    style: str = ""
    units: str = "mm"
    s1: np.ndarray = field(default_factory=lambda: np.zeros(2))
    s2: np.ndarray = field(default_factory=lambda: np.zeros(2))
    beta1: np.ndarray = field(default_factory=lambda: np.zeros(2))
    beta2: np.ndarray = field(default_factory=lambda: np.zeros(2))
    dx: float = 0.0
    dy: float = 0.0
    rot: float = 0.0
    rho: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))    # (nnode, 2)
    e1: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))  # 0-based
    e2: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    fv: np.ndarray = field(default_factory=lambda: np.zeros((3, 0), np.int64))
    fe: np.ndarray = field(default_factory=lambda: np.zeros((3, 0), np.int64))
    fr: np.ndarray = field(default_factory=lambda: np.zeros(0))
    psi1: float = 0.0
    psi2: float = 0.0
    u: float = 0.0
    sheet_class: str = "J"
    info: str = ""
    xieta_check: bool = True
    fufp: bool = False

    @property
    def nface(self) -> int:
        return self.fv.shape[1]

    @property
    def nedge(self) -> int:
        return len(self.e1)

    @property
    def upm(self) -> float:
        return units_per_meter(self.units)


def recttri(rhobl, rhotr, nx: int, ny: int) -> RWGSheet:
    """Structured triangulation of a rectangle (src/Sheets.jl:398-480)."""
    rhobl = np.asarray(rhobl, float)
    rhotr = np.asarray(rhotr, float)
    drho = (rhotr - rhobl) / np.array([nx, ny], float)
    rho = np.empty(((nx + 1) * (ny + 1), 2))
    n = 0
    for j in range(ny + 1):
        yj = j * drho[1]
        for i in range(nx + 1):
            rho[n, 0] = rhobl[0] + i * drho[0]
            rho[n, 1] = rhobl[1] + yj
            n += 1
    e1, e2 = [], []
    for j in range(ny + 1):              # horizontal edges
        kadd = j * (nx + 1)
        for i in range(1, nx + 1):
            e1.append(i + kadd)
            e2.append(i + kadd + 1)
    for j in range(1, ny + 1):           # vertical edges
        kadd = (j - 1) * (nx + 1) + 1
        for i in range(nx + 1):
            e1.append(i + kadd)
            e2.append(i + kadd + nx + 1)
    for j in range(1, ny + 1):           # diagonal edges
        kadd1 = (j - 1) * (nx + 1)
        kadd2 = 1 + j * (nx + 1)
        for i in range(1, nx + 1):
            e1.append(i + kadd1)
            e2.append(i + kadd2)
    nface = 2 * nx * ny
    fv = np.zeros((3, nface), np.int64)
    fe = np.zeros((3, nface), np.int64)
    nhe = nx * ny + nx
    nve = nx * ny + ny
    f = 0
    for j in range(1, ny + 1):
        nadd1 = (j - 1) * (nx + 1)
        nadd2 = 1 + j * (nx + 1)
        for i in range(1, nx + 1):
            fv[:, f] = (i + nadd1, i + nadd2, i + nadd2 - 1)
            fe[:, f] = (i + j * nx, i + nhe + nadd1, i + nhe + nve + (j - 1) * nx)
            f += 1
            fv[:, f] = (fv[0, f - 1], 1 + fv[0, f - 1], fv[1, f - 1])
            fe[:, f] = (1 + fe[1, f - 1], fe[2, f - 1], fe[0, f - 1] - nx)
            f += 1
    sh = RWGSheet()
    sh.rho = rho
    sh.e1 = np.asarray(e1, np.int64) - 1
    sh.e2 = np.asarray(e2, np.int64) - 1
    sh.fv = fv - 1
    sh.fe = fe - 1
    return sh


def _rotate(sh: RWGSheet, rot: float) -> None:
    if rot == 0:
        return
    s = math.sin(math.radians(rot))
    c = math.cos(math.radians(rot))
    R = np.array([[c, -s], [s, c]])
    sh.s1 = R @ sh.s1
    sh.s2 = R @ sh.s2
    sh.beta1 = R @ sh.beta1
    sh.beta2 = R @ sh.beta2
    sh.rho = sh.rho @ R.T
    sh.rot = rot


def rectstrip(*, Lx, Ly, Nx, Ny, Px, Py, units, sheet_class="J", dx=0.0, dy=0.0,
              rot=0.0, Rsheet=0.0, fufp=True) -> RWGSheet:
    """Rectangular strip centred in a rectangular cell (src/Elements.jl:149-186)."""
    for name, v in (("Lx", Lx), ("Ly", Ly), ("Nx", Nx), ("Ny", Ny), ("Px", Px), ("Py", Py)):
        if not v > 0:
            raise ValueError(f"{name} must be positive!")
    if sheet_class not in ("J", "M"):
        raise ValueError(f"Illegal value {sheet_class} for class")
    x0 = 0.5 * (Px - Lx)
    y0 = 0.5 * (Py - Ly)
    sh = recttri([x0, y0], [x0 + Lx, y0 + Ly], Nx, Ny)
    sh.style = "rectstrip"
    sh.units = units
    sh.s1 = np.array([float(Px), 0.0])
    sh.s2 = np.array([0.0, float(Py)])
    sh.beta1, sh.beta2 = s1s2_to_beta(sh.s1, sh.s2)
    sh.fr = np.full(sh.nface, float(Rsheet))
    sh.fufp = bool(fufp)
    sh.sheet_class = sheet_class
    _rotate(sh, rot)
    if dx != 0 or dy != 0:
        sh.rho = sh.rho + np.array([dx, dy], float)
    sh.xieta_check = (Lx == Px or Ly == Py)
    return sh


def _build_edges(fv: np.ndarray):
    """Edge list and face/edge map for a triangle list, edge i of a face opposite vertex i."""
    edge_index = {}
    e1, e2 = [], []
    nface = fv.shape[1]
    fe = np.zeros((3, nface), np.int64)
    for f in range(nface):
        for i in range(3):
            a = int(fv[(i + 1) % 3, f])
            b = int(fv[(i + 2) % 3, f])
            key = (min(a, b), max(a, b))
            e = edge_index.get(key)
            if e is None:
                e = len(e1)
                edge_index[key] = e
                e1.append(key[0])
                e2.append(key[1])
            fe[i, f] = e
    return np.asarray(e1, np.int64), np.asarray(e2, np.int64), fe


def annular_ring(*, a, b, P, nr, nphi, units, sheet_class="J", phi0=0.0, fufp=False) -> RWGSheet:
    """Structured polar triangulation of a metal annulus a <= r <= b centred in a square cell.

    This replaces the Triangle-based `polyring` mesher (src/Elements.jl:368-558,
    src/Meshsub.jl:42-111), which is not available here: an nr x nphi polar grid,
    each quad split by one diagonal, gives 2*nr*nphi faces with no cell-boundary
    edges.  fufp defaults to False as for the reference's unstructured styles.
    """
    nodes = []
    for ir in range(nr + 1):
        r = a + (b - a) * ir / nr
        for ip in range(nphi):
            t = phi0 + 2.0 * math.pi * ip / nphi
            nodes.append((0.5 * P + r * math.cos(t), 0.5 * P + r * math.sin(t)))
    rho = np.asarray(nodes)

    def nid(ir, ip):
        return ir * nphi + (ip % nphi)

    faces = []
    for ir in range(nr):
        for ip in range(nphi):
            p00, p01 = nid(ir, ip), nid(ir, ip + 1)
            p10, p11 = nid(ir + 1, ip), nid(ir + 1, ip + 1)
            faces.append((p00, p11, p10))   # counter-clockwise
            faces.append((p00, p01, p11))
    fv = np.asarray(faces, np.int64).T.copy()
    e1, e2, fe = _build_edges(fv)
    sh = RWGSheet()
    sh.style = "annular_ring"
    sh.units = units
    sh.rho = rho
    sh.e1, sh.e2, sh.fv, sh.fe = e1, e2, fv, fe
    sh.s1 = np.array([float(P), 0.0])
    sh.s2 = np.array([0.0, float(P)])
    sh.beta1, sh.beta2 = s1s2_to_beta(sh.s1, sh.s2)
    sh.fr = np.zeros(sh.nface)
    sh.sheet_class = sheet_class
    sh.fufp = fufp
    sh.xieta_check = False
    return sh
