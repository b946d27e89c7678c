"""ORACLE — test infrastructure only (used by oracle.py, tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg, never by the product).

Independent restatement of the per-point transmission-line side of the path:

  * Floquet modal constants beta, gamma, Y, c, tvec of a layer (setup_modes!,
    src/Modes.jl:298-322), as whole-array numpy expressions;
  * the incident-field partial GSM, the TL Green's functions and the incident
    voltages (currents) of a Gblock holding a J (M) sheet (gsm_electric_gblock /
    gsm_magnetic_gblock, src/GSMs.jl:213-526).  Each mode's stack is solved as a
    nodal network: one unknown voltage per interface, the layers as two-ports
    with admittance parameters coth(gamma d)/Z and -csch(gamma d)/Z, the end
    regions as matched loads, sources as Norton currents.  The reference walks
    impedance recursions (eqs. 6.3-6.38) instead; both give the node voltages of
    the same circuit;
  * the dielectric junction GSM (gsm_slab_interface, src/GSMs.jl:545-620): the
    reflection from the same circuit, the transmission by eq. (3.24b);
  * the star product (cascade, src/GSMs.jl:125-141) from the interconnection
    equations of the two GSMs solved as one linear system; propagation through a
    layer (cascade!, src/GSMs.jl:155-176) as a star product with the layer's own
    diagonal GSM; translation (translate_gsm!, src/GSMs.jl:717-750) as diagonal
    phase matrices; pecgsm / pmcgsm;
  * the steering points and the Gblock walk of _analyze (src/PSSFSS.jl:230-305),
    including its cut-off skip (:252-255); duplicate Gblocks are recomputed,
    not copied.

Only mode and Gblock *selection* (choose_layer_modes!, choose_gblocks, pinned by
the reference's own goldens) is taken from the package.
"""
import math

import numpy as np

MU0 = 1.25663706212 * 1e-6  # src/Constants.jl:6
C0 = 299792458.0
ETA0 = math.sqrt(MU0 / (1.0 / (MU0 * C0 ** 2)))   # sqrt(mu0 / eps0), src/Constants.jl:12-15
TE, TM = 1, 2             # mode-type codes carried by the Layer objects' P lists


def branch_sqrt(z):
    """Principal square root, except that a purely negative-imaginary root is
    negated (src/PGF.jl:116-120): decaying or outgoing waves."""
    r = np.sqrt(np.asarray(z, dtype=complex))
    flip = (r.real == 0.0) & (r.imag < 0.0)
    return np.where(flip, -r, r)


def _bsq(z):
    return complex(branch_sqrt(complex(z)))


class ModalLayer:
    """A layer's electrical data and mode list with the oracle's own per-point
    modal constants."""

    def __init__(self, layer):
        self.epsr = complex(layer.epsr)
        self.mur = complex(layer.mur)
        self.width = float(layer.width)
        self.P = [int(p) for p in layer.P]
        self.M = np.array(layer.M, dtype=float)
        self.N = np.array(layer.N, dtype=float)
        self.beta1 = np.array(layer.beta1, dtype=float)
        self.beta2 = np.array(layer.beta2, dtype=float)
        self.key = {(p, int(m), int(n)): i for i, (p, m, n) in
                    enumerate(zip(self.P, layer.M, layer.N))}

    def setup(self, k0, b00):
        """beta, gamma, Y, c, tvec of every mode at free-space wavenumber k0 and
        incident transverse wavevector b00 (src/Modes.jl:298-322)."""
        nm = len(self.P)
        self.k0 = k0
        if nm == 0:
            self.beta = np.zeros((0, 2))
            self.gamma = self.Y = self.c = np.zeros(0, complex)
            self.tvec = np.zeros((0, 2))
            return
        cell = (2 * math.pi) ** 2 / abs(self.beta1[0] * self.beta2[1] - self.beta1[1] * self.beta2[0])
        b = (np.asarray(b00, float)[None, :] + self.M[:, None] * self.beta1[None, :]) \
            + self.N[:, None] * self.beta2[None, :]
        b2 = b[:, 0] * b[:, 0] + b[:, 1] * b[:, 1]
        small = b2 * cell < 1e-14
        nrm = np.sqrt(np.where(small, 1.0, b2))
        bh = np.where(small[:, None], np.array([[1.0, 0.0]]), b / nrm[:, None])
        g = branch_sqrt(b2 - k0 * k0 * self.epsr * self.mur)
        te = np.array([p == TE for p in self.P])
        Y = np.where(te, g / (1j * k0 * ETA0 * self.mur), 1j * k0 * self.epsr / (ETA0 * g))
        self.tvec = np.where(te[:, None], np.stack([-bh[:, 1], bh[:, 0]], axis=1), bh)
        self.beta, self.gamma, self.Y = b, g, Y
        self.c = branch_sqrt(1.0 / (cell * Y))

    def cell_area_recip(self):
        return abs(self.beta1[0] * self.beta2[1] - self.beta1[1] * self.beta2[0])


# ------------------------------------------------------------ nodal TL network
def _line_impedance(p, beta_sq, k0, epsr, mur):
    g = _bsq(beta_sq - k0 * k0 * epsr * mur)
    Z = (1j * k0 * ETA0 * mur) / g if p == TE else g / (1j * k0 * epsr / ETA0)
    return g, Z


def _two_port(g, Z, d):
    """Admittance parameters (self, mutual) of a line of length d, written with
    exp(-2 g d) so that lossy or evanescent sections stay finite."""
    e2 = np.exp(-2 * g * d)
    coth = (1 + e2) / (1 - e2)
    csch = 2 * np.exp(-g * d) / (1 - e2)
    return coth / Z, -csch / Z


def _network(layers, p, beta_sq, k0, src_end, src_Y):
    """Node admittance matrix of the region's interfaces (node i-1 is interface i)
    for one mode type and transverse wavenumber; the end region on side src_end
    (0 first, 1 last) gets its admittance from the layer data, the other from the
    line formula.  Returns (Ynodal, Y_first_end, Y_last_end)."""
    nl = len(layers)
    nn = nl - 1
    Yn = np.zeros((nn, nn), complex)
    for i in range(1, nl - 1):            # interior layer i between nodes i-1 and i
        L = layers[i]
        g, Z = _line_impedance(p, beta_sq, k0, L.epsr, L.mur)
        if not L.width > 0:
            raise ValueError("interior layers need a positive width")
        ys, ym = _two_port(g, Z, L.width)
        Yn[i - 1, i - 1] += ys
        Yn[i, i] += ys
        Yn[i - 1, i] += ym
        Yn[i, i - 1] += ym
    ends = []
    for side, L in ((0, layers[0]), (1, layers[-1])):
        if side == src_end:
            ends.append(src_Y)
        else:
            ends.append(1.0 / _line_impedance(p, beta_sq, k0, L.epsr, L.mur)[1])
    Yn[0, 0] += ends[0]
    Yn[nn - 1, nn - 1] += ends[1]
    return Yn, ends[0], ends[1]


def gblock_electric(layers, s, k0):
    """Partial GSM, TLGF (voltage at the end interface per unit current injected at
    interface s) and incident voltages at interface s (from a 2 V Thevenin source
    of the matched end) for a J sheet at interface s of `layers` (ModalLayer, set
    up at the point)."""
    L1, LN = layers[0], layers[-1]
    n1, n2 = len(L1.P), len(LN.P)
    nn = len(layers) - 1
    gsm = new_gsm(n1, n2)
    nmx = max(n1, n2)
    vinc = np.zeros((nmx, 2), complex)
    tlgf = np.zeros((nmx, 2), complex)
    root_area = math.sqrt(LN.cell_area_recip() / L1.cell_area_recip())
    for q in range(n1):
        bsq = float(L1.beta[q] @ L1.beta[q])
        Yn, Y1, _ = _network(layers, L1.P[q], bsq, k0, 0, L1.Y[q])
        rhs = np.zeros(nn, complex)
        rhs[0] = 2.0 * Y1
        V = np.linalg.solve(Yn, rhs)
        gsm[0][q, q] = V[0] - 1
        vinc[q, 0] = V[s - 1]
        q2 = LN.key.get((L1.P[q], int(L1.M[q]), int(L1.N[q])))
        if q2 is not None:
            gsm[2][q2, q] = V[nn - 1] * L1.c[q] / LN.c[q2] * root_area
        e = np.zeros(nn, complex)
        e[s - 1] = 1.0
        tlgf[q, 0] = np.linalg.solve(Yn, e)[0]
    for q in range(n2):
        bsq = float(LN.beta[q] @ LN.beta[q])
        Yn, _, YN = _network(layers, LN.P[q], bsq, k0, 1, LN.Y[q])
        rhs = np.zeros(nn, complex)
        rhs[nn - 1] = 2.0 * YN
        V = np.linalg.solve(Yn, rhs)
        gsm[3][q, q] = V[nn - 1] - 1
        vinc[q, 1] = V[s - 1]
        q1 = L1.key.get((LN.P[q], int(LN.M[q]), int(LN.N[q])))
        if q1 is not None:
            gsm[1][q1, q] = V[0] * LN.c[q] / L1.c[q1] / root_area
        e = np.zeros(nn, complex)
        e[s - 1] = 1.0
        tlgf[q, 1] = np.linalg.solve(Yn, e)[nn - 1]
    return gsm, tlgf, vinc


def _shorted_side(layers, p, bsq, k0, src_Y, s, left):
    """The part of the region on one side of interface s, with interface s short
    circuited.  Returns the node admittance matrix over that side's free nodes
    (ordered from the matched end towards s), the end admittance, and the mutual
    admittance coupling the node next to s to s (0 when there is none)."""
    nl = len(layers)
    idx = list(range(1, s)) if left else list(range(nl - 2, s - 1, -1))   # interior layers
    # free nodes: left side nodes 1..s-1 (interfaces), right side s+1..nl-1
    nodes = list(range(1, s)) if left else list(range(nl - 1, s, -1))
    pos = {nd: i for i, nd in enumerate(nodes)}
    Yn = np.zeros((len(nodes), len(nodes)), complex)
    ym_s = 0.0
    for i in idx:
        L = layers[i]                       # layer i (0-based) lies between interfaces i and i+1
        g, Z = _line_impedance(p, bsq, k0, L.epsr, L.mur)
        ys, ym = _two_port(g, Z, L.width)
        a, b = i, i + 1
        for x, y in ((a, b), (b, a)):
            if x in pos:
                Yn[pos[x], pos[x]] += ys
                if y in pos:
                    Yn[pos[x], pos[y]] += ym
                elif y == s:
                    ym_s = ym
    if nodes:
        Yn[0, 0] += src_Y
    return Yn, ym_s, nodes


def gblock_magnetic(layers, s, k0):
    """Partial GSM (interface s short circuited), TLGF (current at the end interface
    per unit voltage source at interface s) and incident currents at interface s
    (from a 2 A Norton source of the matched end) for an M sheet."""
    L1, LN = layers[0], layers[-1]
    n1, n2 = len(L1.P), len(LN.P)
    gsm = new_gsm(n1, n2)
    gsm[1][:] = 0                             # the short decouples the two sides
    gsm[2][:] = 0
    nmx = max(n1, n2)
    iinc = np.zeros((nmx, 2), complex)
    tlgf = np.zeros((nmx, 2), complex)
    for q in range(n1):
        bsq = float(L1.beta[q] @ L1.beta[q])
        Y1 = L1.Y[q]
        Yn, ym_s, nodes = _shorted_side(layers, L1.P[q], bsq, k0, Y1, s, True)
        if nodes:
            rhs = np.zeros(len(nodes), complex)
            rhs[0] = 2.0
            V = np.linalg.solve(Yn, rhs)
            I_in = 2.0 - Y1 * V[0]            # rightward current entering the network
            i_s = -ym_s * V[-1]               # rightward current into the short at s
            rhs = np.zeros(len(nodes), complex)
            rhs[-1] = -ym_s * (-1.0)          # interface s held at -1 V
            Vt = np.linalg.solve(Yn, rhs)
            tl = -Y1 * Vt[0]
        else:                                 # sheet on the first interface
            I_in, i_s, tl = 2.0, 2.0, Y1
        gsm[0][q, q] = 1 - I_in
        iinc[q, 0] = i_s
        tlgf[q, 0] = tl
    for q in range(n2):
        bsq = float(LN.beta[q] @ LN.beta[q])
        YN = LN.Y[q]
        Yn, ym_s, nodes = _shorted_side(layers, LN.P[q], bsq, k0, YN, s, False)
        if nodes:
            rhs = np.zeros(len(nodes), complex)
            rhs[0] = 2.0
            V = np.linalg.solve(Yn, rhs)
            I_left = 2.0 - YN * V[0]          # leftward current entering the network
            i_s = ym_s * V[-1]                # rightward current at s (from the short
                                              # into the first layer beyond it)
            rhs = np.zeros(len(nodes), complex)
            rhs[-1] = -ym_s * 1.0             # interface s held at +1 V
            Vt = np.linalg.solve(Yn, rhs)
            tl = YN * Vt[0]
        else:                                 # sheet on the last interface
            I_left, i_s, tl = 2.0, -2.0, YN
        gsm[3][q, q] = 1 - I_left
        iinc[q, 1] = i_s
        tlgf[q, 1] = tl
    return gsm, tlgf, iinc


def slab_interface(L1, L2, k0):
    """Junction of two semi-infinite media (ModalLayer, set up at the point)."""
    n1, n2 = len(L1.P), len(L2.P)
    gsm = [np.zeros((n1, n1), complex), np.zeros((n1, n2), complex),
           np.zeros((n2, n1), complex), np.zeros((n2, n2), complex)]

    def other_Y(Lsrc, i, Loth):
        j = Loth.key.get((Lsrc.P[i], int(Lsrc.M[i]), int(Lsrc.N[i])))
        if j is not None:
            return Loth.Y[j], j
        ksq = complex(Loth.beta[0] @ Loth.beta[0]) - Loth.gamma[0] ** 2
        g = _bsq(float(Lsrc.beta[i] @ Lsrc.beta[i]) - ksq)
        Y = g / (1j * k0 * ETA0 * Loth.mur) if Lsrc.P[i] == TE else 1j * k0 / ETA0 * Loth.epsr / g
        return Y, None

    for i in range(n1):
        Ya = L1.Y[i]
        Yb, j = other_Y(L1, i, L2)
        V = 2 * Ya / (Ya + Yb)                  # 2 V source behind 1/Ya into 1/Yb
        gsm[0][i, i] = V - 1
        if j is not None:
            gsm[2][j, i] = 2 * _bsq(Yb) * _bsq(Ya) / (Yb + Ya)
    for i in range(n2):
        Yb = L2.Y[i]
        Ya, j = other_Y(L2, i, L1)
        V = 2 * Yb / (Ya + Yb)
        gsm[3][i, i] = V - 1
        if j is not None:
            gsm[1][j, i] = 2 * _bsq(Yb) * _bsq(Ya) / (Yb + Ya)
    return gsm


# --------------------------------------------------------------- GSM algebra
def new_gsm(n1, n2):
    k = min(n1, n2)
    s12 = np.zeros((n1, n2), complex)
    s21 = np.zeros((n2, n1), complex)
    s12[range(k), range(k)] = 1
    s21[range(k), range(k)] = 1
    return [np.zeros((n1, n1), complex), s12, s21, np.zeros((n2, n2), complex)]


def star(a, b):
    """Connect port 2 of `a` to port 1 of `b`: solve for the two wave vectors on
    the junction, u = a21 x1 + a22 v (leaving a) and v = b11 u + b12 x3 (leaving b),
    for unit excitations x1 and x3 at once."""
    a11, a12, a21, a22 = a
    b11, b12, b21, b22 = b
    n1, n2, n3 = a11.shape[0], a22.shape[0], b22.shape[0]
    K = np.block([[np.eye(n2), -a22], [-b11, np.eye(n2)]])
    R = np.block([[a21, np.zeros((n2, n3))], [np.zeros((n2, n1)), b12]])
    W = np.linalg.solve(K, R)
    u, v = W[:n2], W[n2:]
    s11 = a11 + a12 @ v[:, :n1]
    s13 = a12 @ v[:, n1:]
    s31 = b21 @ u[:, :n1]
    s33 = b22 + b21 @ u[:, n1:]
    return [s11, s13, s31, s33]


def propagate(a, layer, width=None):
    """Side-2 ports of `a` carried through `layer` (its own diagonal GSM)."""
    d = layer.width if width is None else width
    if d == 0:
        return a
    t = np.diag(np.exp(-layer.gamma * d))
    n = t.shape[0]
    z = np.zeros((n, n), complex)
    return star(a, [z, t, t, z])


def translate(g, dx, dy, L1, L2):
    d = np.array([dx, dy])
    D1 = np.exp(1j * (L1.beta @ d))
    D2 = np.exp(1j * (L2.beta @ d))
    return [(g[0] * D1[None, :]) / D1[:, None], (g[1] * D2[None, :]) / D1[:, None],
            (g[2] * D1[None, :]) / D2[:, None], (g[3] * D2[None, :]) / D2[:, None]]


def pec(n1, n2):
    return [-np.eye(n1, dtype=complex), np.zeros((n1, n2), complex),
            np.zeros((n2, n1), complex), -np.eye(n2, dtype=complex)]


def pmc(n1, n2):
    return [np.eye(n1, dtype=complex), np.zeros((n1, n2), complex),
            np.zeros((n2, n1), complex), np.eye(n2, dtype=complex)]


# ---------------------------------------------------------------- the sweep
_PHI = ("phi", "ϕ", "Phi", "PHI", "Φ")
_PSI1 = ("psi1", "ψ₁", "Psi1", "PSI1", "ψ1")
_PSI2 = ("psi2", "ψ₂", "Psi2", "PSI2", "ψ2")


def points(setup):
    """(FGHz, k0, beta00, steering) in the reference's loop order: outer steering
    key, inner steering key, frequency (src/PSSFSS.jl:230-250)."""
    k1, _ = setup.stkeys
    out = []
    L0 = setup.layers[0]
    for so in setup.stvalues[0]:
        for si in setup.stvalues[1]:
            if k1 in _PSI1 + _PSI2:
                p1, p2 = (so, si) if k1 in _PSI1 else (si, so)
                sh = setup.sheets[0]
                b00 = (math.radians(p1) * sh.beta1 * sh.upm
                       + math.radians(p2) * sh.beta2 * sh.upm) / (2 * math.pi)
                steer = {"psi1": p1, "psi2": p2}
            else:
                th, ph = (si, so) if k1 in _PHI else (so, si)
                steer = {"theta": th, "phi": ph}
            for f in setup.freqs:
                k0 = 2 * math.pi * f * 1e9 / C0
                if k1 not in _PSI1 + _PSI2:
                    kr = k0 * math.sqrt((L0.epsr * L0.mur).real) * math.sin(math.radians(th))
                    b00 = np.array([kr * math.cos(math.radians(ph)),
                                    kr * math.sin(math.radians(ph))])
                out.append((f, k0, np.array(b00, float), steer))
    return out


def principal_propagates(first, last):
    """src/PSSFSS.jl:252-255: the principal mode of both end media propagates
    (its gamma purely imaginary with positive imaginary part)."""
    a, b = np.angle(first.gamma[0]), np.angle(last.gamma[0])
    tol = math.sqrt(np.finfo(float).eps)
    return (abs(a - b) <= tol * max(abs(a), abs(b))
            and abs(b - math.pi / 2) <= tol * max(abs(b), math.pi / 2))


def walk(setup, ml, k0, sheet_gsm):
    """Cascade of the stack at one point: each Gblock's GSM (sheet_gsm(region, sheet,
    u, rwgdat, s, isheet) for J/M sheets, fixed GSMs for E/H, the junction GSM for
    sheet-free Gblocks), translated if the sheet is offset, star-connected, then
    carried through the layer after it."""
    g = propagate(new_gsm(len(ml[0].P), len(ml[0].P)), ml[0])
    for gbl in setup.gbls:
        i1, i2 = gbl.rng[0], 1 + gbl.rng[-1]
        region = ml[i1 - 1:i2]
        if gbl.j != 0:
            ish = setup.junc[gbl.j - 1]
            sheet = setup.sheets[ish - 1]
            s = gbl.j - i1 + 1
            if sheet.sheet_class in ("J", "M"):
                gb = sheet_gsm(region, sheet, setup.uvec[ish - 1], setup.rwgdat[ish - 1], s,
                               ish - 1)
            elif sheet.sheet_class == "E":
                gb = pec(len(region[0].P), len(region[-1].P))
            elif sheet.sheet_class == "H":
                gb = pmc(len(region[0].P), len(region[-1].P))
            else:
                raise ValueError(sheet.sheet_class)
            if sheet.dx != 0 or sheet.dy != 0:
                gb = translate(gb, sheet.dx / sheet.upm, sheet.dy / sheet.upm, region[0],
                               region[-1])
        else:
            gb = slab_interface(ml[i1 - 1], ml[i2 - 1], k0)
        g = propagate(star(g, gb), ml[i2 - 1])
    return g
