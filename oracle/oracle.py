"""ORACLE — test infrastructure only (imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product).

FP64 CPU restatement of PSSFSS.jl's per-frequency MoM path:
  * liboracle.so (pssfss_oracle.c): jksums / filljk! / spectral tables + FFT /
    Sigma interpolation / fillz / filly / rwgbfft!  (src/PGF.jl, src/Zint.jl,
    src/FillZY.jl, src/RWG.jl:428-529);
  * LAPACK zgetrf/zgetrs through scipy for lu! / ldiv! (src/PSSFSS.jl:401,439);
  * RHS, solve and GSM extraction of calculate_jtype_gsm / calculate_mtype_gsm
    (src/PSSFSS.jl:362-587);
  * tl_oracle.py: modal constants, the Gblock transmission-line partial GSMs and
    Green's functions, junction GSMs, the star product, the steering points and
    the Gblock walk of _analyze (src/Modes.jl:298-322, src/GSMs.jl:16-750,
    src/PSSFSS.jl:230-305), restated independently of the package.
Only mode and Gblock selection (pinned by the reference's own goldens) and the
Result record come from the package.
"""
import ctypes as C
import math
import os
import sys

import numpy as np
import scipy.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            from build import build  # noqa
            build()
        L = C.CDLL(path)
        L.orc_jksums.argtypes = [_d, C.c_double, C.c_double, _d, _d, C.c_int, C.c_double,
                                 _d, _d, C.POINTER(C.c_int)]
        L.orc_filljk.argtypes = [C.c_int64, _d, _i, C.c_int64, _i, C.c_double, C.c_double,
                                 C.c_double, _d, _d, _d, _d, _d]
        L.orc_c3d3.argtypes = [C.c_double, C.c_double, _d, _d, _d, _d, _d, _d]
        L.orc_modal_tables.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_int, _d, _d, _d, C.c_int, _d, _d, _d, C.c_double,
                                       _d, _d, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_modal_tables.restype = C.c_int
        L.orc_sigma.argtypes = [_d, C.c_int, _d, _d, C.c_double, C.c_double, _d, _d]
        L.orc_fill.argtypes = [C.c_int, C.c_int64, _d, C.c_double, _i, _i, _d, C.c_int64, _i,
                               _i, _i, C.c_int64, C.c_void_p, C.c_void_p, _d, _d, _d,
                               C.c_double, C.c_double, _d, _d, _d, _d, _d, _d, C.c_int, _d, _d,
                               C.c_double, C.c_double, _d]
        L.orc_rwgbfft.argtypes = [C.c_int64, _d, C.c_double, _i, _i, C.c_int64, _i, _i, _i,
                                  _d, C.c_double, C.c_double, _d]
        L.orc_fft2.argtypes = [_d, C.c_int]
        L.orc_decisions.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_num_threads.restype = C.c_int
        _LIB = L
    return _LIB


def _c2(z):
    z = complex(z)
    return np.array([z.real, z.imag])


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ii(a):
    return np.ascontiguousarray(a, dtype=np.int64)


# ----------------------------------------------------------------- primitives
def jksums(urho00, psi1, psi2, us1, us2, extract=True, convtest=1e-8):
    js, ks = np.zeros(2), np.zeros(2)
    r = C.c_int(0)
    lib().orc_jksums(_f(urho00), psi1, psi2, _f(us1), _f(us2), int(extract), convtest,
                     js, ks, C.byref(r))
    return complex(*js), complex(*ks), r.value


def c3d3(k0, u, mu1, eps1, mu2, eps2):
    c3, d3 = np.zeros(2), np.zeros(2)
    lib().orc_c3d3(k0, u, _c2(mu1), _c2(eps1), _c2(mu2), _c2(eps2), c3, d3)
    return complex(*c3), complex(*d3)


def filljk(sheet, rwgdat, u_local, psi1, psi2):
    """Frequency-independent J/K integrals and analytic singular terms per UFP class."""
    nufp = rwgdat.nufp
    rep = _ii(rwgdat.ufp_rep())
    JK = np.zeros(12 * nufp)
    rho_r = np.zeros(2 * nufp)
    rinv = np.zeros(nufp)
    lib().orc_filljk(sheet.nface, _f(sheet.rho), _ii(sheet.fv.T), nufp, rep, u_local, psi1,
                     psi2, _f(sheet.s1), _f(sheet.s2), JK, rho_r, rinv)
    return JK.view(np.complex128).reshape(nufp, 6), rho_r.reshape(nufp, 2), rinv


def modal_tables(magnetic, k0, u, psi1, psi2, layers, s, beta1, beta2, beta00, convtest,
                 mmax_cap=None):
    nl = len(layers)
    eps = _f(np.array([[l.epsr.real, l.epsr.imag] for l in layers]).ravel())
    mu = _f(np.array([[l.mur.real, l.mur.imag] for l in layers]).ravel())
    width = _f([l.width for l in layers])
    t1 = np.zeros(2 * 2050 * 2050)
    t2 = np.zeros(2 * 2050 * 2050)
    conv = C.c_int(0)
    cap = C.c_int(mmax_cap if mmax_cap else 2048)
    M = lib().orc_modal_tables(int(magnetic), k0, u, psi1, psi2, nl, eps, mu, width, s,
                               _f(beta1), _f(beta2), _f(beta00), convtest, t1, t2,
                               C.byref(conv), C.byref(cap))
    E = (M + 2) ** 2
    return (M, t1[:2 * E].copy(), t2[:2 * E].copy(), bool(conv.value))


def sigma(table, M, beta1, beta2, psi1, psi2, rho):
    out = np.zeros(2)
    lib().orc_sigma(_f(table), M, _f(beta1), _f(beta2), psi1, psi2, _f(rho), out)
    return complex(*out)


def fill(sheet, rwgdat, JK, rho_r, rinv, k0, u, layers, s, tables, beta1m, beta2m, psi1, psi2):
    """Z (class 'J') or Y (class 'M') via fillz/filly; tables = (M, t1, t2)."""
    magnetic = sheet.sheet_class == "M"
    M, t1, t2 = tables
    nbf = rwgdat.nbf
    Z = np.zeros(2 * nbf * nbf)
    if rwgdat.ufp_identity:
        offs = membs = None
        o_p = m_p = None
    else:
        offs = _ii(rwgdat.ufp_offsets)
        membs = _ii(rwgdat.ufp_members)
        o_p = offs.ctypes.data_as(C.c_void_p)
        m_p = membs.ctypes.data_as(C.c_void_p)
    la, lb = layers[s - 1], layers[s]
    JKf = _f(np.asarray(JK, np.complex128).view(np.float64).ravel())
    lib().orc_fill(int(magnetic), sheet.nface, _f(sheet.rho), sheet.upm, _ii(sheet.fv.T),
                   _ii(sheet.fe.T), _f(sheet.fr), nbf, _ii(rwgdat.bff.T), _ii(rwgdat.ebf),
                   _ii(rwgdat.eci), rwgdat.nufp, o_p, m_p, JKf, _f(rho_r).ravel(),
                   _f(rinv), k0, u, _c2(la.epsr), _c2(la.mur), _c2(lb.epsr), _c2(lb.mur),
                   _f(t1), _f(t2), M, _f(beta1m), _f(beta2m), psi1, psi2, Z)
    return Z.view(np.complex128).reshape(nbf, nbf, order="F")


def rwgbfft(sheet, rwgdat, k, psi1, psi2):
    nbf = rwgdat.nbf
    ft = np.zeros(4 * nbf)
    lib().orc_rwgbfft(sheet.nface, _f(sheet.rho), sheet.upm, _ii(sheet.fv.T), _ii(sheet.fe.T),
                      nbf, _ii(rwgdat.bff.T), _ii(rwgdat.ebf), _ii(rwgdat.eci), _f(k), psi1,
                      psi2, ft)
    return ft.view(np.complex128).reshape(nbf, 2)


class FrozenDecisions:
    """The oracle's discrete choices (rings summed, table cells and Floquet shifts,
    interpolation snaps, singular-integral branches), recorded per analysed point
    and sheet at one geometry and replayed at nearby geometries of the same mesh
    topology, so that central differences of the restatement see a smooth function:

        fz = FrozenDecisions(nufp, npoints)        # nufp: one int per sheet, or an int
        analyze(pk, base, freqs, steer, decisions=(fz, "record"))
        analyze(pk, moved, freqs, steer, decisions=(fz, "replay"))

    (the same frequencies and steering each time: the Gblock partition and the
    table sizes depend on them)."""

    MODES = {"record": 1, "replay": 2}

    def __init__(self, nufp, npoints=1):
        nufps = [int(nufp)] if np.ndim(nufp) == 0 else [int(v) for v in nufp]
        self.slots = [[(np.zeros(7 * n, np.int8), np.zeros(42 * n, np.int16),
                        np.zeros(n, np.uint8)) for n in nufps] for _ in range(npoints)]

    def select(self, mode, ipoint, isheet=0):
        r, c, g = self.slots[ipoint][isheet]
        lib().orc_decisions(self.MODES[mode], r.ctypes.data, c.ctypes.data, g.ctypes.data)

    @staticmethod
    def off():
        lib().orc_decisions(0, None, None, None)


# ---------------------------------------------------------- sheet GSM (one point)
class SheetCache:
    """The reference keeps the J/K integrals on the sheet keyed by (psi1, psi2, u)
    (src/FillZY.jl:97-108); the oracle keeps the same cache per sheet."""

    def __init__(self):
        self.key = None
        self.val = None


_TOL_U = 1e-4    # src/Constants.jl: tolerance of the J/K cache key on u


def sheet_point(region, sheet, u, rwgdat, s, k0, kinc, cache=None, want_z=False):
    """calculate_jtype_gsm / calculate_mtype_gsm (src/PSSFSS.jl:362-587) in FP64.

    region: tl_oracle.ModalLayer list, set up at the point.  Returns the Gblock's
    GSM as [s11, s12, s21, s22] and a dict of extras."""
    import tl_oracle as T
    upm = sheet.upm
    area = abs(sheet.s1[0] * sheet.s2[1] - sheet.s1[1] * sheet.s2[0]) / upm ** 2
    ends = (region[0], region[-1])
    acf = [math.sqrt(((2 * math.pi) * (2 * math.pi) / l.cell_area_recip()) / area) for l in ends]
    magnetic = sheet.sheet_class == "M"
    if magnetic:
        gsm, tlgf, vinc = T.gblock_magnetic(region, s, k0)
    else:
        gsm, tlgf, vinc = T.gblock_electric(region, s, k0)
    psi1 = float(kinc @ sheet.s1) / upm
    psi2 = float(kinc @ sheet.s2) / upm
    b1m, b2m = sheet.beta1 * upm, sheet.beta2 * upm
    b00 = (psi1 * b1m + psi2 * b2m) / (2 * math.pi)
    ulocal = u / upm
    cache = cache if cache is not None else SheetCache()
    if (cache.key is None or psi1 != cache.key[0] or psi2 != cache.key[1] or cache.key[2] == 0
            or abs((ulocal - cache.key[2]) / cache.key[2]) > _TOL_U):
        cache.val = filljk(sheet, rwgdat, ulocal, psi1, psi2)
        cache.key = (psi1, psi2, ulocal)
    JK, rho_r, rinv = cache.val
    tabs = modal_tables(magnetic, k0, u, psi1, psi2, region, s, b1m, b2m, b00, 1e-7)
    Z = fill(sheet, rwgdat, JK, rho_r, rinv, k0, u, region, s, tabs[:3], b1m, b2m, psi1, psi2)
    Zkeep = Z.copy() if want_z else None
    lu = sla.lu_factor(Z, overwrite_a=True, check_finite=False)
    nbf = rwgdat.nbf
    nmax = max(len(l.P) for l in ends)
    bfft = np.zeros((nbf, 2, nmax, 2), complex)
    for sr, l in enumerate(ends):
        for qp in range(len(l.P)):
            kv = l.beta[qp]
            if qp > 0 and _approx(kv, l.beta[qp - 1]):
                bfft[:, sr, qp] = bfft[:, sr, qp - 1]
                continue
            if sr == 1 and len(region[0].beta) > qp and _approx(kv, region[0].beta[qp]):
                bfft[:, 1, qp] = bfft[:, 0, qp]
                continue
            bfft[:, sr, qp] = rwgbfft(sheet, rwgdat, kv, psi1, psi2)
    blk = {(1, 1): 0, (1, 2): 1, (2, 1): 2, (2, 2): 3}
    sigma_ = -1
    for sr, ls in enumerate(ends):
        sigma_ *= -1
        for qp in range(len(ls.P)):
            if magnetic:
                t = ls.tvec[qp]
                zt = np.array([-t[1], t[0]])
                src = vinc[qp, sr] * ls.c[qp] * ls.Y[qp] * zt
            else:
                src = vinc[qp, sr] * ls.c[qp] * acf[sr] * ls.tvec[qp]
            b = bfft[:, sr, qp]
            rhs = np.conj(b[:, 0]) * src[0] + np.conj(b[:, 1]) * src[1]
            x = sla.lu_solve(lu, rhs, check_finite=False)
            for orr, lo in enumerate(ends):
                smat = gsm[blk[(orr + 1, sr + 1)]]
                for q in range(len(lo.P)):
                    FT = x @ bfft[:, orr, q]
                    t = lo.tvec[q]
                    if magnetic:
                        zt = np.array([-t[1], t[0]])
                        smat[q, qp] += (zt[0] * FT[0] + zt[1] * FT[1]) * (sigma_ * tlgf[q, orr] * lo.c[q])
                    else:
                        smat[q, qp] -= (t[0] * FT[0] + t[1] * FT[1]) * (
                            tlgf[q, orr] / (lo.c[q] * acf[orr] * area))
    return gsm, {"Z": Zkeep, "mmax": tabs[0], "converged": tabs[3], "psi": (psi1, psi2)}


def _approx(a, b):
    """Julia isapprox for vectors: norm(a-b) <= sqrt(eps)*max(norm(a), norm(b))."""
    return np.linalg.norm(a - b) <= 1.4901161193847656e-08 * max(np.linalg.norm(a), np.linalg.norm(b))


# ------------------------------------------------------------------- analyze
def analyze(pk, strata, freqs, steering, mode_override=None, want_z=False, dbmin=None,
            decisions=None):
    """The `_analyze` loop (src/PSSFSS.jl:199-334) with the oracle's sheet solver and
    transmission-line side (tl_oracle); mode and Gblock selection from the package's
    prepare().

    steering: dict with keys 'theta','phi' (degrees) or 'psi1','psi2' (degrees);
    the first key is the outer loop.  decisions: (FrozenDecisions, "record" or
    "replay"), one record per analysed point.  Returns a list of result dicts."""
    import tl_oracle as T
    kw = {} if dbmin is None else {"dbmin": dbmin}
    setup = pk.analyze.prepare(strata, freqs, steering, mode_override=mode_override, **kw)
    ml = [T.ModalLayer(l) for l in setup.layers]
    caches = [SheetCache() for _ in setup.sheets]
    results = []
    for fghz, k0, b00, steer in T.points(setup):
        for l in ml:
            l.setup(k0, b00)
        if not T.principal_propagates(ml[0], ml[-1]):
            continue
        extras = []
        ipoint = len(results)

        def sheet_solver(region, sheet, u, rwgdat, s, isheet):
            if decisions is not None:
                decisions[0].select(decisions[1], ipoint, isheet)
            g, ex = sheet_point(region, sheet, u, rwgdat, s, k0, b00, caches[isheet],
                                want_z=want_z)
            extras.append(ex)
            return g
        g = T.walk(setup, ml, k0, sheet_solver)
        res = pk.outputs.make_result(pk.GSM(*g), steer, b00, fghz, setup.layers)
        res.extras = extras
        results.append(res)
    if decisions is not None:
        FrozenDecisions.off()
    return results
